"""Checkpoint / resume and isolated steps through history files on the GPU
(nufi_save_history / nufi_load_history).  A run interrupted at step k, saved,
loaded into a fresh context and continued reproduces the uninterrupted run
bit for bit (binary and 17-digit text files); an isolated step recomputed
from a loaded history (bin/test_nufi_cpu_3d_isolated.cpp:169-228) reproduces
that step's slice and energy bit for bit."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def conf_1d():
    return nufi.landau_1d(Nx=32, Nu=128, T=2.0)            # Nt = 32


def conf_3d():
    return nufi.landau_3d(N=8, V=8, T=1.0)                 # Nt = 8


@pytest.mark.parametrize("fmt", ["binary", "text"])
@pytest.mark.parametrize("prec", [nufi.FP64, nufi.FP32C])
def test_resume_reproduces_the_uninterrupted_run_1d(tmp_path, fmt, prec):
    conf = conf_1d()
    k = 17
    with nufi.GpuScheduler(conf, precision=prec) as s:
        e_full = s.run()
        h_full = s.history()
    path = str(tmp_path / f"ckpt.{fmt}")
    with nufi.GpuScheduler(conf, precision=prec) as a:
        a.run(0, k)
        a.save_history(path, k, fmt=fmt)
    print(f"[history] 1D {fmt} precision={prec}: saved {k} slices", flush=True)
    with nufi.GpuScheduler(conf, precision=prec) as b:
        assert b.load_history(path) == k
        e_tail = b.run(k, conf.Nt + 1)
        h = b.history()
        e = b.energies()
    assert np.array_equal(h, h_full)
    assert np.array_equal(e_tail, e_full[k:])
    if fmt == "binary":                                    # energies of steps < k travel with the file
        assert np.array_equal(e, e_full)


def test_isolated_step_from_a_loaded_history_1d(tmp_path):
    conf = conf_1d()
    path = str(tmp_path / "full.bin")
    with nufi.GpuScheduler(conf) as s:
        e_full = s.run()
        h_full = s.history()
        s.save_history(path)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf) as b:
        assert b.load_history(path) == conf.Nt + 1
        for n in (0, 5, 20, conf.Nt):
            e = b.step(n)
            assert e == e_full[n], n
            assert np.array_equal(b.history(n, n + 1), h_full[n * st:(n + 1) * st]), n


def test_resume_and_isolated_step_3d(tmp_path):
    conf = conf_3d()
    with nufi.GpuScheduler(conf) as s:
        e_full = s.run()
        h_full = s.history()
    path = str(tmp_path / "ckpt3d.bin")
    with nufi.GpuScheduler(conf) as a:
        a.run(0, 4)
        a.save_history(path, 4)
    with nufi.GpuScheduler(conf) as b:
        assert b.load_history(path) == 4
        b.run(4, conf.Nt + 1)
        assert np.array_equal(b.history(), h_full)
        assert np.array_equal(b.energies(), e_full)
        st = conf.stride_t()
        assert b.step(6) == e_full[6]
        assert np.array_equal(b.history(6, 7), h_full[6 * st:7 * st])


def test_load_rejects_another_grid(tmp_path):
    path = str(tmp_path / "h.bin")
    with nufi.GpuScheduler(conf_1d()) as s:
        s.run(0, 3)
        s.save_history(path, 3)
    with nufi.GpuScheduler(nufi.landau_1d(Nx=64, Nu=128, T=2.0)) as b:
        with pytest.raises(nufi.NufiError):
            b.load_history(path)
    with nufi.GpuScheduler(nufi.landau_1d(Nx=32, Nu=128, T=1.0)) as b:      # shorter run: Nt 16 < 32
        with pytest.raises(nufi.NufiRangeError):
            b.load_history(path)
    with nufi.GpuScheduler(conf_1d()) as b:
        with pytest.raises(nufi.NufiIOError):
            b.load_history(str(tmp_path / "missing.bin"))


def write_reference_text(path, conf, hist, n_slices, order=4):
    """The potential file of bin/test_nufi_cpu_3d_isolated.cpp:128-135, 158-162:
    six header lines (std::to_string values), a blank line, then one value per
    line with std::setprecision(16) (C's %.16g)."""
    st = conf.stride_t(order)
    ny, nz = (conf.Ny, conf.Nz) if conf.dim == 3 else (1, 1)
    with open(path, "w") as f:
        f.write(f"Nt = {conf.Nt}\ndt = {conf.dt:.6f}\nNx = {conf.Nx}\nNy = {ny}\nNz = {nz}\norder = {order}\n\n")
        f.write("".join(f"{v:.16g}\n" for v in hist[: n_slices * st]))


@pytest.mark.parametrize("dim", [1, 3])
def test_isolated_steps_from_a_reference_text_file_against_the_oracle(tmp_path, dim):
    """The reference's isolated-step harness (bin/test_nufi_cpu_3d_isolated.cpp:
    169-228): an oracle run writes its history in the reference's 16-digit text
    layout; the file is loaded into a context and steps n are recomputed on the
    device from slices 0..n-1 of the file, against the oracle's step on the
    values the file holds (rho -> Poisson -> fit; energy and slice <= 1e-12).
    Steps run from the last down so each reads only slices the file gave."""
    conf = conf_1d() if dim == 1 else conf_3d()
    st = conf.stride_t()
    hist = O.new_history(conf, conf.Nt + 1)
    O.run(conf, 0, conf.Nt + 1, hist)
    path = str(tmp_path / "potential.txt")
    write_reference_text(path, conf, hist, conf.Nt + 1)
    vals = np.loadtxt(path, skiprows=6)
    assert vals.size == (conf.Nt + 1) * st
    file_hist = np.zeros(hist.size)
    file_hist[: vals.size] = vals
    with nufi.GpuScheduler(conf) as b:
        assert b.load_history(path) == conf.Nt + 1
        for n in (conf.Nt, 5, 1):
            e = b.step(n)
            got = b.history(n, n + 1)
            h = file_hist.copy()
            e_ref, _ = O.step(conf, n, h)
            err_e = abs(e - e_ref) / e_ref
            err_s = np.linalg.norm(got - h[n * st:(n + 1) * st]) / np.linalg.norm(h[n * st:(n + 1) * st])
            print(f"[history text] dim={dim} step {n}: energy rel err {err_e:.3e}, slice rel err {err_s:.3e}",
                  flush=True)
            assert err_e <= 1e-12 and err_s <= 1e-12, (n, err_e, err_s)
