"""Checkpoint / resume and isolated steps through history files on the GPU
(nufi_save_history / nufi_load_history).  A run interrupted at step k, saved,
loaded into a fresh context and continued reproduces the uninterrupted run
bit for bit (binary and 17-digit text files); an isolated step recomputed
from a loaded history (bin/test_nufi_cpu_3d_isolated.cpp:169-228) reproduces
that step's slice and energy bit for bit."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi

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
