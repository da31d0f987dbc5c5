"""History files (numericalflowiteration_amd/csrc/history_io.cpp through the C ABI,
host only): lossless binary round trips with energies, exact text round trips
(17 significant digits), reading a file in the reference's own layout and
precision (bin/test_nufi_cpu_3d_isolated.cpp:128-135, 158-162: header lines
from std::to_string, values at setprecision(16)), partial files, and the
error paths."""
import math

import numpy as np
import pytest

import numericalflowiteration_amd as nufi


def info_1d(n_slices=5, Nt=9, Nx=16):
    return dict(dim=1, order=4, Nx=Nx, Ny=1, Nz=1, Nt=Nt, n_slices=n_slices, stride_t=Nx + 3, dt=1 / 16)


def info_3d(n_slices=3, Nt=4):
    return dict(dim=3, order=4, Nx=4, Ny=5, Nz=6, Nt=Nt, n_slices=n_slices, stride_t=7 * 8 * 9, dt=0.125)


def values(n, seed=1):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 3, n)
    v[::7] = 0.0
    v[1] = -0.0
    v[2] = 5e-324                                   # a subnormal survives both formats
    return v


@pytest.mark.parametrize("make", [info_1d, info_3d])
def test_binary_round_trip_is_bit_exact(tmp_path, make):
    info = make()
    sl = values(info["n_slices"] * info["stride_t"])
    en = values(info["n_slices"], seed=2)
    p = str(tmp_path / "h.bin")
    nufi.write_history(p, info, sl, en, fmt="binary")
    head, fmt = nufi.read_info(p)
    assert fmt == "binary" and head["n_energies"] == info["n_slices"]
    got_info, got, got_e = nufi.read_history(p)
    for k, v in info.items():
        assert got_info[k] == v, k
    assert got.tobytes() == sl.tobytes()
    assert got_e.tobytes() == en.tobytes()


@pytest.mark.parametrize("make", [info_1d, info_3d])
def test_text_round_trip_is_exact(tmp_path, make):
    info = make()
    sl = values(info["n_slices"] * info["stride_t"], seed=3)
    p = str(tmp_path / "h.txt")
    nufi.write_history(p, info, sl, fmt="text")
    head, fmt = nufi.read_info(p)
    assert fmt == "text"
    assert head["n_slices"] == info["n_slices"] and head["dim"] == info["dim"]
    assert head["n_energies"] == 0
    assert abs(head["dt"] - info["dt"]) < 5e-7               # std::to_string: six decimals
    _, got, _ = nufi.read_history(p)
    assert np.array_equal(got, sl)
    lines = open(p).read().split("\n")
    assert lines[:7] == [f"Nt = {info['Nt']}", f"dt = {info['dt']:.6f}", f"Nx = {info['Nx']}",
                         f"Ny = {info['Ny']}", f"Nz = {info['Nz']}", "order = 4", ""]


def test_reads_a_file_in_the_reference_layout(tmp_path):
    """The reference writes std::to_string header fields and the values with
    setprecision(16) (16 significant digits): read back to within one unit
    in the 16th digit."""
    Nt, Nx, Ny, Nz = 2, 3, 4, 5
    stride = (Nx + 3) * (Ny + 3) * (Nz + 3)
    sl = values((Nt + 1) * stride, seed=4)
    p = tmp_path / "bump_on_tail_potential_periodic_3d.txt"
    with open(p, "w") as f:
        f.write(f"Nt = {Nt}\ndt = {0.0625:.6f}\nNx = {Nx}\nNy = {Ny}\nNz = {Nz}\norder = 4\n\n")
        for v in sl:
            f.write(f"{v:.16g}\n")
    info, got, en = nufi.read_history(str(p))
    assert info["dim"] == 3 and info["n_slices"] == Nt + 1 and info["stride_t"] == stride
    assert en.size == 0
    rel = np.abs(got - sl) / np.maximum(np.abs(sl), 1e-300)
    assert np.all(rel <= 1e-15)
    assert np.all(got[sl == 0] == 0)


def test_partial_checkpoint_counts_whole_slices(tmp_path):
    info = info_1d(n_slices=4, Nt=9)
    sl = values(4 * info["stride_t"])
    p = str(tmp_path / "h.txt")
    nufi.write_history(p, info, sl, fmt="text")
    with open(p, "a") as f:                     # a torn write of a fifth slice
        f.write("1.0\n2.0\n")
    head, _ = nufi.read_info(p)
    assert head["n_slices"] == 4
    _, got, _ = nufi.read_history(p)
    assert np.array_equal(got, sl)


def test_error_paths(tmp_path):
    with pytest.raises(nufi.NufiIOError):
        nufi.read_info(str(tmp_path / "missing.bin"))
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTAHISTORYFILE" * 4)
    with pytest.raises(nufi.NufiIOError):
        nufi.read_history(str(bad))
    info = info_1d()
    sl = values(info["n_slices"] * info["stride_t"])
    p = tmp_path / "h.bin"
    nufi.write_history(str(p), info, sl, fmt="binary")
    data = p.read_bytes()
    (tmp_path / "short.bin").write_bytes(data[: len(data) - 16])       # truncated body
    with pytest.raises(nufi.NufiIOError):
        nufi.read_history(str(tmp_path / "short.bin"))
    wrong = dict(info, stride_t=info["stride_t"] + 1)                  # inconsistent header
    with pytest.raises(nufi.NufiError):
        nufi.write_history(str(tmp_path / "w.bin"), wrong, values(wrong["n_slices"] * wrong["stride_t"]))
    too_many = dict(info, n_slices=info["Nt"] + 2)
    with pytest.raises(nufi.NufiError):
        nufi.write_history(str(tmp_path / "w.bin"), too_many, values(too_many["n_slices"] * info["stride_t"]))
    assert math.isfinite(info["dt"])
