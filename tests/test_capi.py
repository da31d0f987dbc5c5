"""The C-ABI library loads and exports every function include/nufi_hip.h declares;
host-side config derivation agrees bit for bit with the Python mirror.  No
compute call is made (no GPU needed)."""
import ctypes
import math
import os
import re
import subprocess

import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import _capi
from numericalflowiteration_amd import config as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nufi_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nufi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for must in ["nufi_create", "nufi_destroy", "nufi_compute_rho", "nufi_download_rho", "nufi_upload_phi",
                 "nufi_compute_metrics", "nufi_download_metrics", "nufi_step", "nufi_run",
                 "nufi_poisson_solve", "nufi_interpolate", "nufi_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.library_path()], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (nufi_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_binding_covers_exactly_the_header():
    assert set(_capi.SIGNATURES) == set(declared_functions())


def test_peak_probe_library_exports_its_probes():
    """bench.py --peak-probe loads libnufi_probes.so (a separate library: the
    microbenchmarks are not part of the rho-path ABI)."""
    path = os.path.join(ROOT, "numericalflowiteration_amd", "libnufi_probes.so")
    lib = ctypes.CDLL(path)
    for name in ("nufi_probe_fma", "nufi_probe_int", "nufi_probe_lds"):
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.library_path()], capture_output=True, text=True,
                         check=True).stdout
    assert "nufi_probe" not in out


def test_struct_layout_matches_header(tmp_path):
    """Compile the header with gcc and compare sizeof/offsetof with the ctypes mirror."""
    fields = [f for f, _ in C.Config._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(nufi_config, {f}));' for f in fields)
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "nufi_hip.h"
int main(void) {{
printf("sizeof %zu\\n", sizeof(nufi_config));
printf("f0size %zu\\n", sizeof(nufi_f0));
{body}
return 0; }}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(out["sizeof"]) == ctypes.sizeof(C.Config)
    assert int(out["f0size"]) == ctypes.sizeof(C.F0)
    for f in fields:
        assert int(out[f]) == getattr(C.Config, f).offset, f


@pytest.mark.parametrize("make", [lambda: nufi.landau_1d(), lambda: nufi.twostream_1d(),
                                  lambda: nufi.multimode_1d(nufi.adjoint_modes()), lambda: nufi.landau_3d()])
def test_derive_matches_c(make):
    py = make()
    c = py.copy()
    for f in ["dx", "dx_inv", "Lx", "Lx_inv", "du", "dy", "dy_inv", "Ly", "Ly_inv", "dz", "dz_inv", "Lz",
              "Lz_inv", "dv", "dw"]:
        setattr(c, f, 0.0)
    assert _capi.load().nufi_config_derive(ctypes.byref(c)) == 0
    for f in ["dx", "dx_inv", "Lx", "Lx_inv", "du"] + (
            ["dy", "dy_inv", "Ly", "Ly_inv", "dz", "dz_inv", "Lz", "Lz_inv", "dv", "dw"] if py.dim == 3 else []):
        assert getattr(c, f) == getattr(py, f), f


def test_derive_rejects_empty_grid():
    c = C.Config()
    assert _capi.load().nufi_config_derive(ctypes.byref(c)) == _capi.ERR_INVALID


def test_create_without_device_fails_cleanly():
    if _capi.device_count() > 0:
        pytest.skip("a device is present; covered by the gpu tests")
    with pytest.raises(nufi.NufiError):
        nufi.GpuScheduler(nufi.landau_1d())
    assert len(_capi.load().nufi_last_error(None)) > 0


def test_workload_presets():
    c1, c2, c5 = nufi.landau_1d(), nufi.twostream_1d(), nufi.landau_3d()
    assert (c1.Nx, c1.Nu, c1.Nt, c1.dt) == (64, 256, 480, 1 / 16)
    assert (c2.Nx, c2.Nu, c2.Nt, c2.dt) == (256, 1024, 1600, 1 / 32)
    assert (c5.Nx, c5.Nu, c5.Nt, c5.dt) == (16, 32, 160, 1 / 8)
    assert c1.x_max == 4 * math.pi and c2.x_max == 10 * math.pi
    # characteristic steps of a whole run: Nx^d Nv^d Nt(Nt+1)/2
    assert c1.char_steps(0, c1.Nt + 1) == 64 * 256 * 480 * 481 // 2
    assert abs(c2.char_steps(0, c2.Nt + 1) / 3.3575e11 - 1) < 1e-4
    assert abs(c5.char_steps(0, c5.Nt + 1) / 1.7287e12 - 1) < 1e-4
    assert c1.char_steps(10, 20) == 64 * 256 * sum(range(10, 20))


def test_sweep_is_seeded_and_in_range():
    a, b = nufi.sweep_runs(), nufi.sweep_runs()
    assert [(r.alpha, r.k, r.Nx, r.twostream) for r in a] == [(r.alpha, r.k, r.Nx, r.twostream) for r in b]
    assert len(a) == 512
    assert all(0.001 <= r.alpha <= 0.5 and 0.3 <= r.k <= 0.6 and r.Nx in (32, 64, 128, 256) for r in a)
    assert 200 < sum(r.twostream for r in a) < 312
    # SplitMix64 known answer for seed 0 (Vigna's reference implementation)
    rng = C.SplitMix64(0)
    assert rng.next_u64() == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("conf,prec", [
    (lambda: nufi.config_1d(30000, 64, 1 / 16, 1.0, 0.0, 1.0, -1.0, 1.0, C.F0_LANDAU_1D), C.FP64),
    (lambda: nufi.config_1d(16384, 64, 1 / 16, 1.0, 0.0, 1.0, -1.0, 1.0, C.F0_LANDAU_1D), C.FP32C),
    (lambda: nufi.config_3d((25, 25, 25), (4, 4, 4), 0.1, 1.0, (0.0, 1.0), (-1.0, 1.0), C.F0_LANDAU_3D), C.FP64),
    (lambda: nufi.config_3d((32, 32, 32), (4, 4, 4), 0.1, 1.0, (0.0, 1.0), (-1.0, 1.0), C.F0_LANDAU_3D), C.FP32C),
])
def test_create_refuses_grids_whose_slice_exceeds_the_lds(conf, prec):
    """One history slice in the trace kernel's LDS layout must fit 160 KiB;
    nufi_create refuses larger grids on the host, before touching a device."""
    lib = _capi.load()
    ctx = ctypes.c_void_p()
    c = conf()
    rc = lib.nufi_create(ctypes.byref(c), None, 4, prec, 0, ctypes.byref(ctx))
    assert rc == _capi.ERR_UNSUPPORTED, rc
    assert b"160 KiB" in lib.nufi_last_error(None)
    assert not ctx.value
