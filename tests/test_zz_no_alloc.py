"""nufi_rho3d_noalloc, the rho entry point that allocates nothing, called
through the C ABI on device buffers, against nufi_compute_rho (all nodes) +
nufi_download_rho of a context holding the same history: bit for bit, on two
calls in a row, the second on a smaller grid (so the shared scratch array
still holds the first call's partial sums beyond the second's)."""
import ctypes

import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import _capi
from numericalflowiteration_amd.config import F0_LANDAU_3D, config_3d

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 8 << 20


def grid(N, V, T):
    return config_3d((N, N, N), (V, V, V), 1.0 / 8.0, T, (0.0, 4 * np.pi), (-6.0, 6.0), F0_LANDAU_3D,
                     alpha=0.01, k=0.5)


def context_rho(conf, n_steps, n):
    """The history of n_steps device steps and rho(n) by the context entry points."""
    with nufi.GpuScheduler(conf) as s:
        s.run(0, n_steps)
        h = s.history()
        rho = np.zeros(conf.n_space())
        s.compute_rho(n)
        s.download_rho(rho)
    return h, rho


def noalloc(conf, n, hist_dev, rho_dev, stream):
    lib = _capi.load()
    return lib.nufi_rho3d_noalloc(ctypes.byref(conf), n, ctypes.c_void_p(hist_dev.data_ptr()), hist_dev.numel(),
                                  ctypes.c_void_p(rho_dev.data_ptr()), ctypes.c_void_p(stream))


def test_noalloc_rho_matches_compute_rho_twice():
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    cases = [(grid(16, 32, 2.0), 9, 8), (grid(8, 8, 1.0), 6, 5)]      # 16^3 x 32^3 at n=8, then 8^3 x 8^3 at n=5
    for conf, n_steps, n in cases:
        print(f"[noalloc] grid {conf.Nx}^3 x {conf.Nu}^3, step {n}", flush=True)
        h, ref = context_rho(conf, n_steps, n)
        hist_dev = torch.tensor(h, dtype=torch.float64, device="cuda")
        rho_dev = torch.full((conf.n_space(),), np.nan, dtype=torch.float64, device="cuda")
        rc = noalloc(conf, n, hist_dev, rho_dev, stream)
        assert rc == 0, _capi.load().nufi_last_error(None)
        torch.cuda.synchronize()
        got = rho_dev.cpu().numpy()
        print(f"[noalloc] max |diff| {np.max(np.abs(got - ref)):.3e}", flush=True)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_on_the_host():
    import torch
    lib = _capi.load()
    conf = grid(8, 8, 1.0)
    st = conf.stride_t()
    hist_dev = torch.zeros(9 * st + 8, dtype=torch.float64, device="cuda")
    rho_dev = torch.zeros(conf.n_space(), dtype=torch.float64, device="cuda")
    # n > Nt
    assert noalloc(conf, conf.Nt + 1, hist_dev, rho_dev, 0) == _capi.ERR_RANGE
    # a history not 16-byte aligned
    rc = lib.nufi_rho3d_noalloc(ctypes.byref(conf), 3, ctypes.c_void_p(hist_dev.data_ptr() + 8), 8 * st,
                                ctypes.c_void_p(rho_dev.data_ptr()), None)
    assert rc == _capi.ERR_INVALID
    # a history shorter than n slices + 1 real
    rc = lib.nufi_rho3d_noalloc(ctypes.byref(conf), 3, ctypes.c_void_p(hist_dev.data_ptr()), 3 * st,
                                ctypes.c_void_p(rho_dev.data_ptr()), None)
    assert rc == _capi.ERR_INVALID
    # partial sums beyond the scratch array: 32^3 points x 32^3 velocities
    big = grid(32, 32, 1.0)
    rc = lib.nufi_rho3d_noalloc(ctypes.byref(big), 0, None, 0, ctypes.c_void_p(rho_dev.data_ptr()), None)
    assert rc == _capi.ERR_UNSUPPORTED
    assert b"scratch" in lib.nufi_last_error(None)
    # 1D configs are not served
    assert lib.nufi_rho3d_noalloc(ctypes.byref(nufi.landau_1d()), 0, None, 0,
                                  ctypes.c_void_p(rho_dev.data_ptr()), None) == _capi.ERR_UNSUPPORTED
