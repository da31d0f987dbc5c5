"""The FP32C frames (csrc/internal.hpp fp32c_frame) against the properties
the kernels rely on.  Positions and velocities (cells, cells per step) are held
as P0 + s and W0 + w in FP32 hi + lo pairs.
- 3D (shared frame, W0 = P0 = 1.5 * 2^k): every hi part stays in the binade
  [2^k, 2^(k+1)), so the offset comes off a velocity exactly and the drift
  sh - (wh - W0) of the hi parts is exact (csrc/trace3d.hip drift_locate32).
- 1D (separate frames, W0 = 1.5 * 2^j, P0 = 1.5 * 2^k): the drift is an
  ordered Fast2Sum whose error term e = (sh - y) - d recovers sh - d exactly
  (csrc/step1d.hpp fp32c_move).
Checked in numpy float32 against the same operations in float64, for the
measured workloads' axes and for random grids; a frame whose binade were too
small, or a shared frame with W0 on a finer grid than P0, fails here."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi


def frame(N, cd, v_lo, nv, dv, shared):
    """fp32c_frame of the library."""
    vmax = max(abs(v_lo), abs(v_lo + nv * dv))
    wcap = 2.0 * cd * vmax + 1.0
    j = k = 1
    while 2.0 ** (j - 1) < wcap:
        j += 1
    while 2.0 ** (k - 1) < N + 17.0 * wcap + 2.0:
        k += 1
    P0 = np.float32(1.5 * 2.0 ** k)
    W0 = P0 if shared else np.float32(1.5 * 2.0 ** j)
    return P0, W0, wcap, j, k


def random_axes(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.integers(3, 5000))
        L = float(rng.uniform(0.5, 200.0))
        vmax = float(rng.uniform(0.5, 20.0))
        nv = int(rng.integers(4, 4096))
        dt = float(rng.uniform(1e-3, 0.5))
        out.append((N, dt * N / L, -vmax, nv, 2 * vmax / nv))
    return out


def axes_3d():
    c = nufi.landau_3d()
    return [(c.Nx, c.dt * c.dx_inv, c.u_min, c.Nu, c.du)] + random_axes(11, 30)


def axes_1d():
    out = []
    for conf in (nufi.twostream_1d(), nufi.landau_1d(), nufi.multimode_1d(nufi.adjoint_modes())):
        out.append((conf.Nx, conf.dt * conf.dx_inv, conf.u_min, conf.Nu, conf.du))
    return out + random_axes(12, 20)


def states(P0, W0, wcap, N, seed):
    """hi parts of positions before a drift (wrapped into [0, N), then at most
    16 drifts of at most wcap cells: the slices between re-wraps) and of
    velocities."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(-16 * wcap - 1, N + 16 * wcap + 1, 20000)
    w = rng.uniform(-wcap, wcap, 20000)
    return (np.float64(P0) + s).astype(np.float32), (np.float64(W0) + w).astype(np.float32)


@pytest.mark.parametrize("N,cd,v_lo,nv,dv", axes_3d())
def test_shared_frame_drift_is_exact(N, cd, v_lo, nv, dv):
    P0, W0, wcap, j, k = frame(N, cd, v_lo, nv, dv, shared=True)
    sh, wh = states(P0, W0, wcap, N, N)
    lo, hi = 2.0 ** k, 2.0 ** (k + 1)
    assert np.all((sh >= lo) & (sh < hi)) and np.all((wh >= lo) & (wh < hi))
    d = wh - W0                                                 # float32
    assert np.array_equal(d.astype(np.float64), wh.astype(np.float64) - np.float64(W0))
    y = sh - d                                                  # float32
    assert np.array_equal(y.astype(np.float64), sh.astype(np.float64) - d.astype(np.float64))
    assert np.all((y >= lo) & (y < hi))
    # the separate frames' finer velocity grid would not give an exact drift
    _, W0s, _, js, _ = frame(N, cd, v_lo, nv, dv, shared=False)
    if k - js >= 4:
        _, whs = states(P0, W0s, wcap, N, N + 1)
        ds = whs - W0s
        ys = sh - ds
        assert not np.array_equal(ys.astype(np.float64), sh.astype(np.float64) - ds.astype(np.float64))


@pytest.mark.parametrize("N,cd,v_lo,nv,dv", axes_1d())
def test_separate_frames_drift_is_a_fast2sum(N, cd, v_lo, nv, dv):
    P0, W0, wcap, j, k = frame(N, cd, v_lo, nv, dv, shared=False)
    sh, wh = states(P0, W0, wcap, N, N)
    assert np.all((wh >= 2.0 ** j) & (wh < 2.0 ** (j + 1)))
    d = wh - W0
    assert np.array_equal(d.astype(np.float64), wh.astype(np.float64) - np.float64(W0))
    y = sh - d
    e = (sh - y) - d
    assert np.array_equal(y.astype(np.float64) + e.astype(np.float64),
                          sh.astype(np.float64) - d.astype(np.float64))
    assert np.all((y >= 2.0 ** k) & (y < 2.0 ** (k + 1)))
