"""The FP32C kernels where their offset frames are tightest (the hi parts of
positions and velocities held at one per-axis offset so that the drift of the
hi parts is exact and the kick an ordered Fast2Sum, csrc/step1d.hpp and
csrc/trace3d.hip): many cells per
step, velocities reversed by a nonlinear field, grids that are not powers of
two (a wrap every step) and non-cubic 3D cells.  Per-step mean-free rho on
the oracle's own history against the oracle with the exact v-sum
(nufi/rho.hpp:31-61, 318-367 arithmetic); the FP32C bar, 1e-6 relative."""
import math

import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import config as C
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

FP32C_TOL = 1e-6


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def mean_free(a):
    return a - a.mean()


def check(conf, steps):
    hist = O.new_history(conf, conf.Nt + 1)
    O.run(conf, 0, conf.Nt, hist)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf, precision=nufi.FP32C) as s:
        s.upload_history(hist[: conf.Nt * st], 0, conf.Nt)
        for n in steps:
            rho = np.zeros(conf.n_space())
            s.compute_rho(n)
            s.download_rho(rho)
            exact = O.rho(conf, n, hist, exact_sum=True)
            err = rel(mean_free(rho), mean_free(exact))
            print(f"  n={n} rel err {err:.3e}", flush=True)
            assert err <= FP32C_TOL, (n, err)


def test_1d_many_cells_per_step_nonlinear():
    """Landau, alpha = 0.3, dt = 1/2 on dx = 4pi/64: up to 25 cells per step."""
    conf = C.config_1d(64, 128, 0.5, 10.0, 0.0, 4 * math.pi, -10.0, 10.0, C.F0_LANDAU_1D, 0.3, 0.5)
    print(f"[fp32c frames] 1D Landau Nx={conf.Nx} dt={conf.dt} Nt={conf.Nt}", flush=True)
    check(conf, [1, 10, 17, conf.Nt])


def test_1d_not_a_power_of_two():
    """Two-stream on Nx = 48 (a wrap every step), alpha = 0.2, several cells per step."""
    conf = C.config_1d(48, 96, 0.25, 5.0, 0.0, 4 * math.pi, -8.0, 8.0, C.F0_TWOSTREAM_1D, 0.2, 0.5)
    print(f"[fp32c frames] 1D two-stream Nx={conf.Nx} Nt={conf.Nt}", flush=True)
    check(conf, [1, 9, conf.Nt])


def test_3d_non_cubic_not_a_power_of_two():
    """12 x 8 x 10 cells (non-cubic, x and z not powers of two), dt = 1/4."""
    conf = C.config_3d((12, 8, 10), (6, 6, 6), 0.25, 2.0, (0.0, 4 * math.pi), (-6.0, 6.0),
                       C.F0_LANDAU_3D, 0.2, 0.5)
    print(f"[fp32c frames] 3D {conf.Nx}x{conf.Ny}x{conf.Nz} Nt={conf.Nt}", flush=True)
    check(conf, [1, 4, conf.Nt])


def test_3d_power_of_two_past_a_rewrap():
    """8^3 cells at dt = 1/2 (about two cells per step), steps past slice 15,
    where the pairs are re-wrapped and renormalised."""
    conf = C.config_3d((8, 8, 8), (6, 6, 6), 0.5, 10.0, (0.0, 4 * math.pi), (-6.0, 6.0),
                       C.F0_LANDAU_3D, 0.2, 0.5)
    print(f"[fp32c frames] 3D 8^3 dt={conf.dt} Nt={conf.Nt}", flush=True)
    check(conf, [16, conf.Nt])
