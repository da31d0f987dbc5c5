"""The multi-workgroup field solve (field_large.hip) for grids whose two complex
buffers exceed one workgroup's LDS (more than 5120 points): whole device runs
against the oracle's (DHT Poisson, exact circulant fit) on a 20^3 grid and a
1D grid of 8192 points, the host-buffer Poisson and fit entry points, and the
1D adjoint at Nx = 4096 against the oracle's tangent.
Tolerances: FP64 <= 1e-12 (energies, slices), FP32C <= 1e-6, adjoint <= 1e-5."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import adjoint as A
from numericalflowiteration_amd.config import F0_LANDAU_1D, F0_LANDAU_3D, config_1d, config_3d
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def landau3d_20():
    return config_3d((20, 20, 20), (8, 8, 8), 1.0 / 8.0, 0.375, (0.0, 4 * np.pi), (-6.0, 6.0), F0_LANDAU_3D,
                     alpha=0.01, k=0.5)                     # 8000 points, steps 0..3


def landau1d_8192():
    """alpha = 0.2: rho = 1 - du sum f cancels to the perturbation, so the
    oracle's sequential FP64 v-sum carries ~1e-16/alpha relative error; at the
    two-stream's alpha = 1e-3 that floor alone is ~3e-12."""
    return config_1d(8192, 32, 1.0 / 16.0, 3.0 / 16.0, 0.0, 4 * np.pi, -8.0, 8.0, F0_LANDAU_1D, alpha=0.2, k=0.5)


@pytest.mark.parametrize("make,prec,tol", [
    (landau3d_20, nufi.FP64, 1e-12),
    (landau3d_20, nufi.FP32C, 1e-6),
    (landau1d_8192, nufi.FP64, 1e-12),
    (landau1d_8192, nufi.FP32C, 1e-6),
], ids=["3d-20^3-fp64", "3d-20^3-fp32c", "1d-8192-fp64", "1d-8192-fp32c"])
def test_large_grid_run_against_oracle(make, prec, tol):
    conf = make()
    print(f"[large field] dim={conf.dim} points={conf.n_space()} Nt={conf.Nt} precision={prec}", flush=True)
    hist = O.new_history(conf, conf.Nt + 1)
    e_ref = O.run(conf, 0, conf.Nt + 1, hist)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf, precision=prec) as s:
        e = s.run()
        h = s.history()
    err_e = float(np.max(np.abs(e - e_ref) / np.abs(e_ref)))
    err_h = rel(h, hist[: (conf.Nt + 1) * st])
    print(f"[large field] energy {err_e:.3e}, history {err_h:.3e}", flush=True)
    assert err_e <= tol and err_h <= tol


@pytest.mark.parametrize("make", [landau3d_20, lambda: nufi.twostream_1d(Nx=8192, Nu=32, T=1.0)],
                         ids=["3d-20^3", "1d-8192"])
def test_large_grid_poisson_and_fit(make):
    conf = make()
    rng = np.random.default_rng(21)
    rho = rng.standard_normal(conf.n_space())
    phi_ref, e_ref = O.poisson(conf, rho)
    sl_ref = O.fit(conf, phi_ref)
    with nufi.GpuScheduler(conf) as s:
        phi = rho.copy()
        e = s.poisson_solve(phi)
        sl = s.interpolate(phi_ref)
    print(f"[large field] Poisson {rel(phi, phi_ref):.3e}, energy {abs(e - e_ref) / e_ref:.3e}, "
          f"fit {rel(sl, sl_ref):.3e}", flush=True)
    assert rel(phi, phi_ref) <= 1e-12 and abs(e - e_ref) <= 1e-12 * e_ref and rel(sl, sl_ref) <= 1e-12


def test_adjoint_at_nx_4096_against_the_oracle_tangent():
    conf = nufi.multimode_1d(nufi.adjoint_modes(), Nx=4096, Nu=64, T=0.25)
    J, g = A.energy_and_grad(conf, nufi.adjoint_modes())
    Jt, dJ = O.tangent_run(conf)
    err = float(np.max(np.abs(g - dJ[:8])) / np.max(np.abs(dJ[:8])))
    print(f"[large field] adjoint Nx=4096: J {abs(J - Jt) / Jt:.3e}, grad {err:.3e}", flush=True)
    assert abs(J - Jt) <= 1e-12 * abs(Jt) and err <= 1e-5
