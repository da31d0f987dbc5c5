"""Parity at the full sizes of the measured workloads, on the GPU's own history.

* 1D two-stream, 256 x 1024 nodes, Nt = 1600 (x-major fast kernel with its
  16-slice re-wrap schedule and 13-slice double-buffered table chunks): the
  whole device-resident run, then isolated steps n = 1, 800, 1600 (rho against
  the oracle with the exact v-sum, and the slice and energy the oracle's step
  makes from the same history, bin/test_nufi_cpu_3d_isolated.cpp:169-228
  pattern).
* 3D Landau, 16^3 x 32^3 nodes: nine device steps, then rho of steps 1, 4 and
  8 against the oracle at a sample of 48 spatial points (the oracle's full 3D
  step at this size takes minutes); and the whole FP32C run against the whole
  FP64 run (the FP32C form's energy target, 1e-6 at every step).

Tolerances: FP64 <= 1e-12 relative (mean-free rho, slice, energy); FP32C
<= 1e-6.
"""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32C_TOL = 1e-6


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def mean_free(a):
    return a - a.mean()


@pytest.fixture(scope="module")
def twostream_run():
    conf = nufi.twostream_1d()
    print(f"[full size] two-stream run {conf.Nx} x {conf.Nu}, Nt = {conf.Nt}", flush=True)
    with nufi.GpuScheduler(conf) as s:
        e = s.run()
        h = s.history()
    hist = np.zeros(h.size + 8)
    hist[: h.size] = h
    ref = {n: O.rho(conf, n, hist, exact_sum=True) for n in (1, 800, 1600)}
    return conf, hist, e, ref


@pytest.mark.parametrize("prec,tol", [(nufi.FP64, FP64_TOL), (nufi.FP32C, FP32C_TOL)])
def test_twostream_fullsize_isolated_rho(twostream_run, prec, tol):
    conf, hist, _, ref = twostream_run
    st = conf.stride_t()
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.upload_history(hist[: (conf.Nt + 1) * st], 0, conf.Nt + 1)
        for n, exact in ref.items():
            rho = np.zeros(conf.Nx)
            s.compute_rho(n)
            s.download_rho(rho)
            err = rel(mean_free(rho), mean_free(exact))
            print(f"[full size] two-stream precision={prec} n={n}: rho rel err {err:.3e}", flush=True)
            assert err <= tol, (n, err)


def test_twostream_fullsize_isolated_step(twostream_run):
    """The slice and energy of steps 800 and 1600 of the device run against the
    oracle's step (rho, DHT Poisson, circulant fit) on the same slices 0..n-1."""
    conf, hist, e, _ = twostream_run
    st = conf.stride_t()
    for n in (800, 1600):
        h = hist.copy()
        e_ref, _ = O.step(conf, n, h)
        err_e = abs(e[n] - e_ref) / e_ref
        err_s = rel(hist[n * st:(n + 1) * st], h[n * st:(n + 1) * st])
        print(f"[full size] two-stream step {n}: energy rel err {err_e:.3e}, slice rel err {err_s:.3e}", flush=True)
        assert err_e <= FP64_TOL and err_s <= FP64_TOL, (n, err_e, err_s)


@pytest.mark.parametrize("prec,tol", [(nufi.FP64, FP64_TOL), (nufi.FP32C, FP32C_TOL)])
def test_landau3d_fullsize_isolated_rho(prec, tol):
    conf = nufi.landau_3d()
    st = conf.stride_t()
    print(f"[full size] 3D Landau 16^3 x 32^3, 9 steps, precision={prec}", flush=True)
    with nufi.GpuScheduler(conf) as s:
        s.run(0, 9)
        h = s.history(0, 9)
    hist = np.zeros(h.size + 8)
    hist[: h.size] = h
    rng = np.random.default_rng(3)
    pts = np.concatenate([[0, conf.n_space() - 1], rng.choice(np.arange(1, conf.n_space() - 1), 46, replace=False)])
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.upload_history(hist[: 9 * st], 0, 9)
        for n in (1, 4, 8):
            rho = np.zeros(conf.n_space())
            s.compute_rho(n)
            s.download_rho(rho)                          # -dv sum f, no "+1" (cuda_kernel.cu:425)
            ref = O.rho3d_points(conf, n, hist, pts, exact_sum=True, nthreads=16)
            # pointwise against the oracle's 1 - dv sum f; the scale is the perturbation
            err = np.linalg.norm((rho[pts] + 1.0) - ref) / np.linalg.norm(mean_free(rho)[pts])
            print(f"[full size] 3D Landau precision={prec} n={n}: rho rel err at {pts.size} points {err:.3e}",
                  flush=True)
            assert err <= tol, (n, err)


def test_landau3d_fp32c_run_meets_the_energy_target():
    """The accuracy target of the FP32-compensated form on the 3D3V config:
    over the whole 3D Landau run (16^3 x 32^3, dt = 1/8, T = 20, 161 steps) the
    electric energy of the FP32C run stays within 1e-6 relative of the FP64
    run's at every step (the maximum over n), and so does E at the nodes.  The
    FP64 run is the reference here; it matches the oracle to 1e-12 (the
    isolated-step tests above).  rho of the mean-free perturbation is reported
    and held to 5e-6: at the minima of the damped energy the perturbation is
    ~1e-5 of its n = 0 size, and there the FP32C rho error reaches ~1.7e-6
    relative to it.  The kernels are deterministic, so these figures do not
    depend on the box."""
    from numericalflowiteration_amd import diagnostics as D
    conf = nufi.landau_3d()
    runs = {}
    for prec, name in ((nufi.FP64, "fp64"), (nufi.FP32C, "fp32c")):
        print(f"[full size] 3D Landau run to T, {name}", flush=True)
        with nufi.GpuScheduler(conf, precision=prec) as s:
            e = s.run()
            runs[name] = (s.history(), e)
    errs = D.run_errors(conf, runs["fp32c"][0], runs["fp32c"][1], runs["fp64"][0], runs["fp64"][1])
    print(f"[full size] 3D Landau FP32C vs FP64 over {errs['steps_compared']} steps: energy "
          f"{errs['energy_max_rel_err']:.3e} (step {errs['energy_err_at_step']}), E {errs['E_max_rel_err']:.3e}, "
          f"rho {errs['rho_max_rel_err']:.3e}", flush=True)
    assert errs["steps_compared"] == conf.Nt + 1
    assert errs["energy_max_rel_err"] <= 1e-6
    assert errs["E_max_rel_err"] <= 1e-6
    assert errs["rho_max_rel_err"] <= 5e-6


def test_twostream_fp32c_run_errors_hold():
    """The whole two-stream FP32C run (256 x 1024, 1601 steps) against the FP64
    run: the per-config errors the FP32C form states (maximum over the steps of
    mean-free rho 1.1e-6, E 2.2e-7, energy 2.1e-7) held to about twice their
    values, so a change that costs the long unstable run accuracy (an offset
    frame shared by positions and velocities raised them to 4.0e-6, 6.9e-7 and
    9.9e-7) fails here.  The kernels are deterministic: the figures do not
    depend on the box."""
    from numericalflowiteration_amd import diagnostics as D
    conf = nufi.twostream_1d()
    runs = {}
    for prec, name in ((nufi.FP64, "fp64"), (nufi.FP32C, "fp32c")):
        with nufi.GpuScheduler(conf, precision=prec) as s:
            e = s.run()
            runs[name] = (s.history(), e)
    errs = D.run_errors(conf, runs["fp32c"][0], runs["fp32c"][1], runs["fp64"][0], runs["fp64"][1])
    print(f"[full size] two-stream FP32C vs FP64 over {errs['steps_compared']} steps: energy "
          f"{errs['energy_max_rel_err']:.3e}, E {errs['E_max_rel_err']:.3e}, rho {errs['rho_max_rel_err']:.3e}",
          flush=True)
    assert errs["steps_compared"] == conf.Nt + 1
    assert errs["energy_max_rel_err"] <= 5e-7
    assert errs["E_max_rel_err"] <= 5e-7
    assert errs["rho_max_rel_err"] <= 2.5e-6
