"""3D3V rho path on the GPU (through the C ABI) against the CPU oracle.

Tolerances: FP64 (strict, fast) <= 1e-12 relative per step on identical
history (normwise, mean-free rho / slice / energy); FP32C <= 1e-6.
"""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import _capi
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def mean_free(a):
    return a - a.mean()


_cache = {}


def small_3d():
    return nufi.landau_3d(N=8, V=8, T=1.0)     # 8^3 x 8^3 nodes, Nt = 8


def oracle_history(conf, n_slices):
    key = n_slices
    if key not in _cache:
        h = O.new_history(conf, n_slices + 1)
        e = O.run(conf, 0, n_slices, h)
        _cache[key] = (h, e)
    return _cache[key]


VARIANTS = [("strict", nufi.FP64, _capi.TRACE_STRICT, 1e-12),
            ("fast", nufi.FP64, _capi.TRACE_FAST, 1e-12),
            ("fp32c", nufi.FP32C, _capi.TRACE_FP32C, 1e-6)]


@pytest.mark.parametrize("name,prec,variant,tol", VARIANTS)
def test_rho_parity_3d(name, prec, variant, tol):
    conf = small_3d()
    hist, _ = oracle_history(conf, 6)
    st = conf.stride_t()
    print(f"[3d rho parity] {name}", flush=True)
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.set_trace_variant(variant)
        s.upload_history(hist[: 6 * st], 0, 6)
        for n in [0, 1, 3, 6]:
            rho = np.zeros(conf.n_space())
            s.compute_rho(n)
            s.download_rho(rho)
            ref = O.rho(conf, n, hist)
            exact = O.rho(conf, n, hist, exact_sum=True)
            floor = rel(mean_free(ref), mean_free(exact))
            err = rel(mean_free(rho), mean_free(exact))
            err_ref = rel(mean_free(rho), mean_free(ref))
            print(f"  n={n} rel err vs exact-sum {err:.3e}, vs reference order {err_ref:.3e} "
                  f"(reference floor {floor:.3e})", flush=True)
            assert err <= tol, (n, err)
            assert err_ref <= tol + 2 * floor, (n, err_ref, floor)


@pytest.mark.parametrize("name,prec,variant,tol", VARIANTS)
def test_step_parity_3d(name, prec, variant, tol):
    conf = small_3d()
    hist, energies = oracle_history(conf, 6)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.set_trace_variant(variant)
        s.upload_history(hist[: 5 * st], 0, 5)
        e = s.step(5)
        got = s.history(5, 6)
    ref = hist[5 * st: 6 * st]
    print(f"[3d step] {name}: slice {rel(got, ref):.3e} energy {abs(e - energies[5]) / energies[5]:.3e}",
          flush=True)
    assert rel(got, ref) <= tol
    assert abs(e - energies[5]) <= tol * energies[5]


def test_poisson_fit_3d_against_oracle():
    conf = nufi.landau_3d()           # 16^3 grid
    rng = np.random.default_rng(5)
    rho = rng.standard_normal(conf.n_space())
    with nufi.GpuScheduler(conf) as s:
        phi = rho.copy()
        e = s.poisson_solve(phi)
        sl = s.interpolate(phi)
    phi_ref, e_ref = O.poisson(conf, rho)
    sl_ref = O.fit(conf, phi_ref)
    assert rel(phi, phi_ref) < 1e-14
    assert abs(e - e_ref) < 1e-13 * e_ref
    assert rel(sl, sl_ref) < 1e-13


def test_q_subranges_3d():
    conf = small_3d()
    hist, _ = oracle_history(conf, 6)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf) as s:
        s.upload_history(hist[: 4 * st], 0, 4)
        full = np.zeros(conf.n_space())
        s.compute_rho(4)
        s.download_rho(full)
        acc = np.zeros(conf.n_space())
        N = conf.n_nodes()
        cuts = [0, 1, 700, 700, 70001, N]
        for a, b in zip(cuts[:-1], cuts[1:]):
            s.compute_rho(4, a, b)
            s.download_rho(acc)
    assert rel(acc, full) < 1e-13


@pytest.mark.parametrize("prec", [nufi.FP64, nufi.FP32C])
def test_rho_bit_identical_across_launch_shapes_3d(prec):
    """Each lane of the x-major FP64 kernel sums its fixed velocity group in iv
    order, and each wave of the velocity-major FP32C kernel its 64 velocities,
    so rho does not depend on the launch shape (for FP32C also not on the
    staging: two LDS buffers filled by direct loads, or, for the 16-slice
    chunk whose two buffers would not fit 160 KiB, one staged buffer)."""
    conf = small_3d()
    hist, _ = oracle_history(conf, 6)
    st = conf.stride_t()
    out = []
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.upload_history(hist[: 5 * st], 0, 5)
        # FP64: eight slices of this grid fit the LDS, and eight nodes a lane run as four off 16^3
        extra = [(512, 4, 16)] if prec == nufi.FP32C else [(512, 4, 8), (512, 8, 1)]
        for shape in [(256, 1, 1), (512, 4, 2), (256, 2, 1), (1024, 2, 1), (1024, 1, 1)] + extra:
            s.set_launch_shape(*shape)
            rho = np.zeros(conf.n_space())
            s.compute_rho(4)
            s.download_rho(rho)
            out.append(rho)
        with pytest.raises(nufi.NufiError):
            s.set_launch_shape(512, 3, 1)          # not built
        if prec == nufi.FP32C:
            with pytest.raises(nufi.NufiError):
                s.set_launch_shape(512, 8, 1)      # FP64 only
    for r in out[1:]:
        assert np.array_equal(r, out[0])


@pytest.mark.parametrize("N", [8, 6])
@pytest.mark.parametrize("name,prec,variant,tol", VARIANTS[1:])
def test_rho_parity_3d_long_history(N, name, prec, variant, tol):
    """Steps past slice 15, where the power-of-two kernels re-wrap positions
    (their cell indices wrap by a mask in between), and a 6^3 grid, whose
    kernels wrap exactly every step; history from the oracle's own run."""
    conf = nufi.config_3d((N, N, N), (6, 6, 6), 1 / 8, 3.0, (0.0, 4 * np.pi), (-6.0, 6.0),
                          nufi.config.F0_LANDAU_3D, 0.05, 0.5)
    hist = O.new_history(conf, 24 + 1)
    O.run(conf, 0, 24, hist)
    print(f"[3d long history] N={N} {name}", flush=True)
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.set_trace_variant(variant)
        s.upload_history(hist[: 24 * conf.stride_t()], 0, 24)
        for n in (16, 24):
            rho = np.zeros(conf.n_space())
            s.compute_rho(n)
            s.download_rho(rho)
            exact = O.rho(conf, n, hist, exact_sum=True)
            err = rel(mean_free(rho), mean_free(exact))
            print(f"  n={n} rel err {err:.3e}", flush=True)
            assert err <= tol, (n, err)


@pytest.mark.parametrize("cubic", [True, False])
@pytest.mark.parametrize("prec,tol", [(nufi.FP64, 1e-12), (nufi.FP32C, 1e-6)])
def test_rho_16cubed_compiled_strides_3d(prec, tol, cubic):
    """A 16^3 grid runs kernels whose slice strides are compile-time constants
    (two, four and, FP64, eight nodes a lane); other node counts run the
    runtime-stride kernels.  On cubic cells the FP32C kernel takes its kicks as
    the gradient itself; with y in [0, 8pi) (dy = 2 dx) it takes the general
    form.  rho at a step past the slice-15 re-wrap is the same bit for bit
    across all of them and matches the oracle's exact sum."""
    conf = nufi.config_3d((16, 16, 16), (8, 8, 8), 1 / 8, 3.0, (0.0, 4 * np.pi), (-6.0, 6.0),
                          nufi.config.F0_LANDAU_3D, 0.05, 0.5)
    if not cubic:
        conf.y_max = 8 * np.pi              # cos(0.5 y) stays periodic
        conf = conf.derive()
    n = 18
    hist = O.new_history(conf, n + 1)
    O.run(conf, 0, n, hist)
    exact = O.rho(conf, n, hist, exact_sum=True)
    shapes = [(512, 4, 1), (512, 2, 1), (1024, 2, 1), (256, 1, 1)]
    if prec == nufi.FP64:
        shapes = [(512, 8, 1), (512, 8, 2)] + shapes
    out = []
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.upload_history(hist[: n * conf.stride_t()], 0, n)
        for shape in shapes:
            s.set_launch_shape(*shape)
            rho = np.zeros(conf.n_space())
            s.compute_rho(n)
            s.download_rho(rho)
            out.append(rho)
            err = rel(mean_free(rho), mean_free(exact))
            print(f"[3d 16^3 strides] cubic={cubic} {shape} rel err {err:.3e}", flush=True)
            assert err <= tol, (shape, err)
    for shape, r in zip(shapes[1:], out[1:]):
        assert np.array_equal(r, out[0]), shape


@pytest.mark.parametrize("prec", [nufi.FP64, nufi.FP32C])
def test_autotune_3d_keeps_rho(prec):
    """The 3D autotuner times every built shape (FP64 also eight nodes a lane)
    on a 16^3 grid, keeps one the library accepts, and rho at the probe step
    is the same bit for bit before and after."""
    conf = nufi.config_3d((16, 16, 16), (8, 8, 8), 1 / 8, 1.0, (0.0, 4 * np.pi), (-6.0, 6.0),
                          nufi.config.F0_LANDAU_3D, 0.05, 0.5)
    n = 6
    hist = O.new_history(conf, n + 1)
    O.run(conf, 0, n, hist)
    with nufi.GpuScheduler(conf, precision=prec) as s:
        s.upload_history(hist[: n * conf.stride_t()], 0, n)
        before = np.zeros(conf.n_space())
        s.compute_rho(n)
        s.download_rho(before)
        shape, ms = s.autotune(n_probe=n, repeats=1)
        print(f"[3d autotune] {'fp64' if prec == nufi.FP64 else 'fp32c'} picked {shape} ({ms:.3f} ms)", flush=True)
        assert ms > 0
        assert s.launch_shape() == shape
        after = np.zeros(conf.n_space())
        s.compute_rho(n)
        s.download_rho(after)
    assert np.array_equal(after, before)
