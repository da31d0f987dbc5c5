"""compute_metrics / download_metrics (cuda_kernel.cu:53-79, 158-189, 428-466)
against the oracle: [int f, int f^2, kinetic energy, entropy] of eval_f (with
the initial half-kick, rho.hpp:63-96) on the metrics grid.  These are sums of
positive terms (no cancellation): tolerance 1e-12 relative."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import _capi
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def hist_of(conf, n):
    h = O.new_history(conf, n + 2)
    O.run(conf, 0, n + 1, h)
    return h


@pytest.mark.parametrize("variant", [_capi.TRACE_FAST, _capi.TRACE_STRICT])
@pytest.mark.parametrize("metrics_grid", ["same", "finer"])
def test_metrics_1d(variant, metrics_grid):
    conf = nufi.landau_1d(Nx=32, Nu=128, T=1.0)
    cm = conf if metrics_grid == "same" else nufi.landau_1d(Nx=48, Nu=200, T=1.0)
    h = hist_of(conf, 12)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf, conf_metrics=cm) as s:
        s.set_trace_variant(variant)
        s.upload_history(h[: 13 * st], 0, 13)
        for n in [0, 1, 5, 12]:
            got = np.zeros(4)
            s.compute_metrics(n, 0, cm.Nx * cm.Nu)
            s.download_metrics(got)
            ref = O.metrics(conf, n, h, conf_metrics=cm)
            err = np.max(np.abs(got - ref) / np.abs(ref))
            print(f"[metrics 1d] grid={metrics_grid} variant={variant} n={n} max rel err {err:.3e}", flush=True)
            assert err < 1e-12, (n, got, ref)


def test_metrics_subranges_accumulate():
    conf = nufi.landau_1d(Nx=32, Nu=128, T=1.0)
    h = hist_of(conf, 6)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf) as s:
        s.upload_history(h[: 7 * st], 0, 7)
        full = np.zeros(4)
        s.compute_metrics(6)
        s.download_metrics(full)
        acc = np.zeros(4)
        N = conf.n_nodes()
        for a, b in [(0, 100), (100, 100), (100, 2500), (2500, N)]:
            s.compute_metrics(6, a, b)
            s.download_metrics(acc)
    assert np.max(np.abs(acc - full) / np.abs(full)) < 1e-13


def test_metrics_fp32c_context_is_fp64():
    conf = nufi.landau_1d(Nx=32, Nu=128, T=1.0)
    h = hist_of(conf, 4)
    st = conf.stride_t()
    out = []
    for prec in (nufi.FP64, nufi.FP32C):
        with nufi.GpuScheduler(conf, precision=prec) as s:
            s.upload_history(h[: 5 * st], 0, 5)
            m = np.zeros(4)
            s.compute_metrics(4)
            s.download_metrics(m)
            out.append(m)
    assert np.max(np.abs(out[0] - out[1]) / np.abs(out[0])) < 1e-14


@pytest.mark.parametrize("variant", [_capi.TRACE_FAST, _capi.TRACE_STRICT])
def test_metrics_3d(variant):
    conf = nufi.landau_3d(N=8, V=8, T=1.0)
    h = hist_of(conf, 4)
    st = conf.stride_t()
    with nufi.GpuScheduler(conf) as s:
        s.set_trace_variant(variant)
        s.upload_history(h[: 5 * st], 0, 5)
        for n in [0, 3, 4]:
            got = np.zeros(4)
            s.compute_metrics(n)
            s.download_metrics(got)
            ref = O.metrics(conf, n, h)
            err = np.max(np.abs(got - ref) / np.abs(ref))
            print(f"[metrics 3d] variant={variant} n={n} max rel err {err:.3e}", flush=True)
            assert err < 1e-12


def test_metrics_range_errors():
    conf = nufi.landau_1d(Nx=32, Nu=128, T=1.0)
    with nufi.GpuScheduler(conf) as s:
        with pytest.raises(nufi.NufiRangeError):
            s.compute_metrics(conf.Nt + 1)
        with pytest.raises(nufi.NufiRangeError):
            s.compute_metrics(0, 0, conf.n_nodes() + 1)
