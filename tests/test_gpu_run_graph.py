"""nufi_run replays a hipGraph of its step range (captured on the first call
for a range and launch shape): the energies and histories must equal those of
the same steps launched one by one, bit for bit, on every replay, after the
history was changed in between, across a change of launch shape, and in each
trace variant; timed runs launch directly (their launches are counted)."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import _capi

pytestmark = pytest.mark.gpu


def run_direct(conf, precision=nufi.FP64, variant=None):
    with nufi.GpuScheduler(conf, precision=precision) as s:
        s.set_run_graph(False)
        if variant is not None:
            s.set_trace_variant(variant)
        e = s.run()
        return e, s.history()


CASES = [
    ("landau 1D fast", lambda: nufi.landau_1d(Nx=32, Nu=128, T=2.0), nufi.FP64, None),
    ("landau 1D strict", lambda: nufi.landau_1d(Nx=32, Nu=128, T=2.0), nufi.FP64, _capi.TRACE_STRICT),
    ("two-stream 1D FP32C", lambda: nufi.twostream_1d(Nx=64, Nu=256, T=2.0), nufi.FP32C, None),
    ("landau 3D fast", lambda: nufi.landau_3d(N=8, V=8, T=1.0), nufi.FP64, None),
    ("landau 3D FP32C", lambda: nufi.landau_3d(N=8, V=8, T=1.0), nufi.FP32C, None),
]


@pytest.mark.parametrize("name,make,prec,variant", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_direct_launches(name, make, prec, variant):
    conf = make()
    e_ref, h_ref = run_direct(conf, prec, variant)
    with nufi.GpuScheduler(conf, precision=prec) as s:
        if variant is not None:
            s.set_trace_variant(variant)
        for rep in range(3):                     # capture, then two replays
            e = s.run()
            h = s.history()
            print(f"[run graph] {name} replay {rep}: max |de| {np.max(np.abs(e - e_ref)):.3e}", flush=True)
            assert np.array_equal(e, e_ref) and np.array_equal(h, h_ref)
            s.upload_history(np.zeros_like(h))   # the next replay recomputes every slice
        # a different range recaptures; the first half then the second equals the whole
        s.run(0, conf.Nt // 2)
        e2 = s.run(conf.Nt // 2, conf.Nt + 1)
        assert np.array_equal(e2, e_ref[conf.Nt // 2:])
        assert np.array_equal(s.history(), h_ref)


def test_graph_follows_a_launch_shape_change():
    conf = nufi.twostream_1d(Nx=64, Nu=256, T=2.0)
    e_ref, h_ref = run_direct(conf)
    with nufi.GpuScheduler(conf) as s:
        s.run()
        s.set_launch_shape(256, 4, 8)
        e = s.run()
        assert s.launch_shape()[0] == 256
        assert np.array_equal(e, e_ref) and np.array_equal(s.history(), h_ref)


def test_timed_runs_launch_directly():
    conf = nufi.landau_1d(Nx=32, Nu=128, T=1.0)
    with nufi.GpuScheduler(conf) as s:
        s.set_timing(True)
        s.run()
        launches, ms, cs = s.trace_timing()
        assert launches == conf.Nt + 1 and ms > 0
        assert cs == conf.char_steps(0, conf.Nt + 1)
