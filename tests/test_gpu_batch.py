"""Batched sweep (nufi_batch_*): many independent 1D runs of uneven grids and
parameters in shared launches.  Every run's history and energies must equal a
single-run context of the same config bit for bit (the same device arithmetic),
with and without the hipGraph and under any launch shape; two runs are also
checked against the oracle."""
import math

import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import config as C
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def small_sweep():
    runs = []
    specs = [(16, C.F0_LANDAU_1D, 0.05, 0.5), (32, C.F0_TWOSTREAM_1D, 0.2, 0.35), (64, C.F0_LANDAU_1D, 0.4, 0.31),
             (24, C.F0_TWOSTREAM_1D, 0.01, 0.55), (128, C.F0_LANDAU_1D, 0.001, 0.45), (32, C.F0_LANDAU_1D, 0.3, 0.6),
             (48, C.F0_TWOSTREAM_1D, 0.1, 0.42), (16, C.F0_TWOSTREAM_1D, 0.5, 0.3)]
    for Nx, kind, a, k in specs:
        runs.append(C.config_1d(Nx, 4 * Nx, 1 / 16, 2.0, 0.0, 2 * math.pi / k, -10.0, 10.0, kind, alpha=a, k=k))
    return runs


@pytest.mark.parametrize("prec", [nufi.FP64, nufi.FP32C])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_runs(prec, graph):
    runs = small_sweep()
    print(f"[batch] prec={prec} graph={graph} runs={len(runs)}", flush=True)
    with nufi.BatchScheduler(runs, precision=prec) as b:
        b.run(graph=graph)
        b.synchronize()
        bh = [b.history(r) for r in range(len(runs))]
        be = [b.energies(r) for r in range(len(runs))]
    for r, conf in enumerate(runs):
        with nufi.GpuScheduler(conf, precision=prec) as s:
            e = s.run()
            h = s.history()
        assert np.array_equal(bh[r], h), r
        assert np.array_equal(be[r], e), r


def test_batch_shapes_and_graph_replay_are_bit_stable():
    runs = small_sweep()
    with nufi.BatchScheduler(runs, precision=nufi.FP64) as b:
        b.run(graph=True)
        ref = [b.history(r) for r in range(len(runs))]
        shape, ms = b.autotune(n_probe=20, repeats=1)
        print(f"[batch] autotuned shape {shape} ({ms:.3f} ms)", flush=True)
        for sh in [(256, 4, 8), shape]:
            b.set_launch_shape(*sh)
            b.run(graph=True)         # recaptured for the new shape
            b.run(graph=True)         # replayed
            b.synchronize()
            for r in range(len(runs)):
                assert np.array_equal(b.history(r), ref[r]), (sh, r)


def test_batch_against_oracle():
    runs = small_sweep()
    with nufi.BatchScheduler(runs) as b:
        b.run()
        b.synchronize()
        for r in [0, 3]:
            conf = runs[r]
            hist = O.new_history(conf, conf.Nt + 1)
            e_ref = O.run(conf, 0, conf.Nt + 1, hist)
            e = b.energies(r)
            h = b.history(r)
            st = conf.stride_t()
            err_e = np.max(np.abs(e - e_ref) / e_ref)
            err_h = np.linalg.norm(h - hist[: (conf.Nt + 1) * st]) / np.linalg.norm(hist)
            print(f"[batch vs oracle] run {r}: energy {err_e:.3e} history {err_h:.3e}", flush=True)
            assert err_e <= 1e-12
            assert err_h <= 1e-12


def test_batch_rejects_bad_input():
    bad = C.config_3d((8, 8, 8), (8, 8, 8), 0.1, 1.0, (0.0, 1.0), (-1.0, 1.0), C.F0_LANDAU_3D)
    with pytest.raises(nufi.NufiError):
        nufi.BatchScheduler([bad])
