"""Python mirror of the reference's GPU rho scheduler.

`GpuScheduler` keeps the member functions and semantics of
``nufi::dim{1,3}::cuda_scheduler<real,order>`` (nufi/cuda_scheduler.hpp:33-164,
287-392): ``compute_rho(n, q_begin, q_end)`` queues the rho partials of a node
range, ``download_rho(rho)`` synchronises and *adds* into ``rho``,
``upload_phi(n, coeffs)`` copies slice ``n`` of a host history.  On top of that
it exposes the device-resident step loop (``step``, ``run``) that replaces the
host round trip of the reference drivers (bin/test_nufi_gpu_1d.cpp:241-253).

One context lives on one device; host arrays are numpy float64.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi
from .config import FP32C, FP64, Config

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


class GpuScheduler:
    """One device context of the rho path (cuda_kernel<real,order> + scheduler)."""

    def __init__(self, conf: Config, conf_metrics: Config | None = None, order: int = 4,
                 precision: int = FP64, device: int = 0):
        self._lib = _capi.load()
        self.conf = conf.copy()
        self.order = order
        self.precision = precision
        self.device = device
        ctx = ctypes.c_void_p()
        cm = ctypes.byref(conf_metrics) if conf_metrics is not None else None
        rc = self._lib.nufi_create(ctypes.byref(self.conf), cm, order, precision, device, ctypes.byref(ctx))
        _capi.check(rc, None)
        self._ctx = ctx

    # ---- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.nufi_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        _capi.check(rc, self._ctx)

    # ---- sizes -----------------------------------------------------------
    @property
    def stride_t(self) -> int:
        return self.conf.stride_t(self.order)

    @property
    def n_nodes(self) -> int:
        return self.conf.n_nodes()

    @property
    def n_space(self) -> int:
        return self.conf.n_space()

    # ---- the reference scheduler contract --------------------------------
    def compute_rho(self, n: int, q_begin: int | None = None, q_end: int | None = None):
        q_begin = 0 if q_begin is None else q_begin
        q_end = self.n_nodes if q_end is None else q_end
        self._check(self._lib.nufi_compute_rho(self._ctx, n, q_begin, q_end))

    def download_rho(self, rho: np.ndarray) -> np.ndarray:
        self._check(self._lib.nufi_download_rho(self._ctx, _ptr(rho)))
        return rho

    def upload_phi(self, n: int, coeffs: np.ndarray):
        if coeffs.size < (n + 1) * self.stride_t:
            raise ValueError("history too short for slice n")
        self._check(self._lib.nufi_upload_phi(self._ctx, n, _ptr(coeffs)))

    def compute_metrics(self, n: int, q_begin: int | None = None, q_end: int | None = None):
        q_begin = 0 if q_begin is None else q_begin
        q_end = self.n_nodes if q_end is None else q_end
        self._check(self._lib.nufi_compute_metrics(self._ctx, n, q_begin, q_end))

    def download_metrics(self, metrics: np.ndarray) -> np.ndarray:
        self._check(self._lib.nufi_download_metrics(self._ctx, _ptr(metrics)))
        return metrics

    # ---- device-resident loop --------------------------------------------
    def step(self, n: int, wait: bool = True) -> float | None:
        """Step n on the device; returns its energy (synchronising), or queues the
        step and returns None with wait=False."""
        if not wait:
            self._check(self._lib.nufi_step(self._ctx, n, None))
            return None
        e = np.zeros(1)
        self._check(self._lib.nufi_step(self._ctx, n, _ptr(e)))
        return float(e[0])

    def run(self, n_begin: int = 0, n_end: int | None = None) -> np.ndarray:
        n_end = self.conf.Nt + 1 if n_end is None else n_end
        e = np.zeros(max(0, n_end - n_begin))
        self._check(self._lib.nufi_run(self._ctx, n_begin, n_end, _ptr(e) if e.size else None))
        return e

    def set_run_graph(self, enabled: bool):
        """Whether run() replays a captured hipGraph of its step range (default on)."""
        self._check(self._lib.nufi_set_run_graph(self._ctx, int(enabled)))

    def history(self, n_begin: int = 0, n_end: int | None = None) -> np.ndarray:
        n_end = self.conf.Nt + 1 if n_end is None else n_end
        out = np.zeros((n_end - n_begin) * self.stride_t)
        if out.size:
            self._check(self._lib.nufi_download_history(self._ctx, n_begin, n_end, _ptr(out)))
        return out

    def upload_history(self, hist: np.ndarray, n_begin: int = 0, n_end: int | None = None):
        n_end = n_begin + hist.size // self.stride_t if n_end is None else n_end
        self._check(self._lib.nufi_upload_history(self._ctx, n_begin, n_end, _ptr(hist)))

    def energies(self, n_begin: int = 0, n_end: int | None = None) -> np.ndarray:
        n_end = self.conf.Nt + 1 if n_end is None else n_end
        out = np.zeros(n_end - n_begin)
        if out.size:
            self._check(self._lib.nufi_download_energies(self._ctx, n_begin, n_end, _ptr(out)))
        return out

    # ---- diagnostics samplers (bin/test_nufi_gpu_1d.cpp:290-361) -------------
    def sample_field(self, n: int, x) -> tuple[np.ndarray, np.ndarray]:
        """(E, rho) = (-phi', -phi'') of slice n at the points x (1D)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        E, rho = np.zeros_like(x), np.zeros_like(x)
        if x.size:
            self._check(self._lib.nufi_sample_field(self._ctx, n, _ptr(x), x.size, _ptr(E), _ptr(rho)))
        return E, rho

    def sample_f(self, n: int, x, v) -> np.ndarray:
        """f(n, x, v) with the initial half-kick (rho.hpp:63-96) at the points (x, v)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if x.shape != v.shape:
            raise ValueError("x and v differ in shape")
        f = np.zeros_like(x)
        if x.size:
            self._check(self._lib.nufi_sample_f(self._ctx, n, _ptr(x), _ptr(v), x.size, _ptr(f)))
        return f

    def phase_flow(self, n: int, x, v) -> tuple[np.ndarray, np.ndarray]:
        """The phase flow of rho.hpp:98-131 at the points (x, v)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if x.shape != v.shape:
            raise ValueError("x and v differ in shape")
        xo, vo = np.zeros_like(x), np.zeros_like(x)
        if x.size:
            self._check(self._lib.nufi_phase_flow(self._ctx, n, _ptr(x), _ptr(v), x.size, _ptr(xo), _ptr(vo)))
        return xo, vo

    # ---- history files (checkpoint / resume, isolated steps) ----------------
    def save_history(self, path: str, n_slices: int | None = None, fmt="binary"):
        """Slices 0..n_slices-1 of the device history (and, binary, their energies)."""
        from .history import _fmt
        n_slices = self.conf.Nt + 1 if n_slices is None else n_slices
        self._check(self._lib.nufi_save_history(self._ctx, os.fsencode(path), _fmt(fmt), n_slices))

    def load_history(self, path: str) -> int:
        """Load a history file of the same grid; returns the number of slices loaded."""
        n = ctypes.c_uint64(0)
        self._check(self._lib.nufi_load_history(self._ctx, os.fsencode(path), ctypes.byref(n)))
        return int(n.value)

    def synchronize(self):
        self._check(self._lib.nufi_synchronize(self._ctx))

    # ---- field solves on host buffers ------------------------------------
    def poisson_solve(self, data: np.ndarray) -> float:
        e = np.zeros(1)
        self._check(self._lib.nufi_poisson_solve(self._ctx, _ptr(data), _ptr(e)))
        return float(e[0])

    def interpolate(self, values: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(values, dtype=np.float64)
        out = np.zeros(self.stride_t)
        self._check(self._lib.nufi_interpolate(self._ctx, _ptr(out), _ptr(v)))
        return out

    # ---- adjoint ---------------------------------------------------------
    def adjoint_grad(self):
        """(J, dJ/da): J = sum_n dt*energy_n over steps 0..Nt and its gradient with
        respect to the f0 parameters (multimode a_m, else alpha).  Runs the forward
        loop and the reverse sweep on the device."""
        g = np.zeros(8)
        J = np.zeros(1)
        self._check(self._lib.nufi_adjoint_grad(self._ctx, _ptr(g), 8, _ptr(J)))
        npar = self.conf.f0.n_modes if self.conf.f0.kind == 3 else 1
        return float(J[0]), g[:npar].copy()

    # ---- kernel selection ------------------------------------------------
    def set_trace_variant(self, variant: int):
        self._check(self._lib.nufi_set_trace_variant(self._ctx, variant))

    def set_launch_shape(self, block: int, nodes_per_thread: int, chunk: int):
        self._check(self._lib.nufi_set_launch_shape(self._ctx, block, nodes_per_thread, chunk))

    def launch_shape(self):
        b, n, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.nufi_get_launch_shape(self._ctx, ctypes.byref(b), ctypes.byref(n), ctypes.byref(c)))
        return b.value, n.value, c.value

    def autotune(self, n_probe: int | None = None, repeats: int = 3):
        """Time the built launch shapes of the trace kernel at step n_probe and keep the fastest."""
        n_probe = self.conf.Nt if n_probe is None else n_probe
        b, n, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(1)
        self._check(self._lib.nufi_autotune(self._ctx, n_probe, repeats, ctypes.byref(b), ctypes.byref(n),
                                            ctypes.byref(c), _ptr(ms)))
        return (b.value, n.value, c.value), float(ms[0])

    # ---- timing ----------------------------------------------------------
    def set_timing(self, enabled: bool):
        self._check(self._lib.nufi_set_timing(self._ctx, int(enabled)))

    def trace_timing(self):
        """(launches, summed trace-kernel ms, summed char-steps) since set_timing."""
        n = ctypes.c_uint64(0)
        ms = np.zeros(1)
        cs = np.zeros(1)
        self._check(self._lib.nufi_trace_timing(self._ctx, ctypes.byref(n), _ptr(ms), _ptr(cs)))
        return int(n.value), float(ms[0]), float(cs[0])




class BatchScheduler:
    """Many independent 1D runs stepped together on one device (nufi_batch_*).

    Each time step is one trace launch over every run's nodes and one field
    launch (a workgroup per run); `run(graph=True)` replays a captured hipGraph
    of the whole step range.  Each run's history equals a single-run
    `GpuScheduler` of the same config bit for bit.
    """

    def __init__(self, confs, order: int = 4, precision: int = FP64, device: int = 0):
        self._lib = _capi.load()
        self.confs = [c.copy() for c in confs]
        arr = (Config * len(self.confs))(*self.confs)
        self._arr = arr
        ctx = ctypes.c_void_p()
        rc = self._lib.nufi_batch_create(arr, len(self.confs), order, precision, device, ctypes.byref(ctx))
        _capi.check(rc, None, batch=True)
        self._b = ctx
        self.order = order
        self.precision = precision

    def close(self):
        if getattr(self, "_b", None):
            self._lib.nufi_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        _capi.check(rc, self._b, batch=True)

    @property
    def n_runs(self) -> int:
        return len(self.confs)

    def run(self, n_begin: int = 0, n_end: int | None = None, graph: bool = True):
        n_end = max(c.Nt for c in self.confs) + 1 if n_end is None else n_end
        self._check(self._lib.nufi_batch_run(self._b, n_begin, n_end, int(graph)))

    def synchronize(self):
        self._check(self._lib.nufi_batch_synchronize(self._b))

    def energies(self, run: int) -> np.ndarray:
        c = self.confs[run]
        out = np.zeros(c.Nt + 1)
        self._check(self._lib.nufi_batch_energies(self._b, run, 0, c.Nt + 1, _ptr(out)))
        return out

    def history(self, run: int) -> np.ndarray:
        c = self.confs[run]
        out = np.zeros((c.Nt + 1) * c.stride_t(self.order))
        self._check(self._lib.nufi_batch_history(self._b, run, 0, c.Nt + 1, _ptr(out)))
        return out

    def set_launch_shape(self, block: int, nodes_per_thread: int, chunk: int):
        self._check(self._lib.nufi_batch_set_launch_shape(self._b, block, nodes_per_thread, chunk))

    def launch_shape(self):
        b, n, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.nufi_batch_get_launch_shape(self._b, ctypes.byref(b), ctypes.byref(n), ctypes.byref(c)))
        return b.value, n.value, c.value

    def autotune(self, n_probe: int, repeats: int = 2):
        b, n, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(1)
        self._check(self._lib.nufi_batch_autotune(self._b, n_probe, repeats, ctypes.byref(b), ctypes.byref(n),
                                                  ctypes.byref(c), _ptr(ms)))
        return (b.value, n.value, c.value), float(ms[0])

    def char_steps(self, n_begin: int, n_end: int) -> int:
        return sum(c.char_steps(n_begin, min(n_end, c.Nt + 1)) for c in self.confs if n_begin <= c.Nt)


__all__ = ["GpuScheduler", "BatchScheduler", "FP64", "FP32C"]
