"""The coefficient history on disk (the file functions of include/nufi_hip.h).

Binary files ("NUFIHST1") are lossless and carry the energies of the stored
steps; text files follow the layout of the reference's potential files
(bin/test_nufi_cpu_3d_isolated.cpp:128-135, 158-162, 190-211: seven header
lines, then one value per line), written here with 17 significant digits.
These functions run on the host only.  ``GpuScheduler.save_history`` /
``load_history`` move a context's device history to and from such files
(checkpoint / resume, isolated steps).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi
from ._capi import HISTORY_BINARY, HISTORY_TEXT, HistoryInfo, NufiError, NufiIOError

_dp = ctypes.POINTER(ctypes.c_double)

FORMATS = {"binary": HISTORY_BINARY, "text": HISTORY_TEXT}


def _fmt(fmt) -> int:
    return FORMATS[fmt] if isinstance(fmt, str) else int(fmt)


def _raise(rc: int, what: str, path: str):
    if rc == _capi.ERR_IO:
        raise NufiIOError(rc, f"{what}: {path}")
    raise NufiError(rc, f"{what}: {path}")


def info_dict(h: HistoryInfo) -> dict:
    return {f: getattr(h, f) for f, _ in HistoryInfo._fields_}


def read_info(path: str) -> tuple[dict, str]:
    """Header of a history file and its format ("binary" or "text")."""
    h = HistoryInfo()
    fmt = ctypes.c_int(0)
    rc = _capi.load().nufi_history_read_info(os.fsencode(path), ctypes.byref(h), ctypes.byref(fmt))
    if rc:
        _raise(rc, "cannot read the history header", path)
    return info_dict(h), ("binary" if fmt.value == HISTORY_BINARY else "text")


def read_history(path: str) -> tuple[dict, np.ndarray, np.ndarray]:
    """(header, slices as one float64 array of n_slices*stride_t, energies)."""
    info, _ = read_info(path)
    slices = np.zeros(info["n_slices"] * info["stride_t"])
    energies = np.zeros(info["n_energies"])
    h = HistoryInfo()
    rc = _capi.load().nufi_history_read(
        os.fsencode(path), ctypes.byref(h),
        slices.ctypes.data_as(_dp) if slices.size else None, info["n_slices"],
        energies.ctypes.data_as(_dp) if energies.size else None, info["n_energies"])
    if rc:
        _raise(rc, "cannot read the history", path)
    return info_dict(h), slices, energies


def write_history(path: str, info: dict, slices: np.ndarray, energies: np.ndarray | None = None,
                  fmt="binary"):
    """Write slices 0..info['n_slices']-1 (and, binary, the energies)."""
    h = HistoryInfo(**{k: v for k, v in info.items() if k in dict(HistoryInfo._fields_)})
    slices = np.ascontiguousarray(slices, dtype=np.float64)
    if slices.size < h.n_slices * h.stride_t:
        raise ValueError("fewer values than n_slices * stride_t")
    e = None
    if energies is not None and _fmt(fmt) == HISTORY_BINARY:
        energies = np.ascontiguousarray(energies, dtype=np.float64)
        h.n_energies = energies.size
        e = energies.ctypes.data_as(_dp) if energies.size else None
    else:
        h.n_energies = 0
    rc = _capi.load().nufi_history_write(os.fsencode(path), _fmt(fmt), ctypes.byref(h),
                                         slices.ctypes.data_as(_dp) if slices.size else None, e)
    if rc:
        _raise(rc, "cannot write the history", path)
