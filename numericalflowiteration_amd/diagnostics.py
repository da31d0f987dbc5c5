"""Nodal field quantities of one history slice (host side, numpy).

A history slice holds the periodic order-4 B-spline coefficients of the
potential phi_n, padded to (N+3)^d reals (nufi/fields.hpp:63-142, 352-490).
At a grid node the cubic B-spline weights are (1/6, 4/6, 1/6, 0) for the
value and (-1/2, 0, 1/2, 0)/dx for the slope (nufi/splines.hpp:39-110 at
t = 0), so

  phi(x_i)  = (c[i] + 4 c[i+1] + c[i+2]) / 6
  E(x_i)    = -(c[i+2] - c[i]) / (2 dx)        (tensorised in 3D)

and, since the fit interpolates the Poisson solution at the nodes and the
Poisson solve divides by |k|^2 (nufi/poisson.cpp:66-89, 328-362), the mean-free
density of the step is rho = IFFT(|k|^2 FFT(phi)).  These are the quantities
the per-step rho / E / energy error of a run is measured on.
"""
from __future__ import annotations

import math

import numpy as np

from .config import Config

_VAL = (1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0)
_DER = (-0.5, 0.0, 0.5)


def _coeffs(conf: Config, slices) -> np.ndarray:
    """Slices as an array indexed [..., k, j, i] (3D) or [..., i] (1D), padded
    extent N+3 per axis; `slices` holds one slice or several (a leading axis)."""
    s = np.asarray(slices, dtype=np.float64)
    if conf.dim == 1:
        return s[..., : conf.Nx + 3]
    p = (conf.Nz + 3, conf.Ny + 3, conf.Nx + 3)
    return s[..., : p[0] * p[1] * p[2]].reshape(s.shape[:-1] + p)


def _stencil(c: np.ndarray, axis: int, w, n: int) -> np.ndarray:
    """sum_j w[j] c[i+j] along `axis` for i = 0..n-1."""
    out = 0.0
    for j, wj in enumerate(w):
        if wj != 0.0:
            out = out + wj * np.take(c, np.arange(j, j + n), axis=axis)
    return out


def _axes(conf: Config):
    """(axis, N, d_inv) of x, y, z in the coefficient array (x fastest)."""
    if conf.dim == 1:
        return [(-1, conf.Nx, conf.dx_inv)]
    return [(-1, conf.Nx, conf.dx_inv), (-2, conf.Ny, conf.dy_inv), (-3, conf.Nz, conf.dz_inv)]


def node_potential(conf: Config, slices) -> np.ndarray:
    """phi at the grid nodes: shape (..., Nx) in 1D, (..., Nz, Ny, Nx) in 3D."""
    c = _coeffs(conf, slices)
    for axis, n, _ in _axes(conf):
        c = _stencil(c, axis, _VAL, n)
    return c


def node_field(conf: Config, slices) -> np.ndarray:
    """E = -grad phi at the grid nodes: shape (..., Nx) in 1D, (..., 3, Nz, Ny, Nx)
    in 3D (components x, y, z)."""
    c = _coeffs(conf, slices)
    comps = []
    for der_axis, _, d_inv in _axes(conf):
        g = c
        for axis, n, _ in _axes(conf):
            g = _stencil(g, axis, _DER if axis == der_axis else _VAL, n)
        comps.append(-g * d_inv)
    return comps[0] if conf.dim == 1 else np.stack(comps, axis=-4)


def _k2(conf: Config) -> np.ndarray:
    def k(n, L):
        return 2.0 * math.pi * np.fft.fftfreq(n, d=L / n)
    if conf.dim == 1:
        return k(conf.Nx, conf.Lx) ** 2
    kz, ky, kx = np.meshgrid(k(conf.Nz, conf.Lz), k(conf.Ny, conf.Ly), k(conf.Nx, conf.Lx), indexing="ij")
    return kx * kx + ky * ky + kz * kz


def node_density(conf: Config, slices) -> np.ndarray:
    """The mean-free rho of the step whose Poisson solve and fit made each slice."""
    phi = node_potential(conf, slices)
    axes = (-1,) if conf.dim == 1 else (-3, -2, -1)
    return np.real(np.fft.ifftn(_k2(conf) * np.fft.fftn(phi, axes=axes), axes=axes))


def _per_step_rel(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """||a_n - b_n|| / ||b_n|| per leading index n (0 where b_n and a_n vanish)."""
    d = np.sqrt(np.sum((a - b).reshape(a.shape[0], -1) ** 2, axis=1))
    r = np.sqrt(np.sum(b.reshape(b.shape[0], -1) ** 2, axis=1))
    return np.where(r > 0, d / np.where(r > 0, r, 1.0), (d > 0).astype(float))


def run_errors(conf: Config, hist, energies, ref_hist, ref_energies) -> dict:
    """Per-step errors of a run against a reference run of the same workload: the
    maximum over the time steps n of the normwise relative error of the
    mean-free rho, of E at the nodes and of the electric energy, with the step
    each maximum is reached at (all steps 0..Nt of both histories)."""
    st = conf.stride_t()
    n = min(len(energies), len(ref_energies), np.size(hist) // st, np.size(ref_hist) // st)
    a = np.asarray(hist, dtype=np.float64)[: n * st].reshape(n, st)
    b = np.asarray(ref_hist, dtype=np.float64)[: n * st].reshape(n, st)
    er = _per_step_rel(node_density(conf, a), node_density(conf, b))
    eE = _per_step_rel(node_field(conf, a), node_field(conf, b))
    e, r = np.asarray(energies[:n], dtype=np.float64), np.asarray(ref_energies[:n], dtype=np.float64)
    ee = np.where(r != 0, np.abs(e - r) / np.where(r != 0, np.abs(r), 1.0), 0.0)
    return {"rho_max_rel_err": float(er.max()), "rho_err_at_step": int(er.argmax()),
            "E_max_rel_err": float(eE.max()), "E_err_at_step": int(eE.argmax()),
            "energy_max_rel_err": float(ee.max()), "energy_err_at_step": int(ee.argmax()),
            "energy_normwise_rel_err": float(np.linalg.norm(e - r) / np.linalg.norm(r)),
            "steps_compared": int(n)}


__all__ = ["node_potential", "node_field", "node_density", "run_errors"]
