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


def _grid_shape(conf: Config):
    if conf.dim == 1:
        return (conf.Nx,)
    return (conf.Nz, conf.Ny, conf.Nx)


def _coeffs(conf: Config, slice_) -> np.ndarray:
    """The slice as an array indexed [k, j, i] (3D) or [i] (1D), padded extent N+3."""
    s = np.asarray(slice_, dtype=np.float64)
    if conf.dim == 1:
        return s[: conf.Nx + 3]
    return s[: (conf.Nx + 3) * (conf.Ny + 3) * (conf.Nz + 3)].reshape(conf.Nz + 3, conf.Ny + 3, conf.Nx + 3)


def _stencil(c: np.ndarray, axis: int, w, n: int) -> np.ndarray:
    """sum_j w[j] c[i+j] along `axis` for i = 0..n-1."""
    out = 0.0
    for j, wj in enumerate(w):
        if wj != 0.0:
            out = out + wj * np.take(c, np.arange(j, j + n), axis=axis)
    return out


def node_potential(conf: Config, slice_) -> np.ndarray:
    """phi at the grid nodes (shape (Nx,) or (Nz, Ny, Nx))."""
    c = _coeffs(conf, slice_)
    if conf.dim == 1:
        return _stencil(c, 0, _VAL, conf.Nx)
    c = _stencil(c, 2, _VAL, conf.Nx)
    c = _stencil(c, 1, _VAL, conf.Ny)
    return _stencil(c, 0, _VAL, conf.Nz)


def node_field(conf: Config, slice_) -> np.ndarray:
    """E = -grad phi at the grid nodes: shape (Nx,) in 1D, (3, Nz, Ny, Nx) in 3D
    (components x, y, z)."""
    c = _coeffs(conf, slice_)
    if conf.dim == 1:
        return -_stencil(c, 0, _DER, conf.Nx) * conf.dx_inv
    comps = []
    for axis_der, d_inv in ((2, conf.dx_inv), (1, conf.dy_inv), (0, conf.dz_inv)):
        g = c
        for axis, n in ((2, conf.Nx), (1, conf.Ny), (0, conf.Nz)):
            g = _stencil(g, axis, _DER if axis == axis_der else _VAL, n)
        comps.append(-g * d_inv)
    return np.stack(comps)


def _k2(conf: Config) -> np.ndarray:
    def k(n, L):
        return 2.0 * math.pi * np.fft.fftfreq(n, d=L / n)
    if conf.dim == 1:
        return k(conf.Nx, conf.Lx) ** 2
    kz, ky, kx = np.meshgrid(k(conf.Nz, conf.Lz), k(conf.Ny, conf.Ly), k(conf.Nx, conf.Lx), indexing="ij")
    return kx * kx + ky * ky + kz * kz


def node_density(conf: Config, slice_) -> np.ndarray:
    """The mean-free rho of the step whose Poisson solve and fit made this slice."""
    phi = node_potential(conf, slice_)
    return np.real(np.fft.ifftn(_k2(conf) * np.fft.fftn(phi)))


def run_errors(conf: Config, hist, energies, ref_hist, ref_energies) -> dict:
    """Per-step errors of a run against a reference run of the same workload: the
    maximum over the time steps n of the normwise relative error of the
    mean-free rho, of E at the nodes and of the electric energy, with the step
    each maximum is reached at (all steps 0..Nt of both histories)."""
    st = conf.stride_t()
    n_slices = min(len(energies), len(ref_energies), np.size(hist) // st, np.size(ref_hist) // st)
    worst = {"rho": (0.0, 0), "E": (0.0, 0), "energy": (0.0, 0)}
    for n in range(n_slices):
        a, b = hist[n * st:(n + 1) * st], ref_hist[n * st:(n + 1) * st]
        for key, fn in (("rho", node_density), ("E", node_field)):
            x, y = fn(conf, a), fn(conf, b)
            den = np.linalg.norm(y)
            e = float(np.linalg.norm(x - y) / den) if den > 0 else float(np.linalg.norm(x - y) > 0)
            if e > worst[key][0]:
                worst[key] = (e, n)
        e = abs(energies[n] - ref_energies[n]) / abs(ref_energies[n]) if ref_energies[n] != 0 else 0.0
        if e > worst["energy"][0]:
            worst["energy"] = (float(e), n)
    return {"rho_max_rel_err": worst["rho"][0], "rho_err_at_step": worst["rho"][1],
            "E_max_rel_err": worst["E"][0], "E_err_at_step": worst["E"][1],
            "energy_max_rel_err": worst["energy"][0], "energy_err_at_step": worst["energy"][1],
            "steps_compared": n_slices}


__all__ = ["node_potential", "node_field", "node_density", "run_errors"]
