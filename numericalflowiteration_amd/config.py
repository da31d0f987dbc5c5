"""Grid descriptions for the NuFI rho path.

`Config` is the ctypes mirror of `nufi_config` (include/nufi_hip.h), which
carries the fields of the reference's ``dim1::config_t`` / ``dim3::config_t``
(nufi/config.hpp:34-53, 168-194) plus a runtime initial-condition family
(the reference hard-codes ``f0`` at config.hpp:72-84, 221-247).

`derive()` fills the dependent fields in the order of the reference ctor
(config.hpp:67-69, 213-218): L = max - min, L_inv = 1/L, d = L/N,
d_inv = 1/d, du = (u_max - u_min)/Nu.

The workload constructors below are the five configurations this build is
measured on (1D Landau, 1D two-stream, the 512-run sweep, the adjoint case and
3D Landau).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

MAX_MODES = 8

F0_LANDAU_1D = 1
F0_TWOSTREAM_1D = 2
F0_MULTIMODE_1D = 3
F0_LANDAU_3D = 11
F0_BUMPONTAIL_3D = 12
F0_TWOSTREAM_3D = 13

FP64 = 0
FP32C = 1


class F0(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_modes", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("k", ctypes.c_double),
        ("a", ctypes.c_double * MAX_MODES),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("Nx", ctypes.c_uint64), ("Ny", ctypes.c_uint64), ("Nz", ctypes.c_uint64),
        ("Nu", ctypes.c_uint64), ("Nv", ctypes.c_uint64), ("Nw", ctypes.c_uint64),
        ("Nt", ctypes.c_uint64),
        ("dt", ctypes.c_double),
        ("x_min", ctypes.c_double), ("x_max", ctypes.c_double),
        ("y_min", ctypes.c_double), ("y_max", ctypes.c_double),
        ("z_min", ctypes.c_double), ("z_max", ctypes.c_double),
        ("u_min", ctypes.c_double), ("u_max", ctypes.c_double),
        ("v_min", ctypes.c_double), ("v_max", ctypes.c_double),
        ("w_min", ctypes.c_double), ("w_max", ctypes.c_double),
        ("dx", ctypes.c_double), ("dx_inv", ctypes.c_double),
        ("Lx", ctypes.c_double), ("Lx_inv", ctypes.c_double),
        ("dy", ctypes.c_double), ("dy_inv", ctypes.c_double),
        ("Ly", ctypes.c_double), ("Ly_inv", ctypes.c_double),
        ("dz", ctypes.c_double), ("dz_inv", ctypes.c_double),
        ("Lz", ctypes.c_double), ("Lz_inv", ctypes.c_double),
        ("du", ctypes.c_double), ("dv", ctypes.c_double), ("dw", ctypes.c_double),
        ("f0", F0),
    ]

    # ---- derived quantities -------------------------------------------
    def derive(self) -> "Config":
        """Fill L*, d*, *_inv, du, dv, dw (config.hpp:67-69, 213-218)."""
        self.Lx = self.x_max - self.x_min
        self.Lx_inv = 1.0 / self.Lx
        self.dx = self.Lx / self.Nx
        self.dx_inv = 1.0 / self.dx
        self.du = (self.u_max - self.u_min) / self.Nu
        if self.dim == 3:
            self.Ly = self.y_max - self.y_min
            self.Ly_inv = 1.0 / self.Ly
            self.Lz = self.z_max - self.z_min
            self.Lz_inv = 1.0 / self.Lz
            self.dy = self.Ly / self.Ny
            self.dy_inv = 1.0 / self.dy
            self.dz = self.Lz / self.Nz
            self.dz_inv = 1.0 / self.dz
            self.dv = (self.v_max - self.v_min) / self.Nv
            self.dw = (self.w_max - self.w_min) / self.Nw
        return self

    def copy(self) -> "Config":
        c = Config()
        ctypes.pointer(c)[0] = self
        return c

    # ---- sizes ----------------------------------------------------------
    def stride_t(self, order: int = 4) -> int:
        """Reals per history slice: (Nx+order-1)^dim (rho.hpp:39, 329)."""
        if self.dim == 1:
            return self.Nx + order - 1
        return (self.Nx + order - 1) * (self.Ny + order - 1) * (self.Nz + order - 1)

    def n_space(self) -> int:
        return self.Nx if self.dim == 1 else self.Nx * self.Ny * self.Nz

    def n_vel(self) -> int:
        return self.Nu if self.dim == 1 else self.Nu * self.Nv * self.Nw

    def n_nodes(self) -> int:
        return self.n_space() * self.n_vel()

    def char_steps(self, n_begin: int, n_end: int) -> int:
        """Characteristic steps of time steps [n_begin, n_end): step n costs n per node."""
        s = (n_end - 1) * n_end // 2 - (n_begin - 1) * n_begin // 2 if n_begin > 0 else (n_end - 1) * n_end // 2
        return self.n_nodes() * s

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        if self.dim == 1:
            return (f"Config(1D Nx={self.Nx} Nu={self.Nu} Nt={self.Nt} dt={self.dt} "
                    f"x=[{self.x_min},{self.x_max}) u=[{self.u_min},{self.u_max}] f0={self.f0.kind})")
        return (f"Config(3D N=({self.Nx},{self.Ny},{self.Nz}) V=({self.Nu},{self.Nv},{self.Nw}) "
                f"Nt={self.Nt} dt={self.dt} f0={self.f0.kind})")


def config_1d(Nx: int, Nu: int, dt: float, T: float, x_min: float, x_max: float,
              u_min: float, u_max: float, f0_kind: int, alpha: float = 0.0, k: float = 0.0,
              modes=None) -> Config:
    c = Config()
    c.dim = 1
    c.Nx, c.Ny, c.Nz = Nx, 1, 1
    c.Nu, c.Nv, c.Nw = Nu, 1, 1
    c.dt = dt
    c.Nt = int(T / dt)
    c.x_min, c.x_max = x_min, x_max
    c.u_min, c.u_max = u_min, u_max
    c.f0.kind = f0_kind
    c.f0.alpha = alpha
    c.f0.k = k
    if modes is not None:
        modes = list(modes)
        if len(modes) > MAX_MODES:
            raise ValueError(f"at most {MAX_MODES} modes")
        c.f0.n_modes = len(modes)
        for i, a in enumerate(modes):
            c.f0.a[i] = a
    return c.derive()


def config_3d(N, V, dt: float, T: float, x_lim, v_lim, f0_kind: int,
              alpha: float = 0.0, k: float = 0.0) -> Config:
    c = Config()
    c.dim = 3
    c.Nx, c.Ny, c.Nz = N
    c.Nu, c.Nv, c.Nw = V
    c.dt = dt
    c.Nt = int(T / dt)
    c.x_min, c.x_max = x_lim
    c.y_min, c.y_max = x_lim
    c.z_min, c.z_max = x_lim
    c.u_min, c.u_max = v_lim
    c.v_min, c.v_max = v_lim
    c.w_min, c.w_max = v_lim
    c.f0.kind = f0_kind
    c.f0.alpha = alpha
    c.f0.k = k
    return c.derive()


# ---- the measured workloads --------------------------------------------------

def landau_1d(Nx: int = 64, Nu: int = 256, T: float = 30.0) -> Config:
    """Linear Landau damping: x in [0,4pi), (1+0.01cos(0.5x))M(v), v in [-10,10], dt=1/16."""
    return config_1d(Nx, Nu, 1.0 / 16.0, T, 0.0, 4 * math.pi, -10.0, 10.0,
                     F0_LANDAU_1D, alpha=0.01, k=0.5)


def twostream_1d(Nx: int = 256, Nu: int = 1024, T: float = 50.0) -> Config:
    """Two-stream instability: x in [0,10pi), (1+0.001cos(0.2x))v^2 M(v), v in [-8,8], dt=1/32."""
    return config_1d(Nx, Nu, 1.0 / 32.0, T, 0.0, 10 * math.pi, -8.0, 8.0,
                     F0_TWOSTREAM_1D, alpha=0.001, k=0.2)


def multimode_1d(a, Nx: int = 128, Nu: int = 512, T: float = 20.0) -> Config:
    """Adjoint case: x in [0,4pi), (1+sum_m a_m cos(0.5 m x))M(v), v in [-10,10], dt=1/16."""
    return config_1d(Nx, Nu, 1.0 / 16.0, T, 0.0, 4 * math.pi, -10.0, 10.0,
                     F0_MULTIMODE_1D, alpha=0.0, k=0.5, modes=a)


def landau_3d(N: int = 16, V: int = 32, T: float = 20.0) -> Config:
    """3D Landau: x in [0,4pi)^3, (1+0.01 sum_d cos(0.5x_d))M3(v), v in [-6,6]^3, dt=1/8."""
    return config_3d((N, N, N), (V, V, V), 1.0 / 8.0, T, (0.0, 4 * math.pi), (-6.0, 6.0),
                     F0_LANDAU_3D, alpha=0.01, k=0.5)


# ---- seeded parameters (SplitMix64 -> U[0,1) = (r >> 11) * 2^-53) ----------------

SWEEP_SEED = 0x4E7546495F4D4935
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform doubles from the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


@dataclass
class SweepRun:
    alpha: float
    k: float
    Nx: int
    twostream: bool


def sweep_runs(n_runs: int = 512, seed: int = SWEEP_SEED):
    """Parameters of the batched sweep: alpha~U[0.001,0.5], k~U[0.3,0.6],
    Nx uniform in {32,64,128,256}, Maxwellian or two-stream 50/50."""
    rng = SplitMix64(seed)
    runs = []
    for _ in range(n_runs):
        alpha = 0.001 + (0.5 - 0.001) * rng.uniform()
        k = 0.3 + (0.6 - 0.3) * rng.uniform()
        Nx = (32, 64, 128, 256)[min(3, int(4 * rng.uniform()))]
        two = rng.uniform() < 0.5
        runs.append(SweepRun(alpha, k, Nx, two))
    return runs


def sweep_config(run: SweepRun, T: float = 20.0) -> Config:
    """One run of the sweep: domain [0, 2pi/k), Nv = 4 Nx, v in [-10,10], dt = 1/16."""
    kind = F0_TWOSTREAM_1D if run.twostream else F0_LANDAU_1D
    return config_1d(run.Nx, 4 * run.Nx, 1.0 / 16.0, T, 0.0, 2 * math.pi / run.k, -10.0, 10.0,
                     kind, alpha=run.alpha, k=run.k)


def adjoint_modes(seed: int = SWEEP_SEED + 1, n: int = 8):
    """The adjoint case's amplitudes a_m ~ U[0, 0.02], m = 1..8."""
    rng = SplitMix64(seed)
    return [0.02 * rng.uniform() for _ in range(n)]
