"""Pins the oracle's 3D rho, eval_f, metrics and time loop (rho.hpp:318-462,
cuda_kernel.cu:392-466, bin/test_nufi_cpu_3d.cpp:63-76) by code that shares
nothing with it:

* an independent 3D NuFI written here in numpy: closed-form cubic B-spline
  pieces, a dense Kronecker collocation solve for the fit, numpy's complex FFT
  for the Poisson solve and its own spectral energy (1/2 int |grad phi|^2 by
  Parseval), vectorised Stormer-Verlet traces over all nodes;
* closed forms: zero-field rho as a brute-force sum of the shifted f0, the
  n = 0 electric energy of the three Landau modes, and the zero-field metrics
  with the weight du*dv*dw the reference's dim3 metrics kernel uses
  (cuda_kernel.cu:457, which leaves out dx*dy*dz).
"""
import math

import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import config as C
from tests import oracle_lib as O

M3_NORM = (2 * math.pi) ** -1.5


# ---- the independent model ----------------------------------------------------

def _basis(t):
    a = 1 - t
    return np.stack([a ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6,
                     t ** 3 / 6], axis=-1)


def _slope(t):
    a = 1 - t
    return np.stack([-a ** 2 / 2, (3 * t ** 2 - 4 * t) / 2, (-3 * t ** 2 + 2 * t + 1) / 2, t ** 2 / 2], axis=-1)


class Model3:
    """3D NuFI on the grid of a Config, order 4, in numpy (nothing of oracle/)."""

    def __init__(self, c):
        self.c = c
        self.N = (c.Nx, c.Ny, c.Nz)
        self.lo = (c.x_min, c.y_min, c.z_min)
        self.L = (c.x_max - c.x_min, c.y_max - c.y_min, c.z_max - c.z_min)
        self.V = (c.Nu, c.Nv, c.Nw)
        self.vlo = (c.u_min, c.v_min, c.w_min)
        self.dv = tuple((hi - lo) / n for lo, hi, n in
                        zip(self.vlo, (c.u_max, c.v_max, c.w_max), self.V))
        self.dt = c.dt
        self.a, self.k = c.f0.alpha, c.f0.k
        self.slices = []                     # padded (Nz+3, Ny+3, Nx+3) coefficient arrays
        # nodes: spatial index l = i + j Nx + k Nx Ny outer, velocity u fastest inner
        g = [lo + np.arange(n) * (L / n) for lo, L, n in zip(self.lo, self.L, self.N)]
        vg = [lo + (np.arange(n) + 0.5) * d for lo, n, d in zip(self.vlo, self.V, self.dv)]
        Z, Y, X = np.meshgrid(g[2], g[1], g[0], indexing="ij")
        W, Vv, U = np.meshgrid(vg[2], vg[1], vg[0], indexing="ij")
        self.x = [A.reshape(-1, 1) * np.ones((1, U.size)) for A in (X, Y, Z)]
        self.v = [np.ones((X.size, 1)) * A.reshape(1, -1) for A in (U, Vv, W)]

    def f0(self, x, v):
        s = 1 + self.a * sum(np.cos(self.k * xd) for xd in x)
        return M3_NORM * s * np.exp(-(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / 2)

    def efield(self, m, x):
        """-grad phi_m at the points x (a list of three arrays)."""
        sl = self.slices[m]
        idx, B, D = [], [], []
        for d in range(3):
            h = self.L[d] / self.N[d]
            s = np.mod(x[d] - self.lo[d], self.L[d]) / h
            i = np.floor(s)
            t = s - i
            i = i.astype(np.int64)
            i = np.where(i >= self.N[d], i - self.N[d], i)
            idx.append(i[..., None] + np.arange(4))
            B.append(_basis(t))
            D.append(_slope(t) / h)
        win = sl[idx[2][..., :, None, None], idx[1][..., None, :, None], idx[0][..., None, None, :]]
        ex = np.einsum("...cba,...a,...b,...c->...", win, D[0], B[1], B[2])
        ey = np.einsum("...cba,...a,...b,...c->...", win, B[0], D[1], B[2])
        ez = np.einsum("...cba,...a,...b,...c->...", win, B[0], B[1], D[2])
        return [-ex, -ey, -ez]

    def f(self, n, half_kick=False, x0=None, v0=None):
        """f of the traced characteristics (rho.hpp:318-367; half_kick: rho.hpp:369-426)."""
        x = [a.copy() for a in (self.x if x0 is None else x0)]
        v = [a.copy() for a in (self.v if v0 is None else v0)]
        if n == 0:
            return self.f0(x, v)
        if half_kick:
            E = self.efield(n, x)
            v = [v[d] + 0.5 * self.dt * E[d] for d in range(3)]
        for m in range(n - 1, 0, -1):
            x = [x[d] - self.dt * v[d] for d in range(3)]
            E = self.efield(m, x)
            v = [v[d] + self.dt * E[d] for d in range(3)]
        x = [x[d] - self.dt * v[d] for d in range(3)]
        E = self.efield(0, x)
        v = [v[d] + 0.5 * self.dt * E[d] for d in range(3)]
        return self.f0(x, v)

    def rho(self, n):
        w = self.dv[0] * self.dv[1] * self.dv[2]
        return 1 - w * self.f(n).sum(axis=1)

    def poisson(self, rho):
        """Nodal phi with -lap phi = rho - mean, and 1/2 int |grad phi|^2 (Parseval)."""
        Nx, Ny, Nz = self.N
        R = np.fft.fftn(rho.reshape(Nz, Ny, Nx))
        kz, ky, kx = np.meshgrid(*(2 * math.pi * np.fft.fftfreq(n, d=L / n) for n, L in
                                   zip((Nz, Ny, Nx), (self.L[2], self.L[1], self.L[0]))), indexing="ij")
        k2 = kx ** 2 + ky ** 2 + kz ** 2
        k2[0, 0, 0] = 1.0
        P = R / k2
        P[0, 0, 0] = 0.0
        energy = 0.5 * self.L[0] * self.L[1] * self.L[2] * float(np.sum(np.abs(R) ** 2 / k2 * (k2 > 0))
                                                                  - np.abs(R[0, 0, 0]) ** 2) / (Nx * Ny * Nz) ** 2
        return np.real(np.fft.ifftn(P)), energy

    def fit(self, phi):
        """Dense solve of sum_{abc} B_a B_b B_c C[k+c, j+b, i+a] = phi[k, j, i], then periodic padding."""
        def circ(n):
            A = np.zeros((n, n))
            for i in range(n):
                for j, w in enumerate((1 / 6, 4 / 6, 1 / 6, 0.0)):
                    A[i, (i + j) % n] += w
            return A
        Nx, Ny, Nz = self.N
        A = np.kron(circ(Nz), np.kron(circ(Ny), circ(Nx)))
        cc = np.linalg.solve(A, phi.reshape(-1)).reshape(Nz, Ny, Nx)
        return cc[np.arange(Nz + 3)[:, None, None] % Nz, np.arange(Ny + 3)[None, :, None] % Ny,
                  np.arange(Nx + 3)[None, None, :] % Nx]

    def step(self, n):
        r = self.rho(n)
        phi, e = self.poisson(r)
        self.slices.append(self.fit(phi))
        assert len(self.slices) == n + 1
        return e, r

    def load(self, hist, n_slices):
        st = np.prod([n + 3 for n in self.N])
        self.slices = [hist[m * st:(m + 1) * st].reshape(self.N[2] + 3, self.N[1] + 3, self.N[0] + 3)
                       for m in range(n_slices)]


def small_3d(alpha=0.2, N=(8, 4, 6), V=(4, 6, 4), T=1.0, kind=C.F0_LANDAU_3D):
    return nufi.config_3d(N, V, 1 / 8, T, (0.0, 4 * math.pi), (-6.0, 6.0), kind, alpha=alpha, k=0.5)


# ---- the time loop, rho, f and metrics against the independent model -----------

@pytest.fixture(scope="module")
def run_pair():
    c = small_3d()
    M = Model3(c)
    e_ref, r_ref = [], []
    for n in range(c.Nt + 1):
        e, r = M.step(n)
        e_ref.append(e)
        r_ref.append(r)
    hist = O.new_history(c, c.Nt + 1)
    e_or, r_or = [], []
    for n in range(c.Nt + 1):
        e, r = O.step(c, n, hist)
        e_or.append(e)
        r_or.append(r)
    return c, M, np.array(e_ref), r_ref, hist, np.array(e_or), r_or


def test_run3d_against_independent_model(run_pair):
    c, M, e_ref, r_ref, hist, e_or, r_or = run_pair
    assert e_ref[0] > 1e-3                                   # the fields are not weak
    assert np.max(np.abs(e_or - e_ref) / e_ref) < 1e-12, (e_or, e_ref)
    for n in range(c.Nt + 1):
        a = r_or[n] - r_or[n].mean()
        b = r_ref[n] - r_ref[n].mean()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), n
    st = c.stride_t()
    for m in range(c.Nt + 1):
        got = hist[m * st:(m + 1) * st]
        ref = M.slices[m].reshape(-1)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), m


def test_rho3d_on_a_history_against_independent_model(run_pair):
    """rho3d at every step on the oracle's own history, both v-sum orders."""
    c, M, _, _, hist, _, _ = run_pair
    M2 = Model3(c)
    M2.load(hist, c.Nt + 1)
    for n in [1, 4, c.Nt]:
        ref = M2.rho(n)
        for exact in (False, True):
            got = O.rho(c, n, hist, exact_sum=exact)
            assert np.linalg.norm((got - got.mean()) - (ref - ref.mean())) <= 1e-12 * np.linalg.norm(ref - ref.mean())
            assert abs(got.mean() - ref.mean()) < 1e-14


def test_f3d_and_ftilda3d_pointwise_against_independent_model(run_pair):
    """trace3d without (rho.hpp:318-367) and with (rho.hpp:369-426) the initial half-kick."""
    c, M, _, _, hist, _, _ = run_pair
    M2 = Model3(c)
    M2.load(hist, c.Nt + 1)
    pts = [(0.3, 2.0, 5.5, 1.2, -0.7, 2.9), (11.9, 0.05, 7.1, -3.3, 0.4, -0.2), (6.2, 9.9, 0.0, 0.0, 5.1, -4.8)]
    for n in [1, 2, 5, c.Nt]:
        for hk in (False, True):
            for p in pts:
                x0 = [np.array([p[0]]), np.array([p[1]]), np.array([p[2]])]
                v0 = [np.array([p[3]]), np.array([p[4]]), np.array([p[5]])]
                ref = float(M2.f(n, half_kick=hk, x0=x0, v0=v0)[0])
                got = O.trace3d(c, n, list(p), hist, half_kick=hk)
                assert abs(got - ref) <= 1e-12 * abs(ref), (n, hk, p, got, ref)


def test_metrics3d_against_independent_model(run_pair):
    """[sum w f, sum w f^2, sum w |v|^2 f/2, -sum w f log f] over every node,
    w = du*dv*dw (cuda_kernel.cu:454-464), f with the initial half-kick."""
    c, M, _, _, hist, _, _ = run_pair
    M2 = Model3(c)
    M2.load(hist, c.Nt + 1)
    w = M2.dv[0] * M2.dv[1] * M2.dv[2]
    v2 = M2.v[0] ** 2 + M2.v[1] ** 2 + M2.v[2] ** 2
    for n in [0, 3, c.Nt]:
        f = M2.f(n, half_kick=True)
        ref = np.array([w * f.sum(), w * (f * f).sum(), w * (v2 * f).sum() / 2,
                        -w * np.where(f > 0, f * np.log(np.where(f > 0, f, 1.0)), 0.0).sum()])
        got = O.metrics(c, n, hist)
        assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref)), (n, got, ref)


# ---- closed forms ----------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 3, 7])
def test_rho3d_zero_field_closed_form(n):
    """All slices zero: free streaming, rho = 1 - du dv dw sum_v f0(x - n dt v, v)."""
    c = small_3d(alpha=0.3, N=(4, 6, 8), V=(6, 4, 4), T=1.0)
    hist = O.new_history(c, c.Nt + 1)
    dv = [(c.u_max - c.u_min) / c.Nu, (c.v_max - c.v_min) / c.Nv, (c.w_max - c.w_min) / c.Nw]
    vg = [lo + (np.arange(m) + 0.5) * d for lo, m, d in zip((c.u_min, c.v_min, c.w_min), (c.Nu, c.Nv, c.Nw), dv)]
    W, Vv, U = np.meshgrid(vg[2], vg[1], vg[0], indexing="ij")
    U, Vv, W = U.ravel(), Vv.ravel(), W.ravel()
    ref = np.zeros(c.Nx * c.Ny * c.Nz)
    for l in range(ref.size):
        i, j, k = l % c.Nx, (l // c.Nx) % c.Ny, l // (c.Nx * c.Ny)
        x, y, z = c.x_min + i * c.dx, c.y_min + j * c.dy, c.z_min + k * c.dz
        t = n * c.dt
        f = M3_NORM * (1 + 0.3 * (np.cos(0.5 * (x - t * U)) + np.cos(0.5 * (y - t * Vv)) + np.cos(0.5 * (z - t * W)))) \
            * np.exp(-(U * U + Vv * Vv + W * W) / 2)
        ref[l] = 1 - dv[0] * dv[1] * dv[2] * math.fsum(f)
    for exact in (False, True):
        got = O.rho(c, n, hist, exact_sum=exact)
        assert np.max(np.abs(got - ref)) < 1e-14, (exact, np.max(np.abs(got - ref)))
        # the mean-free part carries the perturbation
        assert np.linalg.norm((got - got.mean()) - (ref - ref.mean())) <= 1e-13 * np.linalg.norm(ref - ref.mean())


def test_energy3d_at_step0_closed_form():
    """rho(0) = 1 - m (1 + alpha sum_d cos(k x_d)), m the discrete velocity mass;
    each of the three modes carries L^3 (alpha m)^2 / (4 k^2)."""
    c = nufi.landau_3d(N=8, V=16, T=1.0)
    hist = O.new_history(c, c.Nt + 1)
    e0, r0 = O.step(c, 0, hist)
    du = 12.0 / 16
    u = -6 + (np.arange(16) + 0.5) * du
    m = (du * np.sum(np.exp(-u * u / 2)) / math.sqrt(2 * math.pi)) ** 3
    L3 = c.Lx * c.Ly * c.Lz
    ref = 3 * L3 * (0.01 * m) ** 2 / (4 * 0.25)
    assert abs(e0 - ref) <= 1e-12 * ref, (e0, ref)
    assert abs(m - 1) < 1e-8                                  # truncation of [-6, 6]: 3 erfc(6/sqrt 2)
    # rho itself, per point
    g = np.arange(8) * c.dx
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    r_ref = 1 - m * (1 + 0.01 * (np.cos(0.5 * X) + np.cos(0.5 * Y) + np.cos(0.5 * Z)))
    assert np.max(np.abs(r0 - r_ref.ravel())) < 1e-13          # a sequential sum of 4096 values


def _vel_sums(c):
    out = []
    for lo, hi, n in ((c.u_min, c.u_max, c.Nu), (c.v_min, c.v_max, c.Nv), (c.w_min, c.w_max, c.Nw)):
        d = (hi - lo) / n
        u = lo + (np.arange(n) + 0.5) * d
        M = np.exp(-u * u / 2) / math.sqrt(2 * math.pi)
        out.append((d * M.sum(), d * (u * u * M).sum(), d * (M * M).sum()))
    return out


def test_metrics3d_zero_field_closed_forms():
    """alpha = 0, zero field: f = M3(v) at every node; metrics weight du dv dw
    over Nx Ny Nz points (so the mass is Nx Ny Nz, not the volume)."""
    c = nufi.config_3d((4, 6, 8), (32, 24, 32), 1 / 8, 1.0, (0.0, 4 * math.pi), (-6.0, 6.0), C.F0_LANDAU_3D,
                       alpha=0.0, k=0.5)
    hist = O.new_history(c, c.Nt + 1)
    P = c.Nx * c.Ny * c.Nz
    (s0u, s2u, q0u), (s0v, s2v, q0v), (s0w, s2w, q0w) = _vel_sums(c)
    mass = P * s0u * s0v * s0w
    f2 = P * q0u * q0v * q0w
    kin = P * 0.5 * (s2u * s0v * s0w + s0u * s2v * s0w + s0u * s0v * s2w)
    ent = 1.5 * math.log(2 * math.pi) * mass + kin
    ref = np.array([mass, f2, kin, ent])
    for n in [0, 5]:
        got = O.metrics(c, n, hist)
        assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref)), (n, got, ref)
    # and against the continuous limits (truncation of [-6, 6] ~ 2e-9 per axis)
    lim = np.array([P, P / (2 * math.sqrt(math.pi)) ** 3, 1.5 * P, P * (1.5 * math.log(2 * math.pi) + 1.5)])
    assert np.all(np.abs(O.metrics(c, 0, hist) - lim) <= 1e-7 * lim)


def test_metrics3d_free_streaming_of_a_perturbed_f0():
    """Zero field, alpha != 0: f = f0(x - n dt v, v); the brute-force sums over
    the nodes, and the mass stays Nx Ny Nz s0^3 (full periods on the grid)."""
    c = small_3d(alpha=0.3, N=(8, 4, 4), V=(8, 6, 6), T=1.0)
    hist = O.new_history(c, c.Nt + 1)
    M = Model3(c)
    M.slices = [np.zeros((c.Nz + 3, c.Ny + 3, c.Nx + 3))] * (c.Nt + 1)
    w = M.dv[0] * M.dv[1] * M.dv[2]
    v2 = M.v[0] ** 2 + M.v[1] ** 2 + M.v[2] ** 2
    (s0u, _, _), (s0v, _, _), (s0w, _, _) = _vel_sums(c)
    for n in [0, 2, 8]:
        t = n * c.dt
        f = M.f0([M.x[d] - t * M.v[d] for d in range(3)], M.v)
        ref = np.array([w * f.sum(), w * (f * f).sum(), w * (v2 * f).sum() / 2, -w * (f * np.log(f)).sum()])
        got = O.metrics(c, n, hist)
        assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref)), (n, got, ref)
        assert abs(got[0] - c.Nx * c.Ny * c.Nz * s0u * s0v * s0w) <= 1e-12 * got[0]
