"""Algorithmic FLOPs per characteristic step, the figure roofline.achieved uses
(bench.py FLOPS_1D = 30, FLOPS_3D = 475), counted by running one step of the
minimal shared-basis formulation on an op-counting scalar.

Counting rules: add, subtract and multiply 1; a multiply-add 2; floor 1;
sign flips and integer conversions 0.  The 1D step is drift (2), wrap and
locate (8), the four derivative weights (10), the 4-term contraction (one
multiply, three multiply-adds: 7), the dx_inv scale (1) and the kick (2).  The
3D step is drift (6), locate (3 x 8), value and derivative weights per axis
(3 x 20, sharing 1-t, (1-t)^2 and t^2), the contractions as 188 multiply-adds
(2 x 64 over x, 3 x 16 over y, 3 x 4 over z), the scales (3) and the kick (6).

The same code run on plain values is checked against the oracle's spline
evaluation (oracle_field1d / oracle_field3d, the reference's de Boor form), so
the counted formulation computes the field the step needs.  The reference's own
formulation, counted on the oracle's templates with the same rules
(oracle/opcount.hpp; operations among literals and configuration values are
hoistable and free), costs 42 (1D) and 848 (3D) per step."""
import math

import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from tests import oracle_lib as O


class Counter:
    n = 0


class Q:
    """A float that counts the arithmetic done on it."""
    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v)

    def __add__(self, o):
        Counter.n += 1
        return Q(self.v + val(o))

    __radd__ = __add__

    def __sub__(self, o):
        Counter.n += 1
        return Q(self.v - val(o))

    def __rsub__(self, o):
        Counter.n += 1
        return Q(val(o) - self.v)

    def __mul__(self, o):
        Counter.n += 1
        return Q(self.v * val(o))

    __rmul__ = __mul__

    def __neg__(self):
        return Q(-self.v)


def val(a):
    return a.v if isinstance(a, Q) else float(a)


def fma(a, b, c):
    Counter.n += 2
    return Q(val(a) * val(b) + val(c))


def floor(a):
    Counter.n += 1
    return Q(math.floor(val(a)))


def locate(x, x_min, d_inv, N):
    """Cell units, exact periodic wrap into [0, N), cell and offset: 8."""
    y = (x - x_min) * d_inv
    w = floor(y * (1.0 / N))
    y = fma(-N, w, y)
    k = floor(y)
    return int(val(k)), y - k


def deriv_weights(t):
    """N'_0..N'_3 of the uniform cubic B-spline at t, plus the shared terms: 10."""
    a = 1.0 - t
    a2 = a * a
    t2 = t * t
    d0 = -0.5 * a2
    d3 = 0.5 * t2
    d1 = fma(1.5, t, -2.0) * t
    d2 = -(d0 + d1 + d3)
    return (d0, d1, d2, d3), (a, a2, t2)


def value_weights(t, shared):
    """N_0..N_3 from the shared terms: 10."""
    a, a2, t2 = shared
    n0 = (a2 * a) * (1.0 / 6.0)
    t3 = t2 * t
    n3 = t3 * (1.0 / 6.0)
    n1 = fma(0.5, t3, -t2) + (2.0 / 3.0)
    n2 = ((1.0 - n0) - n1) - n3
    return (n0, n1, n2, n3)


def step1d(conf, c, x, u, dt):
    x = fma(-dt, u, x)                                         # drift
    i, t = locate(x, conf.x_min, conf.dx_inv, conf.Nx)
    d, _ = deriv_weights(t)
    g = float(c[i]) * d[0]
    for k in (1, 2, 3):
        g = fma(float(c[i + k]), d[k], g)
    e = g * conf.dx_inv                                        # d phi / dx
    u = fma(dt, e, u)                                          # kick (the sign convention does not change the count)
    return x, u, e


def step3d(conf, c, pos, vel, dt):
    N = (conf.Nx, conf.Ny, conf.Nz)
    lo = (conf.x_min, conf.y_min, conf.z_min)
    dinv = (conf.dx_inv, conf.dy_inv, conf.dz_inv)
    pos = [fma(-dt, vel[a], pos[a]) for a in range(3)]         # drift
    cell, B, D = [], [], []
    for a in range(3):
        i, t = locate(pos[a], lo[a], dinv[a], N[a])
        d, shared = deriv_weights(t)
        cell.append(i)
        D.append(d)
        B.append(value_weights(t, shared))
    sy, sz = conf.Nx + 3, (conf.Nx + 3) * (conf.Ny + 3)
    ix, jy, kz = cell
    A = [[0.0] * 4 for _ in range(4)]
    X = [[0.0] * 4 for _ in range(4)]
    for k in range(4):
        for j in range(4):
            base = (kz + k) * sz + (jy + j) * sy + ix
            for i in range(4):
                A[j][k] = fma(float(c[base + i]), B[0][i], A[j][k])
                X[j][k] = fma(float(c[base + i]), D[0][i], X[j][k])
    P, Qy, R = [0.0] * 4, [0.0] * 4, [0.0] * 4
    for k in range(4):
        for j in range(4):
            P[k] = fma(A[j][k], B[1][j], P[k])
            Qy[k] = fma(A[j][k], D[1][j], Qy[k])
            R[k] = fma(X[j][k], B[1][j], R[k])
    gx = gy = gz = 0.0
    for k in range(4):
        gx = fma(R[k], B[2][k], gx)
        gy = fma(Qy[k], B[2][k], gy)
        gz = fma(P[k], D[2][k], gz)
    e = (gx * dinv[0], gy * dinv[1], gz * dinv[2])
    vel = [fma(dt, e[a], vel[a]) for a in range(3)]            # kick
    return pos, vel, e


def test_1d_char_step_is_30_flops():
    conf = nufi.landau_1d(Nx=16, Nu=32, T=1.0)
    c = np.random.default_rng(1).standard_normal(conf.stride_t())
    Counter.n = 0
    step1d(conf, c, Q(1.3), Q(0.7), conf.dt)
    assert Counter.n == 30


def test_3d_char_step_is_475_flops():
    conf = nufi.landau_3d(N=4, V=4, T=1.0)
    c = np.random.default_rng(2).standard_normal(conf.stride_t())
    Counter.n = 0
    step3d(conf, c, [Q(1.1), Q(2.3), Q(4.7)], [Q(0.4), Q(-0.9), Q(1.6)], conf.dt)
    assert Counter.n == 475


@pytest.mark.parametrize("seed", [3, 4])
def test_1d_counted_form_matches_the_oracle_field(seed):
    conf = nufi.landau_1d(Nx=16, Nu=32, T=1.0)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(conf.stride_t())
    for x0 in rng.uniform(conf.x_min, conf.x_min + conf.Lx, 20):
        u0 = rng.uniform(-3.0, 3.0)
        x, _, e = step1d(conf, c, Q(x0), Q(u0), conf.dt)
        xw = conf.x_min + math.fmod(val(x) - conf.x_min + 10 * conf.Lx, conf.Lx)
        ref = O.field1d(conf, c, xw, deriv=1)
        assert abs(val(e) - ref) <= 1e-12 * max(1.0, abs(ref)), (x0, val(e), ref)


def test_3d_counted_form_matches_the_oracle_field():
    conf = nufi.landau_3d(N=4, V=4, T=1.0)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(conf.stride_t())
    L = conf.Lx
    for _ in range(10):
        p0 = list(rng.uniform(conf.x_min, conf.x_min + L, 3))
        v0 = list(rng.uniform(-3.0, 3.0, 3))
        pos, _, e = step3d(conf, c, [Q(p) for p in p0], [Q(v) for v in v0], conf.dt)
        pw = [conf.x_min + math.fmod(val(p) - conf.x_min + 10 * L, L) for p in pos]
        for a, d in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            ref = O.field3d(conf, c, pw[0], pw[1], pw[2], d=d)
            assert abs(val(e[a]) - ref) <= 1e-12 * max(1.0, abs(ref)), (a, val(e[a]), ref)


def test_reference_formulation_op_counts():
    """fields.hpp:36-61 / 308-350 as restated by the oracle: the drift, the
    wrap and locate as written, de Boor (1D) or the tensor contraction with one
    de Boor (3D, three times for E), and the kick."""
    n1 = O.count_step_ops(nufi.landau_1d(Nx=16, Nu=32, T=1.0))
    n3 = O.count_step_ops(nufi.landau_3d(N=4, V=4, T=1.0))
    assert (n1, n3) == (42, 848)
    assert n1 > 30 and n3 > 475
