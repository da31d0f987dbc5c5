"""The FP32 cubic B-spline bases of the 3D FP32C trace (csrc/trace3d.hip,
cubic_bases_x6(float)), emulated operation by operation in numpy float32.

The kernel's gradient contraction sum_i c_i D_i must carry no term in the
level of the coefficients, so the four FP32 slope values have to sum to
exactly zero; plain FP32 rounding of the four slopes leaves a residual of
about 2^-24 that multiplies the local level of phi (it gave 2e-6 relative
energy error on 3D Landau).  Checked here: the exact zero sum for t over
[-1e-2, 1 + 1e-2] (the lo part of a position pair, unnormalised for up to 16
slices, can push t out of [0, 1) by at most 128 ulp of the frame offset:
0.004 cells on 3D Landau), the slopes and values against the closed forms in FP64
(nufi/splines.hpp:39-110 at order 4, times 2 and 6), and that the naive FP32
slopes do not sum to zero (so the test can fail)."""
from fractions import Fraction

import numpy as np

f32 = np.float32


def fmaf(a, b, c):
    # a*b of two floats is exact in float64; the float64 sum then rounds once
    # more than the hardware FMA, which does not change the grid property
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def bases_fp32(t):
    t = f32(t)
    a = f32(f32(1) - t)
    a2 = f32(fmaf(a, a, f32(1)) - f32(1))
    t2 = f32(fmaf(t, t, f32(1)) - f32(1))
    B0 = f32(a2 * a)
    B3 = f32(t2 * t)
    B1 = fmaf(t2, fmaf(f32(3), t, f32(-6)), f32(4))
    B2 = fmaf(a2, fmaf(f32(3), a, f32(-6)), f32(4))          # B_1(1 - t)
    D1 = f32(fmaf(t, fmaf(f32(3), t, f32(-4)), f32(-1)) + f32(1))
    D2 = f32(f32(a2 - t2) - D1)
    return (B0, B1, B2, B3), (f32(-a2), D1, D2, t2)


def closed_form(t):
    a = 1.0 - t
    B = (a**3, 3 * t**3 - 6 * t * t + 4, -3 * t**3 + 3 * t * t + 3 * t + 1, t**3)
    D = (-a * a, 3 * t * t - 4 * t, -3 * t * t + 2 * t + 1, t * t)
    return B, D


def sample_t():
    rng = np.random.default_rng(7)
    t = rng.uniform(-1e-2, 1 + 1e-2, 20000).astype(np.float32)
    edges = np.array([0.0, 2.0**-24, -2.0**-24, 1e-8, -1e-8, 1 / 3, 0.5, 2 / 3, 1 - 2.0**-24, 1.0,
                      1 + 2.0**-23, -1e-3, 1 + 1e-3, -1e-2, 1 + 1e-2], dtype=np.float32)
    return np.concatenate([t, edges])


def test_slopes_sum_to_exactly_zero():
    for t in sample_t():
        _, D = bases_fp32(t)
        assert sum(Fraction(float(d)) for d in D) == 0, float(t)


def test_bases_against_closed_forms():
    worst_d = worst_b = 0.0
    for t in sample_t():
        B, D = bases_fp32(t)
        Bx, Dx = closed_form(float(t))
        worst_d = max(worst_d, max(abs(float(d) - e) for d, e in zip(D, Dx)))
        worst_b = max(worst_b, max(abs(float(b) - e) for b, e in zip(B, Bx)))
    # slopes: three roundings onto the 2^-23 grid (2^-24 each; for t just below
    # 0, (1-t)^2 + 1 passes 2 and its grid is 2^-22) plus the inner FP32
    # product; values: a few FP32 roundings of numbers up to 6
    assert worst_d < 6 * 2.0**-24
    assert worst_b < 16 * 2.0**-24


def test_plain_fp32_slopes_do_not_sum_to_zero():
    """The residual the quantised slopes remove: rounding each slope to FP32."""
    residuals = 0
    for t in sample_t()[:2000]:
        _, Dx = closed_form(float(t))
        if sum(Fraction(float(f32(d))) for d in Dx) != 0:
            residuals += 1
    assert residuals > 100
