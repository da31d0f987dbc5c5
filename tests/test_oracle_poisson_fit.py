"""Pins the oracle's periodic Poisson solve (poisson.cpp:66-89, 328-362) and the
spline fit (fields.hpp:63-142, 352-490) by closed forms, the analytic case of
bin/test_poisson.cpp:36-50, numpy FFTs and dense linear algebra."""
import math

import numpy as np
import pytest

from numericalflowiteration_amd import config as C
from tests import oracle_lib as O


def conf1(N, L):
    return C.config_1d(N, 16, 0.1, 1.0, 0.0, L, -1.0, 1.0, C.F0_LANDAU_1D)


def conf3(N, L):
    return C.config_3d(N, (4, 4, 4), 0.1, 1.0, (0.0, L), (-1.0, 1.0), C.F0_LANDAU_3D)


@pytest.mark.parametrize("N,m", [(64, 1), (64, 5), (256, 17), (48, 3)])
def test_single_mode_closed_form(N, m):
    L = 4 * math.pi
    c = conf1(N, L)
    x = np.arange(N) * c.dx
    kap = 2 * math.pi * m / L
    alpha = 0.37
    for fn in (np.cos, np.sin):
        rho = alpha * fn(kap * x)
        phi, e = O.poisson(c, rho)
        assert np.max(np.abs(phi - rho / kap ** 2)) < 1e-14 * alpha / kap ** 2 * N
        assert abs(e - L * alpha ** 2 / (4 * kap ** 2)) < 1e-13 * e


def test_nyquist_mode():
    N, L = 32, 2.0
    c = conf1(N, L)
    rho = np.array([(-1.0) ** i for i in range(N)])
    kap = math.pi * N / L
    phi, e = O.poisson(c, rho)
    assert np.allclose(phi, rho / kap ** 2, rtol=0, atol=1e-15)
    assert abs(e - L / (2 * kap ** 2)) < 1e-13 * e


def test_mode_zero_dropped():
    c = conf1(32, 3.0)
    phi, e = O.poisson(c, np.full(32, 2.5))
    assert np.max(np.abs(phi)) < 1e-15 and e < 1e-30


def test_random_vs_numpy_fft_1d():
    rng = np.random.default_rng(1)
    N, L = 128, 10 * math.pi
    c = conf1(N, L)
    rho = rng.standard_normal(N)
    phi, e = O.poisson(c, rho)
    k = 2 * math.pi * np.fft.fftfreq(N, d=L / N)
    F = np.fft.fft(rho)
    k2 = k ** 2
    k2[0] = 1.0
    ph = F / k2
    ph[0] = 0
    assert np.allclose(phi, np.fft.ifft(ph).real, rtol=0, atol=1e-13 * np.abs(phi).max())
    e_ref = L / 2 / N ** 2 * np.sum(np.abs(F[1:]) ** 2 / k2[1:])
    assert abs(e - e_ref) < 1e-12 * e_ref


def test_analytic_3d_case_of_test_poisson():
    """rho = cos(2z) + sin(8y) + cos(42x) on [0,2pi)^3 (bin/test_poisson.cpp:36-50)."""
    c = C.config_3d((128, 64, 32), (4, 4, 4), 0.1, 1.0, (0.0, 2 * math.pi), (-1.0, 1.0), C.F0_LANDAU_3D)
    x = np.arange(128) * c.dx
    y = np.arange(64) * c.dy
    z = np.arange(32) * c.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.cos(2 * Z) + np.sin(8 * Y) + np.cos(42 * X)
    phi, e = O.poisson(c, rho.ravel())
    ref = np.cos(2 * Z) / 4 + np.sin(8 * Y) / 64 + np.cos(42 * X) / 42 ** 2
    assert np.max(np.abs(phi - ref.ravel())) < 1e-14
    # energy: each mode contributes L^3 / (4 kappa^2) for unit amplitude
    L3 = (2 * math.pi) ** 3
    e_ref = L3 / 4 * (1 / 4 + 1 / 64 + 1 / 42 ** 2)
    assert abs(e - e_ref) < 1e-12 * e_ref


def test_random_vs_numpy_fft_3d():
    rng = np.random.default_rng(3)
    c = conf3((8, 16, 4), 4 * math.pi)
    rho = rng.standard_normal(8 * 16 * 4)
    phi, e = O.poisson(c, rho)
    L = 4 * math.pi
    kz = 2 * math.pi * np.fft.fftfreq(4, d=L / 4)
    ky = 2 * math.pi * np.fft.fftfreq(16, d=L / 16)
    kx = 2 * math.pi * np.fft.fftfreq(8, d=L / 8)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    F = np.fft.fftn(rho.reshape(4, 16, 8))
    k2s = k2.copy()
    k2s[0, 0, 0] = 1
    ph = F / k2s
    ph[0, 0, 0] = 0
    assert np.allclose(phi, np.fft.ifftn(ph).real.ravel(), rtol=0, atol=1e-13 * np.abs(phi).max())
    e_ref = L ** 3 / 2 / rho.size ** 2 * np.sum((np.abs(F) ** 2 / k2s).ravel()[1:])
    assert abs(e - e_ref) < 1e-12 * e_ref


def collocation_matrix(N, order):
    B = O.basis(order, 0, 0.0)
    A = np.zeros((N, N))
    for i in range(N):
        for j in range(order):
            A[i, (i + j) % N] += B[j]
    return A


@pytest.mark.parametrize("N", [16, 64, 100])
def test_fit_1d_dense_and_residual(N):
    rng = np.random.default_rng(N)
    c = conf1(N, 2.0)
    v = rng.standard_normal(N)
    s = O.fit(c, v)
    assert s.size == N + 3
    A = collocation_matrix(N, 4)
    assert np.max(np.abs(A @ s[:N] - v)) < 1e-14 * np.abs(v).max() * 4
    assert np.allclose(s[:N], np.linalg.solve(A, v), rtol=0, atol=1e-13)
    assert np.array_equal(s[N:], s[:3])          # periodic padding (fields.hpp:140-141)


def test_fit_odd_order_singular_is_min_norm():
    N = 16
    c = conf1(N, 2.0)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(N)
    s = O.fit(c, v, order=3)
    A = collocation_matrix(N, 3)
    ref = np.linalg.lstsq(A, v, rcond=None)[0]
    assert np.allclose(s[:N], ref, rtol=0, atol=1e-12)


def test_fit_3d_tensor_residual():
    rng = np.random.default_rng(9)
    c = conf3((8, 4, 6), 3.0)
    v = rng.standard_normal(8 * 4 * 6)
    s = O.fit(c, v)
    ex, ey, ez = 11, 7, 9
    S = s.reshape(ez, ey, ex)
    B = O.basis(4, 0, 0.0)
    out = np.zeros((6, 4, 8))
    for k in range(6):
        for j in range(4):
            for i in range(8):
                out[k, j, i] = sum(S[k + c_, j + b_, i + a_] * B[a_] * B[b_] * B[c_]
                                   for a_ in range(4) for b_ in range(4) for c_ in range(4))
    assert np.max(np.abs(out.ravel() - v)) < 1e-13
    # periodic padding in every axis
    assert np.array_equal(S[6:, :, :], S[:3, :, :])
    assert np.array_equal(S[:, 4:, :], S[:, :3, :])
    assert np.array_equal(S[:, :, 8:], S[:, :, :3])
