"""Pins the oracle's forward-mode tangent (dual numbers through the trace,
componentwise Poisson and fit, product-rule energy) against central finite
differences of the oracle's own time loop, and its J against the plain loop."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import adjoint as A
from numericalflowiteration_amd import config as C
from tests import oracle_lib as O


def J_of(conf):
    h = O.new_history(conf, conf.Nt + 1)
    return conf.dt * float(np.sum(O.run(conf, 0, conf.Nt + 1, h)))


def test_tangent_matches_fd_multimode():
    conf = nufi.multimode_1d(nufi.adjoint_modes(), Nx=32, Nu=128, T=1.0)
    J, dJ = O.tangent_run(conf)
    assert abs(J - J_of(conf)) <= 1e-15 * abs(J)          # same loop, summation order aside
    a0, h = nufi.adjoint_modes(), 1e-6
    for m in [0, 3, 7]:
        ap, am = list(a0), list(a0)
        ap[m] += h
        am[m] -= h
        fd = (J_of(A._with_params(conf, ap)) - J_of(A._with_params(conf, am))) / (2 * h)
        assert abs(dJ[m] - fd) <= 1e-7 * abs(fd), (m, dJ[m], fd)


@pytest.mark.parametrize("kind", [C.F0_LANDAU_1D, C.F0_TWOSTREAM_1D])
def test_tangent_matches_fd_alpha(kind):
    conf = C.config_1d(24, 96, 1 / 16, 1.0, 0.0, 4 * np.pi, -8.0, 8.0, kind, alpha=0.05, k=0.5)
    J, dJ = O.tangent_run(conf)
    h = 1e-6
    fd = (J_of(A._with_params(conf, [0.05 + h])) - J_of(A._with_params(conf, [0.05 - h]))) / (2 * h)
    assert abs(dJ[0] - fd) <= 1e-7 * abs(fd)
