"""Nodal phi / E / rho of a history slice (numericalflowiteration_amd.diagnostics)
against the oracle: the slice is made by the oracle's Poisson solve and fit
from a known density, so the density and the nodal potential must come back,
and the nodal field must equal the oracle's spline evaluator at the nodes."""
import numpy as np

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import diagnostics as D
from tests import oracle_lib as O


def slice_from_density(conf, rho):
    phi, energy = O.poisson(conf, rho)
    return phi, O.fit(conf, phi)


def test_1d_potential_density_field():
    conf = nufi.twostream_1d(Nx=32, Nu=64, T=1.0)
    rng = np.random.default_rng(11)
    rho = rng.standard_normal(conf.Nx)
    phi, sl = slice_from_density(conf, rho)
    assert np.max(np.abs(D.node_potential(conf, sl) - phi)) <= 1e-13 * np.max(np.abs(phi))
    mf = rho - rho.mean()
    assert np.linalg.norm(D.node_density(conf, sl) - mf) <= 1e-12 * np.linalg.norm(mf)
    E = D.node_field(conf, sl)
    ref = np.array([-O.field1d(conf, sl, conf.x_min + i * conf.dx, deriv=1) for i in range(conf.Nx)])
    assert np.max(np.abs(E - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_3d_potential_density_field():
    conf = nufi.config.config_3d((8, 4, 6), (4, 4, 4), 0.125, 1.0, (0.0, 4 * np.pi), (-6.0, 6.0),
                                 nufi.config.F0_LANDAU_3D, alpha=0.01, k=0.5)
    conf.y_max = 2 * np.pi                      # a non-cubic box: the axes must not be mixed up
    conf.derive()
    rng = np.random.default_rng(12)
    rho = rng.standard_normal(conf.n_space())
    phi, sl = slice_from_density(conf, rho)
    shape = (conf.Nz, conf.Ny, conf.Nx)
    phi = phi.reshape(shape)
    assert np.max(np.abs(D.node_potential(conf, sl) - phi)) <= 1e-13 * np.max(np.abs(phi))
    mf = (rho - rho.mean()).reshape(shape)
    assert np.linalg.norm(D.node_density(conf, sl) - mf) <= 1e-12 * np.linalg.norm(mf)
    E = D.node_field(conf, sl)
    assert E.shape == (3,) + shape
    for c, d in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        for l in range(0, conf.n_space(), 7):
            i, j, k = l % conf.Nx, (l // conf.Nx) % conf.Ny, l // (conf.Nx * conf.Ny)
            ref = -O.field3d(conf, sl, conf.x_min + i * conf.dx, conf.y_min + j * conf.dy,
                             conf.z_min + k * conf.dz, d=d)
            assert abs(E[c, k, j, i] - ref) <= 1e-13 * np.max(np.abs(E[c])), (c, l)


def test_run_errors_finds_a_planted_difference():
    conf = nufi.landau_1d(Nx=16, Nu=32, T=0.5)
    hist = O.new_history(conf, conf.Nt + 1)
    e = O.run(conf, 0, conf.Nt + 1, hist)
    st = conf.stride_t()
    h = hist[: (conf.Nt + 1) * st]
    r = D.run_errors(conf, h, e, h, e)
    assert r["rho_max_rel_err"] == 0 and r["E_max_rel_err"] == 0 and r["energy_max_rel_err"] == 0
    h2 = h.copy()
    h2[5 * st:6 * st] *= 1 + 1e-7                 # step 5's field 1e-7 too strong
    e2 = e.copy()
    e2[5] *= 1 + 2e-7
    r = D.run_errors(conf, h2, e2, h, e)
    assert r["rho_err_at_step"] == 5 and abs(r["rho_max_rel_err"] - 1e-7) < 1e-9
    assert r["E_err_at_step"] == 5 and abs(r["E_max_rel_err"] - 1e-7) < 1e-9
    assert r["energy_err_at_step"] == 5 and abs(r["energy_max_rel_err"] - 2e-7) < 1e-12
