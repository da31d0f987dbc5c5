"""Device adjoint (nufi_adjoint_grad) of J = sum_n dt * energy_n with respect to
the f0 parameters: J against the oracle, the gradient against central finite
differences (h = 1e-6) of the oracle (the reference arithmetic) and of the
device forward model, and through the torch.autograd wrapper.
Tolerance: <= 1e-5 relative (the bar of the adjoint case) against the oracle."""
import numpy as np
import pytest

import numericalflowiteration_amd as nufi
from numericalflowiteration_amd import adjoint as A
from numericalflowiteration_amd import config as C
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def J_oracle(conf):
    h = O.new_history(conf, conf.Nt + 1)
    e = O.run(conf, 0, conf.Nt + 1, h)
    return conf.dt * float(np.sum(e))


def J_device(conf):
    with nufi.GpuScheduler(conf) as s:
        return conf.dt * float(np.sum(s.run()))


def small_multimode():
    return nufi.multimode_1d(nufi.adjoint_modes(), Nx=32, Nu=128, T=2.0)      # 33 steps, 8 parameters


def test_adjoint_multimode_against_fd():
    conf = small_multimode()
    J, g = A.energy_and_grad(conf, nufi.adjoint_modes())
    Jo = J_oracle(conf)
    print(f"[adjoint] J device {J:.15e} oracle {Jo:.15e}", flush=True)
    assert abs(J - Jo) < 1e-11 * abs(Jo)
    h = 1e-6
    a0 = nufi.adjoint_modes()
    for m in range(8):
        ap, am = list(a0), list(a0)
        ap[m] += h
        am[m] -= h
        fd_o = (J_oracle(A._with_params(conf, ap)) - J_oracle(A._with_params(conf, am))) / (2 * h)
        fd_d = (J_device(A._with_params(conf, ap)) - J_device(A._with_params(conf, am))) / (2 * h)
        print(f"  a_{m+1}: adjoint {g[m]:.10e} fd(oracle) {fd_o:.10e} fd(device) {fd_d:.10e}", flush=True)
        scale = max(abs(fd_o), 1e-3 * np.max(np.abs(g)))
        assert abs(g[m] - fd_o) <= 1e-5 * scale
        assert abs(g[m] - fd_d) <= 1e-6 * scale


def test_adjoint_against_forward_tangent():
    conf = small_multimode()
    J, g = A.energy_and_grad(conf, nufi.adjoint_modes())
    Jt, dJ = O.tangent_run(conf)
    err = np.max(np.abs(g - dJ[:8]) / np.max(np.abs(dJ[:8])))
    print(f"[adjoint vs tangent] max rel err {err:.3e}", flush=True)
    assert abs(J - Jt) < 1e-11 * abs(Jt)
    assert err < 1e-8


@pytest.mark.parametrize("kind", [C.F0_LANDAU_1D, C.F0_TWOSTREAM_1D])
def test_adjoint_alpha(kind):
    conf = C.config_1d(32, 128, 1 / 16, 1.5, 0.0, 4 * np.pi, -8.0, 8.0, kind, alpha=0.05, k=0.5)
    J, g = A.energy_and_grad(conf, [0.05])
    h = 1e-6
    fd = (J_oracle(A._with_params(conf, [0.05 + h])) - J_oracle(A._with_params(conf, [0.05 - h]))) / (2 * h)
    print(f"[adjoint alpha] kind={kind} adjoint {g[0]:.10e} fd {fd:.10e}", flush=True)
    assert abs(g[0] - fd) <= 1e-5 * abs(fd)


def test_torch_autograd_wrapper():
    import torch
    conf = small_multimode()
    a = torch.tensor(nufi.adjoint_modes(), dtype=torch.float64, requires_grad=True)
    J = A.time_integrated_energy(a, conf)
    (2.0 * J).backward()
    _, g = A.energy_and_grad(conf, nufi.adjoint_modes())
    assert np.allclose(a.grad.numpy(), 2.0 * g, rtol=1e-12, atol=0)


def test_adjoint_of_a_strict_context():
    """The forward pass runs the fast kernel whatever variant the context is
    set to (with a launch shape that kernel takes), and the variant and shape
    are restored after it."""
    from numericalflowiteration_amd import _capi
    conf = small_multimode()
    J0, g0 = A.energy_and_grad(conf, nufi.adjoint_modes())
    with nufi.GpuScheduler(A._with_params(conf, nufi.adjoint_modes())) as s:
        s.set_trace_variant(_capi.TRACE_STRICT)
        shape = s.launch_shape()
        J, g = s.adjoint_grad()
        assert s.launch_shape() == shape
        e = s.run()                                     # the strict kernel still runs
    print(f"[adjoint strict context] J {J:.15e} vs {J0:.15e}; grad max rel diff "
          f"{np.max(np.abs(g - g0) / np.abs(g0)):.3e}", flush=True)
    # the forward run is deterministic; the slice adjoints are summed with
    # atomics, so the gradient agrees to rounding, not bit for bit
    assert J == J0 and np.allclose(g, g0, rtol=1e-12, atol=0)
    assert abs(conf.dt * float(np.sum(e)) - J0) <= 1e-12 * abs(J0)


def test_fp32c_adjoint_against_the_oracle_tangent_and_fd():
    """The FP32C adjoint differentiates the FP32-compensated forward model.  Its
    J is the FP32C run's (against the oracle's FP64 J: <= 1e-6), its gradient is
    checked against the oracle's forward tangent, which central differences of
    the oracle pin (test above), at <= 1e-5 of the gradient's scale, and against
    the FP64 device adjoint; and against central differences (h = 1e-6) of the
    FP32C device model itself, whose rounding noise divided by h dominates that
    comparison (measured 6.9e-6; the FP32C forward is deterministic, so the
    figure is too) — held to the adjoint bar, 1e-5."""
    conf = small_multimode()
    a0 = nufi.adjoint_modes()
    J, g = A.energy_and_grad(conf, a0, precision=nufi.FP32C)
    J64, g64 = A.energy_and_grad(conf, a0)
    Jt, dJ = O.tangent_run(conf)
    scale = np.max(np.abs(dJ[:8]))
    err_t = np.max(np.abs(g - dJ[:8])) / scale
    err_64 = np.max(np.abs(g - g64)) / scale
    h = 1e-6
    fd = []
    for m in range(8):
        ap, am = list(a0), list(a0)
        ap[m] += h
        am[m] -= h
        fd.append((A.energy_and_grad(conf, ap, precision=nufi.FP32C)[0] -
                   A.energy_and_grad(conf, am, precision=nufi.FP32C)[0]) / (2 * h))
    err_fd = np.max(np.abs(g - np.array(fd))) / scale
    print(f"[adjoint FP32C] J {J:.12e} vs oracle {Jt:.12e}; grad vs tangent {err_t:.3e}, vs FP64 adjoint "
          f"{err_64:.3e}, vs FD of the FP32C model {err_fd:.3e}", flush=True)
    assert abs(J - Jt) <= 1e-6 * abs(Jt)
    assert err_t <= 1e-5 and err_64 <= 1e-5
    assert err_fd <= 1e-5


def test_fp32c_adjoint_non_power_of_two_grid():
    """Nx = 24 (positions wrapped every step) against the oracle's tangent."""
    conf = nufi.multimode_1d(nufi.adjoint_modes(), Nx=24, Nu=128, T=1.5)
    J, g = A.energy_and_grad(conf, nufi.adjoint_modes(), precision=nufi.FP32C)
    Jt, dJ = O.tangent_run(conf)
    err = np.max(np.abs(g - dJ[:8])) / np.max(np.abs(dJ[:8]))
    print(f"[adjoint FP32C Nx=24] grad vs tangent {err:.3e}", flush=True)
    assert abs(J - Jt) <= 1e-6 * abs(Jt) and err <= 1e-5


@pytest.mark.parametrize("prec,tol", [(nufi.FP64, 1e-8), (nufi.FP32C, 1e-5)])
def test_adjoint_full_grid_against_the_oracle_tangent(prec, tol):
    """The adjoint case's own grid (Nx = 128, Nu = 512, dt = 1/16) over 81 steps
    (the oracle's tangent to T = 20 takes minutes): FP64 to the tangent bar of
    the small case, FP32C to the adjoint bar."""
    conf = nufi.multimode_1d(nufi.adjoint_modes(), T=5.0)
    J, g = A.energy_and_grad(conf, nufi.adjoint_modes(), precision=prec)
    Jt, dJ = O.tangent_run(conf)
    err = np.max(np.abs(g - dJ[:8])) / np.max(np.abs(dJ[:8]))
    print(f"[adjoint full grid] precision={prec} J rel err {abs(J - Jt) / abs(Jt):.3e}, grad vs tangent {err:.3e}",
          flush=True)
    assert abs(J - Jt) <= (1e-11 if prec == nufi.FP64 else 1e-6) * abs(Jt)
    assert err <= tol


def test_adjoint_rejects_unsupported():
    with nufi.GpuScheduler(nufi.landau_3d(N=8, V=8, T=1.0)) as s:
        with pytest.raises(nufi.NufiError):
            s.adjoint_grad()
