"""Pins for the oracle pieces that smooth-data convergence tests leave open:
the upwind side and the face velocity of the WENO fluxes along each axis, and
the passive scalar's coupling to the RK3 stages.  Each check is shown to fail
when the corresponding mistake is planted in the oracle (or_set_mutation)."""
import math

import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2


def _line_setup(o, axis, kind):
    """Velocity with only component `axis`, varying in sign along that axis;
    scalar varying along that axis only (kind 'step' or 'smooth')."""
    n = (o.Nx, o.Ny, o.Nz)[axis]
    L = (o.Lx, o.Ly, o.Lz)[axis]
    s = np.arange(n) * L / n
    vel = np.sin(s + 0.3) + 0.2                 # changes sign twice on the line, never 0 at a grid point
    if kind == "step":
        th = ((s > 0.9) & (s < 3.7)).astype(float)
    else:
        th = np.exp(np.cos(s))
    shape = [1, 1, 1]
    shape[axis] = n
    u = [np.zeros(o.rshape) for _ in range(3)]
    u[axis] = np.broadcast_to(vel.reshape(shape), o.rshape).copy()
    x = np.broadcast_to(th.reshape(shape), o.rshape).copy()
    return u, x, vel


def _weno_step_overshoot(o, axis, scheme, cfl=0.4):
    """Largest excursion outside [0, 1] after one forward-Euler step of the step function."""
    u, x, vel = _line_setup(o, axis, "step")
    d = (o.Lx / o.Nx, o.Ly / o.Ny, o.Lz / o.Nz)[axis]
    dt = cfl * d / np.abs(vel).max()
    y = x - dt * o.weno_advect(scheme, u, x)
    return max(float(y.max()) - 1.0, -float(y.min()), 0.0)


def _weno_telescope(o, axis, scheme):
    """sum_j out_j / u_j along each line, relative to the face values' scale:
    zero when both cells of every face use the same face value."""
    u, x, vel = _line_setup(o, axis, "smooth")
    out = o.weno_advect(scheme, u, x) / u[axis]
    d = (o.Lx / o.Nx, o.Ly / o.Ny, o.Lz / o.Nz)[axis]
    return float(np.abs(out.sum(axis=axis) * d).max())


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("scheme", [1, 3, 5])
def test_weno_upwind_side_keeps_step_bounded(oracle_mod, axis, scheme):
    o = oracle_mod.Oracle(16, 32, 24, weno_zfix=False)
    assert _weno_step_overshoot(o, axis, scheme) < 1e-12
    # the downwind choice along this axis overshoots by about the CFL number
    with oracle_mod.mutation(1 + axis):
        assert _weno_step_overshoot(o, axis, scheme) > 0.1


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("scheme", [1, 3, 5])
def test_weno_face_values_shared_by_both_cells(oracle_mod, axis, scheme):
    for zfix in (False, True):
        o = oracle_mod.Oracle(16, 32, 24, weno_zfix=zfix)
        assert _weno_telescope(o, axis, scheme) < 1e-12
    o = oracle_mod.Oracle(16, 32, 24)
    # selecting by the cell velocity breaks it where the velocity changes sign
    with oracle_mod.mutation(4):
        assert _weno_telescope(o, axis, scheme) > 1e-8


def _shear_scalar(oracle_mod, N, dt, T, Re=1.0):
    """theta advected by the exact decaying shear u = (e^{-t/Re} sin y, 0, 0)
    with RK3 + WENO5 to time T; returns (theta, max error against the closed
    form theta0(x - sin y Re (1 - e^{-t/Re})))."""
    Nx, Ny, Nz = N, 8, 4
    o = oracle_mod.Oracle(Nx, Ny, Nz, Re=Re, dt=dt)
    x = np.arange(Nx) * 2 * np.pi / Nx
    y = np.arange(Ny) * 2 * np.pi / Ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    sy = np.broadcast_to(np.sin(Y)[:, :, None], (Nx, Ny, Nz))
    u = [o.fft_r2c(sy.copy()), np.zeros(o.cshape, np.complex128), np.zeros(o.cshape, np.complex128)]
    th0 = lambda s: np.exp(np.sin(s))                           # noqa: E731
    th = np.broadcast_to(th0(X)[:, :, None], (Nx, Ny, Nz)).copy()
    zero = [np.zeros(o.cshape, np.complex128)] * 3
    steps = int(round(T / dt))
    u, th = o.rk3_scalar(u, zero, th, scheme=5, steps=steps)
    # the velocity stays the decaying mode: amplification 1 + z + z^2/2 + z^3/6 per step, z = -dt/Re
    z = -dt / Re
    amp = (1 + z + z * z / 2 + z ** 3 / 6) ** steps
    assert np.abs(o.fft_c2r(u[0]) - amp * sy).max() < 1e-13
    shift = np.sin(Y) * Re * (1.0 - math.exp(-T / Re))
    exact = np.broadcast_to(th0(X - shift)[:, :, None], (Nx, Ny, Nz))
    return th, float(np.abs(th - exact).max())


def _time_order(oracle_mod, N, T, dts):
    """Temporal order from three step sizes halving each time (the spatial
    error is the same in all three runs and cancels in the differences)."""
    th = [_shear_scalar(oracle_mod, N, dt, T)[0] for dt in dts]
    return math.log2(np.abs(th[0] - th[1]).max() / np.abs(th[1] - th[2]).max())


def test_passive_scalar_against_decaying_shear(oracle_mod):
    """The scalar stages of RK3_SSP with WENO5 (or_rk3_ssp_scalar) against the
    closed form, and third order in time.  Each stage must use its own stage
    velocity and the SSP weights: advecting every stage with u^n's velocity
    (mutation 5) drops to first order; swapping the second stage's weights
    (mutation 6) leaves a first-order error."""
    N, T, dts = 64, 0.8, (0.04, 0.02, 0.01)
    _, err = _shear_scalar(oracle_mod, N, dts[-1], T)
    assert err < 2e-4                                   # WENO5 at 64 points, smooth data
    assert _time_order(oracle_mod, N, T, dts) > 2.7
    with oracle_mod.mutation(5):
        assert _time_order(oracle_mod, N, T, dts) < 1.5
    with oracle_mod.mutation(6):
        assert _time_order(oracle_mod, N, T, dts) < 1.5
