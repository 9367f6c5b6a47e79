"""FP32 paths against the FP64 path (bar: <= 1e-5 relative L2 after 1000 steps).

fp32: FP32 transforms and products, FP64 state.  fp32s: FP32 state inside the RK3 loop too.
fp32i: FP32 transforms and FP32 stage increments over the FP64 state.

Workload: 256^3 decaying turbulence, nu = 2e-3, the seeded k^4 exp(-2(k/4)^2)
spectrum, 1000 RK3 steps in each precision from identical initial bits.  Measured on
MI355X: 1.3e-6 (energy agrees to ~4e-8, dissipation to ~3e-7).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rel(a, b):
    import torch
    num = sum(float(torch.sum(torch.abs(x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(torch.sum(torch.abs(y) ** 2)) for y in b)
    return math.sqrt(num / den)


# fp32s rounds the state to FP32 every step (eps*|u| per step instead of
# eps*dt*|L(u)|): measured 1.3e-5 after 1000 steps, above the 1e-5 bar that the
# FP64-state fp32 path meets.  It is held to the same bar and expected to miss it.
FP32S_MISSES = pytest.mark.xfail(strict=True, reason="fp32s: FP32 state rounding, measured 1.31e-5 > 1e-5 "
                                                      "after 1000 steps at 256^3")


@pytest.mark.parametrize("precision,tol", [("fp32", TOL), pytest.param("fp32s", TOL, marks=FP32S_MISSES),
                                           ("fp32i", TOL)])
def test_fp32_bulk_error_after_1000_steps_256(precision, tol):
    import torch
    from ns3dperiodic_amd import NS3D, fields
    print(f"\n[case] {precision} vs fp64, 256^3, 1000 steps", flush=True)
    N, Re, dt = 256, 500.0, 1.9e-3
    u0 = fields.random_solenoidal(N, seed=1)
    g64 = NS3D(N, Re=Re, dt=dt, precision="fp64")
    g32 = NS3D(N, Re=Re, dt=dt, precision=precision)
    a = [torch.from_numpy(c).cuda() for c in u0]
    b = [c.clone() for c in a]
    for _ in range(10):
        g64.RK3_SSP(a, None, steps=100)
        g32.RK3_SSP(b, None, steps=100)
    err = _rel(b, a)
    e64 = g64.get_kinetic_energy([g64.iFFTN(c) for c in a])
    e32 = g64.get_kinetic_energy([g64.iFFTN(c) for c in b])
    print(f"[{precision}] rel L2 after 1000 steps = {err:.3e}; energy rel diff = {abs(e32 - e64) / e64:.3e}", flush=True)
    assert bool(torch.isfinite(a[0]).all()) and bool(torch.isfinite(b[0]).all())
    assert err < tol
    assert abs(e32 - e64) < 1e-6 * e64
