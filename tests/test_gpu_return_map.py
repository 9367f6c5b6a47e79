"""Return map on the GPU: the one-pass spectral probe against an inverse
transform, and a plane crossing located by bisection, replayed step for step
with the oracle."""
import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 48), (48, 64, 32)], ids=lambda s: "x".join(map(str, s)))
def test_probe_matches_inverse_transform(shape):
    import torch
    from ns3dperiodic_amd import NS3D
    print(f"\n[case] probe {shape}", flush=True)
    g = NS3D(*shape)
    u = fields.random_solenoidal(*shape, seed=61)
    # a non-Hermitian spectrum exercises the documented C2R convention too
    rng = np.random.default_rng(2)
    u[1] = u[1] + 0.3 * (rng.standard_normal(u[1].shape) + 1j * rng.standard_normal(u[1].shape))
    phys = [np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in u]
    ud = [torch.from_numpy(c).cuda() for c in u]
    for pt in [(0, 0, 0), (1, 2, 3), (shape[0] - 1, shape[1] // 2, shape[2] - 1), (5, 7, shape[2] // 2)]:
        got = g.probe(ud, *pt)
        ref = np.array([p[pt] for p in phys])
        assert np.allclose(got, ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max())), (pt, got, ref)
    with pytest.raises(Exception):
        g.probe(ud, shape[0], 0, 0)


def test_crossing_replayed_by_oracle():
    import torch
    from ns3dperiodic_amd import NS3D, return_map
    from oracle.oracle import Oracle
    print("\n[case] return-map crossing, 16^3", flush=True)
    N, Re, dt, pt = 16, 50.0, 2e-3, (3, 5, 7)
    u0 = fields.random_solenoidal(N, seed=62)
    f = fields.kolmogorov_forcing_sin(N, n=4)
    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).cuda() for c in u0]
    fd = [torch.from_numpy(c).cuda() for c in f]
    sysm = return_map.NSSystem(g, u, fd, point=pt)
    p0, v = sysm.probe(), sysm.probe_rhs()
    normal = v / np.linalg.norm(v)
    x0 = p0 + 6.5 * dt * v                      # crossed in the 7th step, moving along v
    out = return_map.find_crossing(sysm, x0, normal, dt, rho=1e9, max_steps=50)
    print(f"[return_map] steps {out['steps']}, bisections {out['bisections']}, |s| {out['error']:.1e}", flush=True)
    assert out["found"] and out["steps"] >= 5 and out["error"] <= 1e-12
    # replay the exact step sequence with the oracle
    uo = [c.copy() for c in u0]
    for h in out["step_sizes"]:
        uo = Oracle(N, Re=Re, dt=h).rk3(uo, f)
    ug = [c.cpu().numpy() for c in u]
    assert rel_l2(ug, uo) < 1e-11
    po = np.array([np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))[pt] for c in uo])
    assert np.allclose(out["point"], po, rtol=0, atol=1e-11)
    assert abs(return_map.plane_value(po, x0, normal)) < 1e-10
