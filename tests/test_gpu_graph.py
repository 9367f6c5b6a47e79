"""HIP-graph replay of K RK3 steps (ns3d_rk3_ssp_graph) against eager RK3_SSP.

The graph is captured once and replayed several times, with unrelated eager
torch work and host synchronisations between the replays; every replay must
reproduce the eager kernels' result bit for bit (same kernels, same order).
A re-capture (changed dt) and the invalidation after a scratch reallocation
(B(u,v) doubles the context's scratch) are checked the same way.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,precision", [((32, 32, 32), "fp64"), ((48, 32, 64), "fp64"), ((64, 64, 64), "fp32"),
                                             ((32, 48, 32), "fp32s")])
def test_graph_replays_match_eager(shape, precision):
    import torch
    from ns3dperiodic_amd import NS3D, fields
    print(f"\n[case] graph {shape} {precision}", flush=True)
    Nx, Ny, Nz = shape
    s = torch.cuda.Stream()
    g = NS3D(Nx, Ny, Nz, Re=200.0, dt=2e-3, precision=precision, stream=s)
    e = NS3D(Nx, Ny, Nz, Re=200.0, dt=2e-3, precision=precision)
    u0 = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(Nx, Ny, Nz, seed=3)]
    f = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(Nx, Ny, Nz, seed=4, energy=0.01)]
    ug = [c.clone() for c in u0]
    ue = [c.clone() for c in u0]
    torch.cuda.synchronize()
    K = 3
    for rep in range(4):
        with torch.cuda.stream(s):
            g.RK3_SSP_graph(ug, f, steps=K)
        e.RK3_SSP(ue, f, steps=K)
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for a, b in zip(ug, ue):
            assert torch.equal(a, b), f"replay {rep} differs"
    # new dt -> re-capture
    g.set_params(1e-3, 200.0)
    e.set_params(1e-3, 200.0)
    with torch.cuda.stream(s):
        g.RK3_SSP_graph(ug, f, steps=K)
    e.RK3_SSP(ue, f, steps=K)
    torch.cuda.synchronize()
    for a, b in zip(ug, ue):
        assert torch.equal(a, b), "re-captured graph differs"
    # B(u,v) reallocates the scratch: the cached graph must not replay stale pointers
    with torch.cuda.stream(s):
        g.calculate_convolution_2p3_UV(ug, ue)
        g.RK3_SSP_graph(ug, f, steps=K)
    e.RK3_SSP(ue, f, steps=K)
    torch.cuda.synchronize()
    for a, b in zip(ug, ue):
        assert torch.equal(a, b), "graph after scratch reallocation differs"
    assert all(bool(torch.isfinite(c).all()) for c in ug)
    assert np.isfinite(g.get_kinetic_energy([g.iFFTN(c) for c in ug]))
