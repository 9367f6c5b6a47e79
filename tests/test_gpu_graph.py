"""HIP-graph replay of K RK3 steps (ns3d_rk3_ssp_graph) against eager RK3_SSP.

The graph is captured once and replayed several times, with unrelated eager
torch work and host synchronisations between the replays; every replay must
reproduce the eager kernels' result bit for bit (same kernels, same order).
A re-capture (changed dt) and the invalidation after a scratch reallocation
(B(u,v) doubles the context's scratch) are checked the same way, and so is
StepGraph, which captures the steps of several contexts of uneven sizes into
one graph of concurrent branches.
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


def test_step_graph_of_uneven_sizes():
    """One captured graph advancing three contexts of different sizes and
    precisions on concurrent branches (StepGraph, two branches for three
    items, so one branch runs two contexts in order): four replays with eager
    work and host syncs between them, each state bitwise equal to stepping
    it eagerly."""
    import torch
    from ns3dperiodic_amd import NS3D, fields
    from ns3dperiodic_amd import StepGraph
    print("\n[case] step graph of 32^3 fp64 + 48x64x32 fp32 + 16^3 fp64", flush=True)
    specs = [((32, 32, 32), "fp64", True), ((48, 64, 32), "fp32", False), ((16, 16, 16), "fp64", True)]
    items, refs = [], []
    for i, (sh, prec, forced) in enumerate(specs):
        g = NS3D(*sh, Re=200.0, dt=2e-3, precision=prec)
        e = NS3D(*sh, Re=200.0, dt=2e-3, precision=prec)
        u = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(*sh, seed=30 + i)]
        f = [torch.from_numpy(c).cuda() for c in fields.kolmogorov_forcing_sin(*sh, n=2, amp=0.1)] if forced else None
        items.append((g, u, f))
        refs.append((e, [c.clone() for c in u], f))
    K = 3
    sg = StepGraph(items, steps=K, branches=2)
    for rep in range(4):
        sg.replay()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device="cuda").sum()      # unrelated eager work between replays
        for e, r, f in refs:
            e.RK3_SSP(r, f, steps=K)
        torch.cuda.synchronize()
        for (g, u, _), (_, r, _) in zip(items, refs):
            for a, b in zip(u, r):
                assert torch.equal(a, b), rep
        del junk
    # a context that changes after the capture makes replay() refuse, not replay stale buffers
    items[1][0].set_params(dt=1e-3)
    with pytest.raises(RuntimeError, match="build a new StepGraph"):
        sg.replay()
    items[1][0].set_params(dt=2e-3)
    items[0][0].calculate_convolution_2p3_UV(items[0][1], items[0][1])   # B(u,v) grows context 0's workspace
    with pytest.raises(RuntimeError, match="build a new StepGraph"):
        sg.replay()
    for g, _, _ in items + refs:
        g.close()
