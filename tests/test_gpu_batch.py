"""Batched RK3 (ns3d_rk3_ssp_batch): several same-size fields advanced by one
set of launches give each member bitwise the result of stepping it alone,
and the batched finite-difference J.v of the time-T map is the one computed
member by member and is linear in the perturbation to O(eps)."""
import numpy as np
import pytest

from ns3dperiodic_amd import fields

pytestmark = pytest.mark.gpu

SHAPE = (32, 48, 16)


def _members(shape, n, seed):
    return [fields.random_solenoidal(*shape, seed=seed + i, energy=0.5) for i in range(n)]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32s"])
def test_batch_equals_members_alone(precision):
    import torch
    from ns3dperiodic_amd import NS3D
    print(f"\n[case] batch of 3 at {SHAPE} {precision}", flush=True)
    f = fields.kolmogorov_forcing_sin(*SHAPE, n=4, amp=0.1)
    fd = [torch.from_numpy(c).cuda() for c in f]
    mem = _members(SHAPE, 3, seed=40)
    g = NS3D(*SHAPE, Re=200.0, dt=2e-3, precision=precision)
    ub = torch.from_numpy(np.stack([np.stack(m) for m in mem])).cuda().contiguous()
    g.RK3_SSP_batch(ub, fd, steps=3)
    for b, m in enumerate(mem):
        u = [torch.from_numpy(c).cuda() for c in m]
        g.RK3_SSP(u, fd, steps=3)
        for k in range(3):
            assert torch.equal(ub[b, k], u[k]), (b, k)
    g.close()


def test_batch_member_against_oracle():
    import torch
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    print(f"\n[case] batch member vs oracle at {SHAPE}", flush=True)
    mem = _members(SHAPE, 2, seed=50)
    g = NS3D(*SHAPE, Re=200.0, dt=2e-3)
    o = Oracle(*SHAPE, Re=200.0, dt=2e-3)
    ub = torch.from_numpy(np.stack([np.stack(m) for m in mem])).cuda().contiguous()
    g.RK3_SSP_batch(ub, None, steps=2)
    zero = [np.zeros(o.cshape, np.complex128)] * 3
    for b, m in enumerate(mem):
        ref = o.rk3(m, zero, steps=2)
        got = ub[b].cpu().numpy()
        err = np.sqrt(sum(np.sum(np.abs(got[k] - ref[k]) ** 2) for k in range(3))
                      / sum(np.sum(np.abs(ref[k]) ** 2) for k in range(3)))
        print(f"[batch] member {b}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
    g.close()


def test_fd_jv_batch():
    import torch
    from ns3dperiodic_amd import NS3D, krylov
    print(f"\n[case] batched FD J.v at {SHAPE}", flush=True)
    u = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(*SHAPE, seed=60, energy=0.5)]
    V = torch.from_numpy(np.stack([np.stack(m) for m in _members(SHAPE, 4, seed=70)])).cuda().contiguous()
    g = NS3D(*SHAPE, Re=200.0, dt=2e-3)
    steps, eps = 5, 1e-7
    J = krylov.fd_jv_batch(g, u, V, eps=eps, steps=steps)
    # member by member: the same quotient from single-field steps, bitwise
    base = [c.clone() for c in u]
    g.RK3_SSP(base, None, steps=steps)
    for i in range(V.shape[0]):
        w = [(u[k] + eps * V[i, k]).contiguous() for k in range(3)]
        g.RK3_SSP(w, None, steps=steps)
        for k in range(3):
            assert torch.equal(J[i, k], (w[k] - base[k]) / eps)
    # linearity to O(eps): J(2v) = 2 J(v); central and forward quotients agree
    J2 = krylov.fd_jv_batch(g, u, 2.0 * V, eps=eps, steps=steps)
    Jc = krylov.fd_jv_batch(g, u, V, eps=1e-5, steps=steps, central=True)
    nrm = float(torch.linalg.vector_norm(J))
    assert float(torch.linalg.vector_norm(J2 - 2.0 * J)) < 1e-5 * nrm
    assert float(torch.linalg.vector_norm(Jc - J)) < 1e-5 * nrm
    g.close()


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp64", True), ("fp32", True)])
def test_batch_members_per_member_dt_re_and_steps(precision, graph):
    """ns3d_rk3_ssp_batch_members: each member takes its own number of the
    launch's steps with its own dt and Re, bitwise as it would alone; an idle
    member (0 steps) is left untouched; a second call with new parameters
    (the same graph replayed) continues each member correctly."""
    import torch
    from ns3dperiodic_amd import NS3D
    print(f"\n[case] batch members {precision} graph={graph} at {SHAPE}", flush=True)
    f = fields.kolmogorov_forcing_sin(*SHAPE, n=4, amp=0.1)
    fd = [torch.from_numpy(c).cuda() for c in f]
    mem = _members(SHAPE, 4, seed=80)
    dts = [2e-3, 1e-3, 3e-3, 2e-3]
    res = [200.0, 80.0, 350.0, 120.0]
    steps1 = [4, 2, 0, 3]
    steps2 = [1, 4, 3, 0]
    s = torch.cuda.Stream()
    gb = NS3D(*SHAPE, Re=1.0, dt=1.0, precision=precision, stream=s)
    ub = torch.from_numpy(np.stack([np.stack(m) for m in mem])).cuda().contiguous()
    dt_d = torch.tensor(dts, dtype=torch.float64, device="cuda")
    re_d = torch.tensor(res, dtype=torch.float64, device="cuda")
    with torch.cuda.stream(s):
        left = torch.tensor(steps1, dtype=torch.int32, device="cuda")
        gb.RK3_SSP_batch_members(ub, fd, dt_d, re_d, left, 4, graph=graph)
        left.copy_(torch.tensor(steps2, dtype=torch.int32))
        gb.RK3_SSP_batch_members(ub, fd, dt_d, re_d, left, 4, graph=graph)
    torch.cuda.synchronize()
    for b, m in enumerate(mem):
        g = NS3D(*SHAPE, Re=res[b], dt=dts[b], precision=precision)
        u = [torch.from_numpy(c).cuda() for c in m]
        g.RK3_SSP(u, fd, steps=steps1[b] + steps2[b])
        for k in range(3):
            assert torch.equal(ub[b, k], u[k]), (b, k)
        g.close()
    gb.close()
