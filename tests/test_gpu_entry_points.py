"""GPU tests of the exported entry points that the operator tests do not reach:
the Newton-shifted and exponential J.v callbacks, the full-length Krylov
packing, the stage-combination kernels, the physical divergence, the
non-finite checks, and the host-side ordering/invalidation rules of StepGraph.
Every comparison is against the CPU oracle or an exact identity (FP64)."""
import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pair(N=16, **kw):
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    kw.setdefault("Re", 100.0)
    kw.setdefault("dt", 2e-3)
    return NS3D(N, **kw), Oracle(N, **{k: v for k, v in kw.items() if k in ("Re", "dt")})


def test_jv_reduced_newton_matches_oracle():
    """ns3d_jv_reduced_newton (NSCallMatrixVector_reduced_Newton_Globalization,
    Source/NS3D_periodic.cu:364-389): J.v - vec/tau on the reduced packing."""
    import torch
    from tests.gpu_util import to_dev
    print("\n[case] jv_reduced_newton 16^3", flush=True)
    g, o = _pair(16)
    u = fields.random_solenoidal(16, seed=11)
    ud = to_dev(u)
    tau = 0.37
    g.set_base_state(ud, None, timesteps_period=1, tau=tau)
    rng = np.random.default_rng(3)
    vec = rng.standard_normal(g.vector_length_reduced())
    out = g.jv_reduced(torch.from_numpy(vec).cuda(), newton=True).cpu().numpy()
    ref = o.jv_reduced_newton(u, vec, tau)
    assert rel_l2(out, ref) < 1e-12
    # the shift itself, against the plain callback on the same input
    plain = g.jv_reduced(torch.from_numpy(vec).cuda()).cpu().numpy()
    assert np.max(np.abs((out - plain) + vec / tau)) <= 1e-12 * np.max(np.abs(vec / tau))
    g.close()


def test_jv_exponential_matches_oracle():
    """ns3d_jv_exponential (NSCallMatrixVector_exponential, :444-470):
    full-length unpack, Timesteps_period tangent steps RK3_SSP_UV (the base
    flow advances too), full-length pack with index 0 of each component left
    untouched."""
    import torch
    from tests.gpu_util import to_dev, to_host
    print("\n[case] jv_exponential 16^3, 3 steps", flush=True)
    g, o = _pair(16)
    u = fields.random_solenoidal(16, seed=12)
    f = fields.kolmogorov_forcing_sin(16, n=2, amp=0.3)
    ud, fd = to_dev(u), to_dev(f)
    K = 3
    g.set_base_state(ud, fd, timesteps_period=K, tau=1.0)
    N = g.n_complex
    rng = np.random.default_rng(4)
    vec = rng.standard_normal(3 * N) * 1e-3
    sentinel = np.full(3 * N, 7.25)
    vin, vout = torch.from_numpy(vec).cuda(), torch.from_numpy(sentinel.copy()).cuda()
    g.jv_exponential(vin, vout)
    out = vout.cpu().numpy()
    ref, u_adv = o.jv_exponential(u, f, K, vec, sentinel)
    for c in range(3):
        assert out[c * N] == 7.25                      # index 0 untouched
    assert rel_l2(out, ref) < 1e-12
    assert rel_l2(to_host(ud), u_adv) < 1e-12          # base flow advanced K steps (Q10)
    g.close()


def test_pack_unpack_full_bit_exact():
    import torch
    from tests.gpu_util import to_dev, to_host
    print("\n[case] pack/unpack full 32x16x48", flush=True)
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    g, o = NS3D(32, 16, 48), Oracle(32, 16, 48)
    N = g.n_complex
    rng = np.random.default_rng(5)
    vec = rng.standard_normal(3 * N)
    v = g.cfield(3)
    from ns3dperiodic_amd._lib import check
    from ns3dperiodic_amd.solver import _ptr
    check(g.lib.ns3d_unpack_full(g.ctx, _ptr(torch.from_numpy(vec).cuda()), *g._c3(v)), "unpack_full")
    vh = to_host(v)
    for a, b in zip(vh, o.unpack_full(vec)):
        np.testing.assert_array_equal(a, b)
    w = fields.random_solenoidal(32, 16, 48, seed=6)
    dst = torch.from_numpy(np.full(3 * N, -3.5)).cuda()
    check(g.lib.ns3d_pack_full(g.ctx, *g._c3(to_dev(w)), _ptr(dst)), "pack_full")
    np.testing.assert_array_equal(dst.cpu().numpy(), o.pack_full(w, np.full(3 * N, -3.5)))
    g.close()


def test_stage_combination_kernels():
    """ns3d_copy3 (copy_arrays) and ns3d_axpby3 (single_RK_step,
    Source/RK_time_step.cu:82-115) against host arithmetic."""
    from tests.gpu_util import to_dev, to_host
    print("\n[case] copy3 / axpby3 32x16x48", flush=True)
    from ns3dperiodic_amd import NS3D
    g = NS3D(32, 16, 48)
    a = fields.random_solenoidal(32, 16, 48, seed=7)
    b = fields.random_solenoidal(32, 16, 48, seed=8)
    ad, bd = to_dev(a), to_dev(b)
    dst = g.cfield(3)
    g.copy_arrays(ad, dst)
    for x, y in zip(to_host(dst), a):
        np.testing.assert_array_equal(x, y)
    g.single_RK_step(ad, 0.75, bd, 0.25, dst)
    ref = [0.75 * x + 0.25 * y for x, y in zip(a, b)]
    assert rel_l2(to_host(dst), ref) < 1e-15
    # in place on an operand, as RK3_SSP's u = 1/3 u + 2/3 u2 (Source/RK_time_step.cu:324)
    g.single_RK_step(ad, 1.0 / 3.0, bd, 2.0 / 3.0, ad)
    assert rel_l2(to_host(ad), [x / 3.0 + 2.0 * y / 3.0 for x, y in zip(a, b)]) < 1e-15
    g.close()


def test_divergence_physical_matches_oracle():
    """ns3d_divergence_physical: solve_divergence then devergence_to_double
    (Source/divergence.cu:8-47), against the oracle's divergence and C2R."""
    from tests.gpu_util import to_dev, to_host
    print("\n[case] divergence_physical 32x16x48", flush=True)
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    g, o = NS3D(32, 16, 48, dt=2e-3), Oracle(32, 16, 48, dt=2e-3)
    rng = np.random.default_rng(9)
    # a field that is not solenoidal, so the divergence is not round-off
    u = [np.fft.rfftn(rng.standard_normal((32, 16, 48))) for _ in range(3)]
    dh, dr = g.divergence_physical(to_dev(u))
    dref = o.divergence(u)
    assert rel_l2(to_host([dh])[0], dref) < 1e-14
    assert rel_l2(to_host([dr])[0], o.fft_c2r(dref)) < 1e-13
    g.close()


def test_nonfinite_detection():
    """The reference checks its state for NaN (check_nans_kernel,
    Source/cuda_supp.cu:242-325; NS3D_periodic.cu:1357-1360).  Here:
    ns3d_check_finite on any array, the sticky flag the fused epilogues set,
    and NS3D_FLAG_CHECK_FINITE, which makes the RK3 entry points return
    NS3D_ERR_NONFINITE.  A NaN in the input does not fault the GPU: it only
    propagates through the arithmetic."""
    import torch
    from ns3dperiodic_amd import NS3D
    from ns3dperiodic_amd._lib import NonFiniteError
    from tests.gpu_util import to_dev
    print("\n[case] non-finite detection 16^3", flush=True)
    u = fields.random_solenoidal(16, seed=13)
    g = NS3D(16, Re=100.0, dt=2e-3)
    ud = to_dev(u)
    assert g.check_nans(ud[0]) and g.check_nans(g.iFFTN(ud[1]))
    g.RK3_SSP(ud, None, steps=2)
    assert g.finite_status()                       # clean steps leave the flag clear
    bad = [c.clone() for c in ud]
    bad[1][3, 2, 1] = complex(float("nan"), 0.0)
    assert not g.check_nans(bad[1])
    inf = g.rfield(1)
    inf[5, 6, 7] = float("inf")
    assert not g.check_nans(inf)
    g.RK3_SSP(bad, None, steps=1)
    assert not g.finite_status(clear=True)         # the epilogue flagged it
    assert g.finite_status()                       # and the flag was cleared
    # asynchronous copy of the flag into pinned host memory
    host = torch.zeros(1, dtype=torch.int32).pin_memory()
    g.RK3_SSP(bad, None, steps=1)
    g.finite_flag_async(host)
    torch.cuda.synchronize()
    assert int(host[0]) == 1
    g.finite_status(clear=True)
    g.close()
    # NS3D_FLAG_CHECK_FINITE: the stepping call itself reports
    gc = NS3D(16, Re=100.0, dt=2e-3, check_finite=True)
    ok = to_dev(u)
    gc.RK3_SSP(ok, None, steps=3)                  # finite: no error
    bad = [c.clone() for c in ok]
    bad[0][1, 1, 1] = complex(float("inf"), 0.0)
    with pytest.raises(NonFiniteError):
        gc.RK3_SSP(bad, None, steps=1)
    gc.RK3_SSP(ok, None, steps=1)                  # the failing call cleared the flag
    gc.close()


def test_step_graph_orders_against_eager_work_without_sync():
    """StepGraph.replay() after eager work on the branch streams and before
    eager work on the current stream, with no host synchronisation between
    them: every result bitwise equal to eager stepping.  A WENO diagnostic or a
    registered Krylov base state between replays (buffers the step does not
    use) leaves the graph valid."""
    import torch
    from ns3dperiodic_amd import NS3D, StepGraph
    print("\n[case] step graph ordering, 32^3 + 16^3", flush=True)
    specs = [(32, 32, 32), (16, 16, 16)]
    items, refs = [], []
    for i, sh in enumerate(specs):
        u = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(*sh, seed=40 + i)]
        items.append((NS3D(*sh, Re=200.0, dt=2e-3), u, None))
        refs.append((NS3D(*sh, Re=200.0, dt=2e-3), [c.clone() for c in u], None))
    K = 2
    sg = StepGraph(items, steps=K, branches=2)
    for rep in range(3):
        for g, u, _ in items:                          # eager step queued on the branch stream
            g.RK3_SSP(u, None, steps=1)
        sg.replay()
        e_items = [g.kinetic_energy_spectral(u) for g, u, _ in items]   # eager, right after the replay
        for e, r, _ in refs:
            e.RK3_SSP(r, None, steps=1 + K)
        e_refs = [e.kinetic_energy_spectral(r) for e, r, _ in refs]
        assert e_items == e_refs, rep
        for (_, u, _), (_, r, _) in zip(items, refs):
            for a, b in zip(u, r):
                assert torch.equal(a, b), rep
        # buffers outside the step: WENO scratch and the J.v unpack buffer
        g0, u0, _ = items[0]
        g0.WenoAdvection(5, u0)
        g0.set_base_state(u0, None)
    torch.cuda.synchronize()
    for g, _, _ in items + refs:
        g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32s"])
def test_serve_ledger_matches_workspace(precision):
    """serve.field_bytes / request_bytes against the library's own count of
    what a context allocated (ns3d_workspace_bytes) plus the resident tensors."""
    import torch
    from ns3dperiodic_amd import NS3D, serve
    from tests.gpu_util import to_dev
    N = 48
    print(f"\n[case] serve ledger {N}^3 {precision}", flush=True)
    nc = N * N * (N // 2 + 1)
    g = NS3D(N, precision=precision, dt=1e-3)
    u = to_dev(fields.random_solenoidal(N, seed=1))
    g.RK3_SSP(u, None, steps=1)
    torch.cuda.synchronize()
    resident = 16 * 4 * nc                              # state + forcing x
    assert int(g.lib.ns3d_workspace_bytes(g.ctx)) + resident == serve.field_bytes(N, precision)
    g.close()
    if precision == "fp32s":
        return
    g = NS3D(N, precision=precision, dt=1e-3)
    u, v = to_dev(fields.random_solenoidal(N, seed=1)), to_dev(fields.random_solenoidal(N, seed=2))
    g.RK3_SSP_UV(u, v)
    torch.cuda.synchronize()
    req = serve.Request(0, "jv", N, 100.0, 1.0, 0, 0.0)
    assert int(g.lib.ns3d_workspace_bytes(g.ctx)) + resident + 16 * 3 * nc == serve.request_bytes(req, precision)
    g.close()
