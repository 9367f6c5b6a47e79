"""Continuous-batching server on the GPU.

* Each request's result is bitwise the same alone and in a mix (FP32-bulk
  and FP64), batched with the other requests of its grid or on a slot of its
  own, and the FP64 server reproduces one-at-a-time FP64 bitwise.
* FP32-bulk results stay within 1e-5 relative L2 of FP64 over these short
  horizons (the stated tolerance of the FP32-bulk path).
* A served request matches the CPU oracle (FP64, <= 1e-12).
* A small HBM budget forces queueing and still serves everything.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(n=8, seed=11, sizes=(32, 48, 64), t_range=(0.05, 0.15), jv_T=0.05):
    from ns3dperiodic_amd import serve
    reqs = serve.poisson_stream(n, rate=400.0, seed=seed, sizes=sizes, t_range=t_range, jv_fraction=0.3, jv_T=jv_T,
                                rm_fraction=0.2)
    io = {}
    for r in reqs:
        io[r.rid] = serve.pin_io(*serve.request_inputs(r))
    return reqs, io


def _rel(a, b):
    num = sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(np.abs(y) ** 2)) for y in b)
    return (num / den) ** 0.5


def _copy(results):
    return {k: ([c.copy() for c in r.u], None if r.v is None else [c.copy() for c in r.v], r.info)
            for k, r in results.items()}


def _same_info(a, b):
    if a is None or b is None:
        return a is b
    return (a["found"] == b["found"] and a["steps"] == b["steps"] and a["time"] == b["time"]
            and np.array_equal(a["point"], b["point"]))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_alone_equals_mix(precision):
    from ns3dperiodic_amd import serve
    print(f"\n[case] serve alone vs mix {precision}", flush=True)
    reqs, io = _mix()
    assert {r.kind for r in reqs} == {"jv", "integrate", "return_map"}
    srv = serve.Server(precision=precision, max_active=6, chunk=4)
    mixed = _copy(srv.run(reqs, io))
    srv.close()
    assert sorted(mixed) == [r.rid for r in reqs]
    # the batching groups (integrate requests of one grid in one launch per pass) against
    # every request on a slot of its own
    srv0 = serve.Server(precision=precision, max_active=6, chunk=4, batch=False)
    unbatched = _copy(srv0.run(reqs, io))
    srv0.close()
    for r in reqs:
        for a, b in zip(unbatched[r.rid][0], mixed[r.rid][0]):
            assert np.array_equal(a, b), f"request {r.rid} ({r.kind}, N={r.N}) differs batched vs on its own slot"
    for r in reqs:
        srv1 = serve.Server(precision=precision, max_active=1, chunk=4)
        alone = srv1.run([dataclasses.replace(r, arrival=0.0)], io)[r.rid]
        srv1.close()
        u, v, info = mixed[r.rid]
        for a, b in zip(alone.u, u):
            assert np.array_equal(a, b), f"request {r.rid} ({r.kind}, N={r.N}) differs alone vs mix"
        assert _same_info(alone.info, info)
        assert (info is not None) == (r.kind == "return_map")
        if r.kind == "jv":
            for a, b in zip(alone.v, v):
                assert np.array_equal(a, b)
    if precision == "fp64":
        base = serve.one_at_a_time(reqs, io, precision="fp64")
        for r in reqs:
            for a, b in zip(base[r.rid].u, mixed[r.rid][0]):
                assert np.array_equal(a, b), f"request {r.rid} differs from one-at-a-time FP64"
            assert _same_info(base[r.rid].info, mixed[r.rid][2])


def test_fp32_served_vs_fp64_and_oracle():
    from ns3dperiodic_amd import serve
    from oracle.oracle import Oracle
    print("\n[case] served fp32 vs fp64 vs oracle", flush=True)
    reqs, io = _mix(n=6, seed=5)
    s32 = serve.Server(precision="fp32", max_active=4, chunk=4)
    r32 = _copy(s32.run(reqs, io))
    s32.close()
    s64 = serve.Server(precision="fp64", max_active=4, chunk=4)
    r64 = _copy(s64.run(reqs, io))
    s64.close()
    for r in reqs:
        if r.kind == "return_map":
            continue    # the two precisions may stop at different steps
        err = _rel(r32[r.rid][0], r64[r.rid][0])
        print(f"[serve] rid {r.rid} {r.kind} N={r.N} steps={r.nsteps} fp32 vs fp64 u: {err:.2e}", flush=True)
        assert err < 1e-5
        if r.kind == "jv":
            assert _rel(r32[r.rid][1], r64[r.rid][1]) < 1e-5
    # one integrate request against the oracle (FP64, forcing f_x = sin 4y)
    r = next(q for q in reqs if q.kind == "integrate" and q.N == min(p.N for p in reqs if p.kind == "integrate"))
    o = Oracle(r.N, Re=r.Re, dt=r.dt, ipq=False, mask_nyq=True)   # the server's physical mode
    u0 = [c.numpy() for c in io[r.rid]["u"]]
    f = [io[r.rid]["f"].numpy(), np.zeros_like(u0[0]), np.zeros_like(u0[0])]
    ref = o.rk3(u0, f, steps=r.nsteps)
    err = _rel(r64[r.rid][0], ref)
    print(f"[serve] rid {r.rid} FP64 served vs oracle after {r.nsteps} steps: {err:.2e}", flush=True)
    assert err < 1e-12 * max(1, r.nsteps) ** 0.5


def test_small_budget_queues_and_serves_all():
    from ns3dperiodic_amd import serve
    print("\n[case] serve under a small HBM budget", flush=True)
    reqs, io = _mix(n=6, seed=2)
    biggest = max(serve.request_bytes(r, "fp32") for r in reqs)
    srv = serve.Server(precision="fp32", max_active=8, budget_bytes=biggest, chunk=4)
    res = srv.run(reqs, io)
    assert sorted(res) == [r.rid for r in reqs]
    assert srv.ledger.used == 0 and srv.ledger.active == 0
    assert srv.ledger.peak <= biggest
    assert srv.ledger.peak_active < len(reqs)     # the budget, not the slots, limited residency
    srv.close()
