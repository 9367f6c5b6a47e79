"""Host-side logic of the request server (no GPU): request stream, dt rule,
HBM-budget admission and its fairness."""
import math

import numpy as np
import pytest

from ns3dperiodic_amd import serve


def test_poisson_stream_is_seeded_and_in_range():
    a = serve.poisson_stream(200, rate=5.0, seed=3)
    b = serve.poisson_stream(200, rate=5.0, seed=3)
    assert [(r.kind, r.N, r.Re, r.T, r.arrival) for r in a] == [(r.kind, r.N, r.Re, r.T, r.arrival) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    # mean inter-arrival 1/rate within sampling error
    assert abs(np.mean(np.diff(arr)) - 0.2) < 0.05
    assert {r.N for r in a} == set(serve.GRID_SIZES)
    assert all(50.0 <= r.Re <= 500.0 for r in a)
    assert all(0.5 <= r.T <= 5.0 for r in a if r.kind == "integrate")
    assert all(r.T == 1.0 for r in a if r.kind == "jv")
    assert 0.1 < np.mean([r.kind == "jv" for r in a]) < 0.4


def test_stream_with_return_map_requests():
    a = serve.poisson_stream(400, rate=5.0, seed=4, jv_fraction=0.25, rm_fraction=0.25)
    frac = {k: np.mean([r.kind == k for r in a]) for k in ("jv", "return_map", "integrate")}
    assert abs(frac["jv"] - 0.25) < 0.07 and abs(frac["return_map"] - 0.25) < 0.07
    assert all(0.5 <= r.T <= 5.0 for r in a if r.kind == "return_map")
    r = next(q for q in a if q.kind == "return_map")
    i = serve.Request(0, "integrate", r.N, 1.0, 1.0, 0, 0.0)
    nc = r.N * r.N * (r.N // 2 + 1)
    assert serve.request_bytes(r, "fp32") == serve.request_bytes(i, "fp32") + 48 * nc
    assert all(0 <= p < r.N for p in serve.probe_point(r.N))


def test_choose_dt_divides_horizon_and_respects_cfl():
    from ns3dperiodic_amd import fields
    N = 16
    u = fields.random_solenoidal(N, seed=5)
    n, dt = serve.choose_dt(u, 1.3, N, cfl=0.2)
    assert n * dt == pytest.approx(1.3, rel=1e-14)
    umax = max(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))).max() for c in u)
    assert dt <= 0.2 * (2 * math.pi / N) / umax * (1 + 1e-12)
    assert (n - 1) * 0.2 * (2 * math.pi / N) / umax < 1.3   # smallest such n


def test_request_inputs_unit_perturbation():
    r = serve.Request(0, "jv", 16, 100.0, 1.0, seed=9, arrival=0.0)
    u, fx, v = serve.request_inputs(r)
    assert r.nsteps >= 1 and r.dt > 0
    assert math.isclose(sum(float(np.vdot(c, c).real) for c in v), 1.0, rel_tol=1e-12)
    assert fx[0, 4, 0] == pytest.approx(-0.5j * 16 ** 3)


def test_request_bytes():
    r = serve.Request(0, "integrate", 64, 100.0, 1.0, 0, 0.0)
    j = serve.Request(1, "jv", 64, 100.0, 1.0, 0, 0.0)
    nc = 64 * 64 * 33
    n8 = 64 * 64 * 40                          # rows of Mz = 33 padded to 8 modes (FP64 and FP32 fields)
    # 4 resident FP64 fields (state + forcing x), 9 FP64 stage fields, 15 scratch fields
    assert serve.request_bytes(r, "fp64") == 16 * 4 * nc + 16 * 9 * n8 + 16 * 15 * n8
    assert serve.request_bytes(r, "fp32") == 16 * 4 * nc + 16 * 9 * n8 + 8 * 15 * n8
    # FP32 state loop: 9 FP32 fields replace the FP64 stages
    assert serve.request_bytes(r, "fp32s") == 16 * 4 * nc + 8 * 9 * n8 + 8 * 15 * n8
    # J.v: perturbation (3 FP64), tangent stage (6 FP64 padded), 15 more scratch fields
    assert serve.request_bytes(j, "fp32") == serve.request_bytes(r, "fp32") + 16 * 3 * nc + 16 * 6 * n8 + 8 * 15 * n8


def test_ledger_limits():
    L = serve.Ledger(100, 2)
    L.admit(60)
    assert not L.fits(50) and L.fits(40)
    L.admit(40)
    assert not L.fits(0)          # slot limit
    L.release(60)
    assert L.fits(50)
    with pytest.raises(RuntimeError):
        L.admit(70)


def _q(costs):
    return [serve.Request(i, "integrate", 32, 100.0, 1.0, 0, 0.0) for i in range(len(costs))], dict(enumerate(costs))


def test_admission_fifo_and_overtaking_bound():
    reqs, cost = _q([50, 80, 10, 10, 10, 10])
    L = serve.Ledger(100, 8)
    c = lambda r: cost[r.rid]   # noqa: E731
    picked = serve.admit_order(reqs, L, c, max_overtakes=2)
    # 50 fits; 80 does not and is overtaken by the next two 10s, then blocks the rest
    assert [r.rid for r in picked] == [0, 2, 3]
    assert reqs[1].overtaken == 2
    # once the running work ends, the blocked request goes first
    L2 = serve.Ledger(100, 8)
    rest = [r for r in reqs if r.rid not in (0, 2, 3)]
    assert [r.rid for r in serve.admit_order(rest, L2, c)] == [1, 4, 5]


def test_admission_slot_limit():
    reqs, cost = _q([1] * 10)
    L = serve.Ledger(1 << 40, 3)
    assert [r.rid for r in serve.admit_order(reqs, L, lambda r: cost[r.rid])] == [0, 1, 2]


def test_summarize():
    rs = [serve.Result(i, 0.0, 1.0 + i, 1.0 + i, None) for i in range(100)]
    s = serve.summarize(rs)
    assert s["solves"] == 100 and s["span_s"] == 100.0
    assert s["p50_latency_s"] == pytest.approx(50.5)
    assert s["p99_latency_s"] == pytest.approx(99.01)
