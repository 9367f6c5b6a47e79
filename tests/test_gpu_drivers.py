"""GPU tests of the drivers around the hot path: the batched probe,
execute_return_map's seed disc (batched and the reference's chained order),
the time_dependant.dat series, and the Kolmogorov-flow workload at 128^3
(FP64 against the oracle; FP32 modes against FP64 over 2000 steps)."""
import math

import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_probe_batch_matches_single_probes():
    import torch
    from ns3dperiodic_amd import NS3D
    print("\n[case] probe_batch 32x16x48, 5 members", flush=True)
    shape = (32, 16, 48)
    g = NS3D(*shape)
    ub = torch.stack([torch.stack([torch.from_numpy(c) for c in fields.random_solenoidal(*shape, seed=70 + b)])
                      for b in range(5)]).cuda().contiguous()
    for pt in [(0, 0, 0), (5, 7, 11), (31, 15, 47)]:
        got = g.probe_batch(ub, *pt)
        for b in range(5):
            ref = g.probe([ub[b, 0], ub[b, 1], ub[b, 2]], *pt)
            assert np.allclose(got[b], ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max())), (pt, b)
    with pytest.raises(Exception):
        g.probe_batch(ub, 32, 0, 0)


def _replay_seed(N, Re, dt, u0, f, pt, seed, steps, t_cross):
    """The oracle's version of one seed: the seed value imposed at the probe
    point in physical space, then steps - 1 steps of dt and one of the bisected length."""
    from oracle.oracle import Oracle
    us = []
    for c, v in zip(u0, seed):
        x = np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))
        x[pt] = v
        us.append(np.fft.rfftn(x, axes=(0, 1, 2)))
    h = t_cross - (steps - 1) * dt
    uo = Oracle(N, Re=Re, dt=dt).rk3(us, f, steps=steps - 1) if steps > 1 else us
    uo = Oracle(N, Re=Re, dt=h).rk3(uo, f)
    return np.array([np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))[pt] for c in uo])


def test_execute_return_map_batched_disc():
    """The 7x7 seed disc advanced as one batch: every seed crosses the shifted
    plane, the crossings lie on it, and one seed's crossing is replayed by the
    oracle (seed imposed, same step sequence)."""
    import torch
    from ns3dperiodic_amd import NS3D, return_map
    print("\n[case] execute_return_map 16^3, 49 seeds batched", flush=True)
    N, Re, dt, pt = 16, 50.0, 2e-3, (3, 5, 7)
    u0 = fields.random_solenoidal(N, seed=62)
    f = fields.kolmogorov_forcing_sin(N, n=4)
    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).cuda() for c in u0]
    fd = [torch.from_numpy(c).cuda() for c in f]
    res = return_map.execute_return_map(g, u, fd, point=pt, shift_steps=8, rho=0.5, max_steps=60, trajectories=True)
    print(f"[return_map] found {int(res['found'].sum())}/49, steps {sorted(set(res['steps'].tolist()))}, "
          f"max |s| {np.nanmax(res['errors']):.1e}", flush=True)
    assert res["found"].all()
    assert np.all(res["errors"] <= 1e-12)
    x0s, ns = res["x0_shift"], res["normal_shift"]
    for p in res["points"]:
        assert abs(return_map.plane_value(p, x0s, ns)) <= 1e-12
    # the seeds lie on the plane through x0 normal to the RHS there
    for s in res["seeds"]:
        assert abs(return_map.plane_value(s, res["x0"], res["normal"])) <= 1e-12 * max(1.0, np.linalg.norm(res["normal"]))
    # the caller's state is untouched
    assert rel_l2([c.cpu().numpy() for c in u], u0) == 0.0
    j = 17
    po = _replay_seed(N, Re, dt, u0, f, pt, res["seeds"][j], int(res["steps"][j]), float(res["time"][j]))
    assert np.allclose(res["points"][j], po, rtol=0, atol=1e-10), (res["points"][j], po)
    assert len(res["trajectories"][j]) == int(res["steps"][j]) + 1


def test_execute_return_map_chained_like_the_reference():
    """chain=True: seeds one after another, each from the previous seed's crossing state."""
    import torch
    from ns3dperiodic_amd import NS3D, return_map
    print("\n[case] execute_return_map 16^3, 2x2 seeds chained", flush=True)
    N, Re, dt, pt = 16, 50.0, 2e-3, (3, 5, 7)
    u0 = fields.random_solenoidal(N, seed=62)
    f = fields.kolmogorov_forcing_sin(N, n=4)
    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).cuda() for c in u0]
    fd = [torch.from_numpy(c).cuda() for c in f]
    res = return_map.execute_return_map(g, u, fd, point=pt, n_rad=2, n_phi=2, shift_steps=8, rho=0.5, max_steps=400,
                                        chain=True)
    print(f"[return_map] chained steps {res['steps'].tolist()}", flush=True)
    assert res["found"].all() and np.all(res["errors"] <= 1e-12)
    # the first seed starts from u itself, so it matches the batched disc's first seed
    rb = return_map.execute_return_map(g, u, fd, point=pt, n_rad=2, n_phi=2, shift_steps=8, rho=0.5, max_steps=60)
    assert rb["steps"][0] == res["steps"][0]
    assert np.allclose(rb["points"][0], res["points"][0], rtol=0, atol=1e-12)


def test_time_series_columns():
    """time_dependant.dat rows against direct reads of the same fields."""
    import torch
    from ns3dperiodic_amd import NS3D
    from ns3dperiodic_amd.timeseries import TimeSeries
    print("\n[case] time series 32^3, 3 steps", flush=True)
    N, dt = 32, 2e-3
    g = NS3D(N, Re=100.0, dt=dt)
    u0 = fields.random_solenoidal(N, seed=80)
    u = [torch.from_numpy(c).cuda() for c in u0]
    ts = TimeSeries(g)
    for _ in range(3):
        g.RK3_SSP(u, None)
        row = ts.record(u, dt)
    uh = [c.cpu().numpy() for c in u]
    phys = [np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2)) for c in uh]
    assert row[0] == pytest.approx(2 * 3 * dt)
    assert np.allclose(row[1:4], [p[ts.P1] for p in phys], atol=1e-13)
    assert np.allclose(row[8:11], [p[ts.P2] for p in phys], atol=1e-13)
    assert np.allclose(row[15:18], [p[ts.P3] for p in phys], atol=1e-13)
    E = g.get_kinetic_energy([g.iFFTN(c) for c in u])
    D = g.get_dissipation([g.iFFTN(c) for c in u])
    assert row[7] == pytest.approx(E, rel=1e-14) and row[28] == pytest.approx(E, rel=1e-14)
    assert row[14] == pytest.approx(D, rel=1e-14) and row[35] == pytest.approx(D, rel=1e-14)
    assert row[21] == pytest.approx(E - D / 100.0, rel=1e-13)
    c = uh[0][ts.S0]
    assert row[29] == c.real and row[30] == c.imag
    assert row[36] == uh[0][ts.F1].imag and row[37] == uh[0][ts.F2].imag and row[38] == uh[1][ts.F1].imag
    assert row[42] == uh[0][2, 2, 2].imag and row[44] == uh[0][0, 0, 0].imag and row[49] == uh[0][N - 1, N - 1, N // 2].imag
    # vorticity columns: central differences of the physical field (get_curl)
    dx = 2 * math.pi / N
    wz = (np.roll(phys[1], -1, 0) - np.roll(phys[1], 1, 0)) / (2 * dx) - \
         (np.roll(phys[0], -1, 1) - np.roll(phys[0], 1, 1)) / (2 * dx)
    assert row[6] == pytest.approx(wz[ts.P1], abs=1e-11)
    assert ts.array().shape == (3, 50)


def _config2_dt(u0, N, steps, fmax=1.0, cfl=0.2):
    """dt with cfl*dx / (max|u0| + steps*dt*fmax) = dt, rounded down to 2 digits."""
    umax = max(float(np.max(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))))) for c in u0)
    dx = 2 * math.pi / N
    dt = (-umax + math.sqrt(umax * umax + 4 * steps * fmax * cfl * dx)) / (2 * steps * fmax)
    e = math.floor(math.log10(dt))
    return math.floor(dt / 10 ** (e - 1)) * 10 ** (e - 1)


def test_config2_kolmogorov_128_fp64_vs_oracle():
    """Kolmogorov flow at 128^3 (f_x = sin 4y, Re = 100, the seeded k^4 exp(-2(k/4)^2)
    field of energy 0.5): two FP64 RK3 steps against the oracle, <= 1e-12."""
    import torch
    from ns3dperiodic_amd import NS3D
    from oracle.oracle import Oracle
    print("\n[case] config 2, 128^3 FP64 vs oracle, 2 steps", flush=True)
    N, Re = 128, 100.0
    u0 = fields.random_solenoidal(N, seed=1)
    f = fields.kolmogorov_forcing_sin(N, n=4)
    dt = _config2_dt(u0, N, 2000)
    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).cuda() for c in u0]
    g.RK3_SSP(u, [torch.from_numpy(c).cuda() for c in f], steps=2)
    ref = Oracle(N, Re=Re, dt=dt).rk3(u0, f, steps=2)
    err = rel_l2([c.cpu().numpy() for c in u], ref)
    print(f"[config2] dt {dt:g}, rel L2 vs oracle after 2 steps {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("precision", ["fp32", "fp32i"])
def test_config2_kolmogorov_128_series_fp32_vs_fp64(precision):
    """2000 steps of the 128^3 Kolmogorov flow in FP64 and in an FP32 mode from
    identical bits: the energy and dissipation series (physical, every step)
    agree to 1e-5 relative, the final fields to 1e-4."""
    import torch
    from ns3dperiodic_amd import NS3D
    print(f"\n[case] config 2, 128^3, 2000 steps, {precision} vs fp64", flush=True)
    N, Re, steps = 128, 100.0, 2000
    u0 = fields.random_solenoidal(N, seed=1)
    f = fields.kolmogorov_forcing_sin(N, n=4)
    dt = _config2_dt(u0, N, steps)
    fd = [torch.from_numpy(c).cuda() for c in f]
    series = {}
    finals = {}
    for prec in ("fp64", precision):
        g = NS3D(N, Re=Re, dt=dt, precision=prec, check_finite=True)
        u = [torch.from_numpy(c).cuda() for c in u0]
        E, D = [], []
        for _ in range(steps):
            g.RK3_SSP(u, fd)
            E.append(g.kinetic_energy_spectral(u))
            D.append(g.dissipation_spectral(u))
        series[prec] = (np.array(E), np.array(D))
        finals[prec] = [c.cpu().numpy() for c in u]
        g.close()
    e64, d64 = series["fp64"]
    e32, d32 = series[precision]
    de = float(np.max(np.abs(e32 - e64) / np.abs(e64)))
    dd = float(np.max(np.abs(d32 - d64) / np.abs(d64)))
    err = rel_l2(finals[precision], finals["fp64"])
    print(f"[config2] {precision}: dt {dt:g}, T {steps * dt:.3f}, energy {e64[0]:.4f} -> {e64[-1]:.4f}, "
          f"max rel diff energy {de:.2e}, dissipation {dd:.2e}, final field rel L2 {err:.2e}", flush=True)
    assert de < 1e-5 and dd < 1e-5
    assert err < 1e-4
