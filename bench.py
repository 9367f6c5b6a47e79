#!/usr/bin/env python3
"""RK3 steps/s of the fused gfx950 pseudo-spectral step (NS3DPeriodic hot path).

Workload: decaying turbulence at 256^3 on [0,2pi]^3 (nu = 2e-3, i.e. Re = 500),
seeded solenoidal field with E(k) ~ k^4 exp(-2 (k/4)^2), energy 0.5, reference
mode (2/3 mask, IPQ rule).  A step is one RK3_SSP call on state resident in HBM.

Prints ONE JSON line.  `value` is steps/s of the primary precision (default
FP64, the parity-pinned path); the other precision is reported beside it.
`roofline` is for the dominant kernel: algorithmic bytes per launch / average
launch duration, both measured here with HIP events on the library's stream.
`cpu_baseline` times the CPU oracle (FP64 restatement, OpenMP) on a bounded
sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
PASS_NAMES = ["x_inverse", "y_inverse", "z_product", "y_forward", "x_forward_epilogue"]
PASS_KERNELS = ["k_x_inv_grad", "k_y_inv_grad", "k_z_product", "k_axis_c2c", "k_x_fwd_epi"]


def pass_bytes(nc, bulk_bytes, state_bytes=16):
    """Algorithmic HBM bytes per launch of each pass of one RK3 stage.

    Field traffic of the five-pass dataflow (S = one spectral field of nc
    modes): P1 reads 3 state fields and writes 6; P2 reads 6, writes 9; P3
    reads 9, writes 3; P4 reads 3, writes 3; P5 reads 3 bulk + 3 state (+3
    u^n in stages 2-3) and writes 3 state.  State fields are FP64 complex
    (16 B) except in the FP32-state loop (8 B); bulk fields use the bulk
    precision.  With one precision throughout the RK3 step totals 159 S, the
    algorithmic byte count of the path.
    """
    s, b = state_bytes * nc, bulk_bytes * nc
    p5_avg = 3 * b + 3 * s + 3 * s + (2.0 / 3.0) * 3 * s
    return [3 * s + 6 * b, 15 * b, 12 * b, 6 * b, p5_avg]


def step_bytes(nc, bulk_bytes, state_bytes=16):
    pb = pass_bytes(nc, bulk_bytes, state_bytes)
    return 3 * sum(pb)


def pass_bytes_incr(nc):
    """Algorithmic bytes per launch of each pass, averaged over the three
    stages, of the FP32-increment RK3 loop (FP32 bulk; the state u^n as the
    FP32 pair hi + lo; FP32 stage increments d; S = 8 nc bytes):
    P1 reads hi (3) [+ d (3) in stages 2-3] and writes 6; P2-P4 as in the
    FP32 path (15, 12, 6); P5 reads Q (3) and hi (3) and writes d (3) in
    stage 1, also reads d in stage 2 (12), and in stage 3 reads Q, hi, lo, d
    (12) and writes hi, lo (6) (18).  Per step 3 x (11 + 15 + 12 + 6 + 13) =
    171 S (the split and join of the caller's FP64 state, once per call, are
    not counted)."""
    S = 8 * nc
    return [11 * S, 15 * S, 12 * S, 6 * S, 13 * S]


def profiled_traffic(precision, kernel_prefix):
    """HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    of a kernel at 256^3, from the committed profile summary; None if absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "rk3_256_passes_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if not line.startswith("{"):
            continue
        d = json.loads(line)
        if d["precision"] == precision and d["kernel"].startswith(kernel_prefix + "<"):
            return d["fetch_bytes"] + d["write_bytes"]
    return None


def run_gpu(N, precision, steps, warmup, u0, Re, dt):
    import torch
    from ns3dperiodic_amd import NS3D

    g = NS3D(N, Re=Re, dt=dt, precision=precision)
    u = [torch.from_numpy(c).to("cuda:0") for c in u0]
    g.RK3_SSP(u, None, steps=warmup)
    torch.cuda.synchronize()
    lib = g.lib
    import ctypes
    # the timed steps run without per-pass events; a shorter run with them
    # (HIP events around every pass, on the library's stream) gives the pass times
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    g.RK3_SSP(u, None, steps=steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lib.ns3d_profile_enable(g.ctx, 1)
    g.RK3_SSP(u, None, steps=max(1, min(steps, 10)))
    torch.cuda.synchronize()
    ms = (ctypes.c_double * 5)()
    calls = (ctypes.c_longlong * 5)()
    lib.ns3d_profile_read(g.ctx, ms, calls)
    lib.ns3d_profile_enable(g.ctx, 0)
    total_ms = e0.elapsed_time(e1)
    finite = bool(torch.isfinite(u[0]).all().item())
    g.close()
    bulk = 16 if precision == "fp64" else 8
    state = 8 if precision == "fp32s" else 16
    nc = N * N * (N // 2 + 1)
    pb = pass_bytes_incr(nc) if precision == "fp32i" else pass_bytes(nc, bulk, state)
    passes = []
    for i in range(5):
        avg = ms[i] / calls[i] if calls[i] else float("nan")
        gbs = pb[i] / (avg * 1e-3) / 1e9 if calls[i] else float("nan")
        passes.append({"pass": PASS_NAMES[i], "kernel": PASS_KERNELS[i], "calls": calls[i], "avg_ms": avg,
                       "total_ms": ms[i], "alg_bytes": pb[i], "GBps": gbs})
    ms_step = total_ms / steps
    sb = 3 * sum(pb) if precision == "fp32i" else step_bytes(nc, bulk, state)
    return {
        "precision": precision,
        "steps_per_s": steps / (total_ms * 1e-3),
        "ms_per_step": ms_step,
        "wall_s": wall,
        "step_alg_bytes": sb,
        "step_GBps": sb / (ms_step * 1e-3) / 1e9,
        "passes": passes,
        "finite": finite,
    }


def run_scalar(N, steps, warmup, u0, theta0, Re, dt):
    """RK3 of the velocity plus a WENO5-advected passive scalar (FP64), timed
    like run_gpu; algorithmic bytes 159 S + 26 R per step (R = one physical
    field: the stage's velocity written by the z pass (3), read by the scalar
    update with theta_s (3 + 1), theta written (1), theta^n in stages 2-3)."""
    import torch
    from ns3dperiodic_amd import NS3D

    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).to("cuda:0") for c in u0]
    th = torch.from_numpy(theta0).to("cuda:0")
    g.RK3_SSP_scalar(u, th, None, steps=warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    g.RK3_SSP_scalar(u, th, None, steps=steps)
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    finite = bool(torch.isfinite(th).all().item())
    g.close()
    nc = N * N * (N // 2 + 1)
    sb = step_bytes(nc, 16) + 26 * 8 * N ** 3
    ms_step = total_ms / steps
    return {"steps_per_s": steps / (total_ms * 1e-3), "ms_per_step": ms_step, "step_alg_bytes": sb,
            "step_roofline_frac": sb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "finite": finite}


def run_accuracy(N, steps, u0, Re, dt, precisions=("fp32", "fp32s", "fp32i")):
    """Relative L2 distance of each FP32 mode from the FP64 path after `steps`
    RK3 steps from identical initial bits (the FP32 accuracy bar is 1e-5
    after 1000 steps)."""
    import torch
    from ns3dperiodic_amd import NS3D
    g = NS3D(N, Re=Re, dt=dt, precision="fp64")
    ref = [torch.from_numpy(c).to("cuda:0") for c in u0]
    g.RK3_SSP(ref, None, steps=steps)
    g.close()
    den = sum(float(torch.sum(torch.abs(y) ** 2)) for y in ref)
    out = {"steps": steps}
    for p in precisions:
        g = NS3D(N, Re=Re, dt=dt, precision=p)
        u = [torch.from_numpy(c).to("cuda:0") for c in u0]
        g.RK3_SSP(u, None, steps=steps)
        g.close()
        out[names_of(p)] = math.sqrt(sum(float(torch.sum(torch.abs(x - y) ** 2)) for x, y in zip(u, ref)) / den)
        del u
    del ref
    torch.cuda.empty_cache()
    return out


def run_tangent(N, precision, steps, u0, Re, dt):
    """Tangent-map steps (RK3_SSP_UV: the base flow advanced together with a
    perturbation through the bilinear passes, 24 real fields per point in the
    z pass) per second, beside the plain step rate."""
    import torch
    from ns3dperiodic_amd import NS3D, fields
    g = NS3D(N, Re=Re, dt=dt, precision=precision)
    u = [torch.from_numpy(c).to("cuda:0") for c in u0]
    v = [torch.from_numpy(c).to("cuda:0") for c in fields.random_solenoidal(N, seed=77, energy=1e-6)]
    g.RK3_SSP_UV(u, v)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    e0.record(s)
    for _ in range(steps):
        g.RK3_SSP_UV(u, v)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    g.close()
    return {"precision": precision, "steps_per_s": 1e3 / ms, "ms_per_step": ms}


def run_fd_jv_batch(N=64, m=64, Re=150.0, T=1.0, eps=1e-7, seed=1):
    """Batched Krylov work: finite-difference J.v of the time-T map for m
    seeded unit-norm solenoidal perturbations of the N^3 Kolmogorov snapshot
    (f_x = sin 4y, Re = 150, 2000 steps from the seeded k^4 exp(-2(k/4)^2)
    field), all m + 1 fields advanced together (ns3d_rk3_ssp_batch), against
    the same m + 1 solves one field at a time.  FP64 (eps = 1e-7 is below
    FP32 resolution)."""
    import torch
    from ns3dperiodic_amd import NS3D, fields, krylov
    u0, f, t_snap = kolmogorov_snapshot(N, Re, seed=seed)
    umax = max(float(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))).max()) for c in u0)
    steps = int(math.ceil(T / (0.2 * (2 * math.pi / N) / (umax + T))))
    dt = T / steps
    vs = []
    for i in range(m):
        v = np.stack(fields.random_solenoidal(N, seed=1000 + i, kind="k4exp", k0=4.0, energy=0.5))
        vs.append(v / np.sqrt(np.sum(np.abs(v) ** 2)))
    V = torch.from_numpy(np.stack(vs)).cuda().contiguous()
    u = [torch.from_numpy(c).cuda() for c in u0]
    fd = [torch.from_numpy(c).cuda() for c in f]
    g = NS3D(N, Re=Re, dt=dt)
    krylov.fd_jv_batch(g, u, V[:2], eps=eps, steps=1, f=fd)     # warm-up (buffers, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = krylov.fd_jv_batch(g, u, V, eps=eps, steps=steps, f=fd)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # one field at a time: the base and each perturbed field stepped alone
    # (the quotients are formed after the timed region, for the comparison)
    t0 = time.perf_counter()
    base = [c.clone() for c in u]
    g.RK3_SSP(base, fd, steps=steps)
    ws = []
    for i in range(m):
        w = [(u[k] + eps * V[i, k]).contiguous() for k in range(3)]
        g.RK3_SSP(w, fd, steps=steps)
        ws.append(w)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    same = all(torch.equal(J[i, k], (ws[i][k] - base[k]) / eps) for i in range(m) for k in range(3))
    del ws
    g.close()
    return {"workload": f"{m} finite-difference J.v (eps={eps:g}) of the T={T:g} map at {N}^3, Re={Re:g}, "
                        f"f_x = sin 4y, from the snapshot at t = {t_snap:.2f}, {steps} RK3 steps of dt={dt:.4g}; "
                        f"{m + 1} fields per batch",
            "jv_per_s_batched": m / t_batch, "jv_per_s_one_at_a_time": m / t_seq,
            "speedup": t_seq / t_batch, "batched_s": t_batch, "one_at_a_time_s": t_seq,
            "field_steps_per_s_batched": (m + 1) * steps / t_batch, "bitwise_same_as_alone": bool(same)}


def config_dt(u0, N, steps, fmax=1.0, cfl=0.2):
    """Fixed dt of a forced run: cfl * dx / (max|u0| + steps * dt * fmax) = dt
    (the velocity bound the forcing allows over the run keeps the CFL number
    <= cfl), rounded down to two significant digits."""
    umax = max(float(np.max(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))))) for c in u0)
    dx = 2 * math.pi / N
    if fmax == 0.0:
        dt = cfl * dx / umax
    else:
        dt = (-umax + math.sqrt(umax * umax + 4 * steps * fmax * cfl * dx)) / (2 * steps * fmax)
    e = math.floor(math.log10(dt))
    return math.floor(dt / 10 ** (e - 1)) * 10 ** (e - 1)


def run_config2(N=128, Re=100.0, steps=2000, seed=1, precisions=("fp32", "fp32i")):
    """Kolmogorov flow (f_x = sin 4y, Re = 100, the seeded k^4 exp(-2(k/4)^2)
    field of energy 0.5) at N^3 for `steps` RK3 steps in FP64 and in the FP32
    modes from identical bits; physical energy and dissipation (the spectral
    sums, every step) compared as series, and the final fields."""
    import torch
    from ns3dperiodic_amd import NS3D, fields
    u0 = fields.random_solenoidal(N, seed=seed, kind="k4exp", k0=4.0, energy=0.5)
    f = fields.kolmogorov_forcing_sin(N, n=4, amp=1.0)
    dt = config_dt(u0, N, steps)
    fd = [torch.from_numpy(c).cuda() for c in f]
    ser, fin, rate = {}, {}, {}
    for p in ("fp64",) + tuple(precisions):
        g = NS3D(N, Re=Re, dt=dt, precision=p, check_finite=True)
        u = [torch.from_numpy(c).cuda() for c in u0]
        E, D = np.empty(steps), np.empty(steps)
        t0 = time.perf_counter()
        for i in range(steps):
            g.RK3_SSP(u, fd)
            E[i] = g.kinetic_energy_spectral(u)
            D[i] = g.dissipation_spectral(u)
        rate[p] = steps / (time.perf_counter() - t0)
        ser[p], fin[p] = (E, D), u
        g.close()
    den = sum(float(torch.sum(torch.abs(y) ** 2)) for y in fin["fp64"])
    e64, d64 = ser["fp64"]
    out = {"workload": f"Kolmogorov flow {N}^3, f_x = sin 4y, Re = {Re:g}, seeded k^4 exp(-2(k/4)^2) field (energy 0.5), "
                       f"{steps} RK3 steps of dt = {dt:g} (T = {steps * dt:.3f}); energy and dissipation every step",
           "dt": dt, "energy_start_end": [float(e64[0]), float(e64[-1])],
           "dissipation_start_end": [float(d64[0]), float(d64[-1])],
           "steps_per_s_with_diagnostics": {names_of(p): r for p, r in rate.items()}}
    for p in precisions:
        e, d = ser[p]
        out[names_of(p)] = {
            "max_rel_diff_energy": float(np.max(np.abs(e - e64) / np.abs(e64))),
            "max_rel_diff_dissipation": float(np.max(np.abs(d - d64) / np.abs(d64))),
            "final_field_rel_l2": math.sqrt(sum(float(torch.sum(torch.abs(x - y) ** 2))
                                                for x, y in zip(fin[p], fin["fp64"])) / den)}
    del fin
    torch.cuda.empty_cache()
    return out


def kolmogorov_snapshot(N=64, Re=150.0, steps=2000, seed=1):
    """The state after `steps` RK3 steps of the Kolmogorov-flow recipe above
    at N^3 and Re (FP64): the base state of the batched Krylov work."""
    import torch
    from ns3dperiodic_amd import NS3D, fields
    u0 = fields.random_solenoidal(N, seed=seed, kind="k4exp", k0=4.0, energy=0.5)
    f = fields.kolmogorov_forcing_sin(N, n=4, amp=1.0)
    dt = config_dt(u0, N, steps)
    g = NS3D(N, Re=Re, dt=dt)
    u = [torch.from_numpy(c).cuda() for c in u0]
    g.RK3_SSP(u, [torch.from_numpy(c).cuda() for c in f], steps=steps)
    g.close()
    return [c.cpu().numpy() for c in u], f, dt * steps


# BiCGStab(L) of the Newton leg; 6 is the reference main's BiCG_L.
L_NEWTON = 6


def run_config4_krylov(N=64, Re=150.0, T=1.0, eps=1e-7, k=6, m=30, seed=1):
    """The Krylov legs on the 64^3 Re = 150 Kolmogorov snapshot:
    * Arnoldi (IRA, m = 30, 6 wanted, "LM", FP64 Gram-Schmidt with DGKS) on
      the finite-difference J.v of the T = 1 map, v -> (Phi(u + eps v) -
      Phi(u)) / eps with Phi(u) computed once: the leading Floquet multipliers;
    * Newton-Krylov for the steady Kolmogorov state at the same grid size
      (``_newton_laminar``), with an FP32-bulk and an FP64 Jacobian."""
    import torch
    from ns3dperiodic_amd import NS3D, fields, krylov
    us, f, t_snap = kolmogorov_snapshot(N, Re, seed=seed)
    umax = max(float(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))).max()) for c in us)
    steps = int(math.ceil(T / (0.2 * (2 * math.pi / N) / (umax + T))))
    dt = T / steps
    g = NS3D(N, Re=Re, dt=dt)
    fd = [torch.from_numpy(c).cuda() for c in f]
    u = [torch.from_numpy(c).cuda() for c in us]
    base = [c.clone() for c in u]
    g.RK3_SSP(base, fd, steps=steps)
    n = g.vector_length_reduced()
    ops = krylov.GpuOps(g)
    nmv = [0]

    def fdjv(vin, vout):
        v = g.unpack_reduced(vin)
        w = [(a + eps * b).contiguous() for a, b in zip(u, v)]
        g.RK3_SSP(w, fd, steps=steps)
        vout.copy_(g.pack_reduced([(a - b) / eps for a, b in zip(w, base)]))
        nmv[0] += 1
        return vout

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = krylov.eigs(ops, fdjv, n, k=k, m=m, which="LM", tol=1e-6, max_iter=60, residuals=False)
    t_arn = time.perf_counter() - t0
    g.close()
    # Newton-Krylov: F(x) = L(x) (return_RHS) in FP64, J.v - v/tau from an FP32-bulk or FP64 context
    newton = {}
    for jp in ("fp32", "fp64"):
        newton[f"jacobian_{jp}"] = _newton_laminar(N, RE_NEWTON, f, seed, jp)
    return {"workload": f"{N}^3 Kolmogorov snapshot (Re = {Re:g}, f_x = sin 4y, t = {t_snap:.2f} from the seeded field); "
                        f"J.v of the T = {T:g} map by forward differences (eps = {eps:g}, {steps} RK3 steps of "
                        f"dt = {dt:.4g})",
            "arnoldi": {"k": k, "m": m, "which": "LM", "restarts": r["iterations"], "ritz_norm": r["ritz_norm"],
                        "matvecs": nmv[0], "seconds": t_arn, "matvecs_per_s": nmv[0] / t_arn,
                        "multipliers_abs": [float(abs(z)) for z in r["values"]]},
            "newton": newton}


RE_NEWTON = 5.0


def _newton_laminar(N, Re, f, seed, jac_precision):
    """Damped Newton (alpha = 0.8) for the steady Kolmogorov state of f_x = sin 4y
    at N^3, from the laminar profile (Re/16) sin 4y plus a 1e-3 relative
    random perturbation of every reduced coordinate: residual F(x) = L(x) in
    FP64, inner solves BiCGStab(6) to 1e-6 (FP32 J) or 1e-10 (FP64 J) on the
    reference's globalised operator J.v - v/tau, tau = 5
    (NSCallMatrixVector_reduced_Newton_Globalization).  The laminar state is
    the exact solution, so the final distance from it measures the result.
    Re = 5 keeps the unpreconditioned inner solves convergent: at Re >= 20 on
    this grid BiCGStab(L) on that operator stagnates (advection outweighs
    diffusion at the forcing scale), in FP64 as in FP32."""
    import torch
    from ns3dperiodic_amd import NS3D, krylov
    fd = [torch.from_numpy(c).cuda() for c in f]
    ust = [(Re / 16.0) * c for c in f]          # laminar state: -|k|^2/Re u + f = 0
    g64 = NS3D(N, Re=Re, dt=1e-3)
    gj = g64 if jac_precision == "fp64" else NS3D(N, Re=Re, dt=1e-3, precision=jac_precision)
    op64 = krylov.NSOperators(g64, None, fd)
    opj = krylov.NSOperators(gj, None, fd, tau=5.0)
    n = g64.vector_length_reduced()
    rng = np.random.default_rng(seed)
    xs = g64.pack_reduced([torch.from_numpy(c).cuda() for c in ust])
    x = xs + 1e-3 * torch.from_numpy(rng.standard_normal(n)).cuda() * float(torch.linalg.vector_norm(xs)) / math.sqrt(n)
    F0 = op64.rhs(x, torch.empty_like(x))
    f0 = float(torch.linalg.vector_norm(F0))
    lin_tol = 1e-6 if jac_precision != "fp64" else 1e-10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nt = krylov.newton(op64.ops, op64.rhs, opj.jacobian, x, tol=1e-9 * f0, max_iter=30, L=L_NEWTON, lin_tol=lin_tol,
                       lin_iter=300)
    torch.cuda.synchronize()
    t_newton = time.perf_counter() - t0
    err = float(torch.linalg.vector_norm(x - xs) / torch.linalg.vector_norm(xs))
    g64.close()
    if gj is not g64:
        gj.close()
    return {"workload": f"{N}^3, Re = {Re:g}, steady Kolmogorov state from the laminar profile + 1e-3 relative "
                        f"random perturbation ({n} unknowns)",
            "jacobian": f"{jac_precision} J.v - v/tau (tau = 5), BiCGStab({L_NEWTON}) to {lin_tol:g}, at most 300 "
                        f"iterations", "residual": "FP64 F(x) = L(x)", "tol": "||F|| < 1e-9 ||F(x0)||",
            "converged": nt["converged"], "breakdown_flag": nt.get("breakdown"), "iterations": nt["iterations"],
            "residual_history": [float(h) for h in nt["history"]],
            "bicgstab_iterations": [int(a) for a, _ in nt.get("lin_history", [])],
            "rel_distance_from_laminar": err, "seconds": t_newton}


def run_accuracy_512(u512, Re, dt, steps, precisions=("fp32", "fp32s", "fp32i")):
    """Config 3 at 512^3: relative L2 distance of each FP32 mode from FP64
    after `steps` steps from identical bits (one precision resident at a time)."""
    return run_accuracy(512, steps, u512, Re, dt, precisions)


def pad_spectrum(u, M):
    """The same physical field on an M^3 grid (M > N): the N^3 half spectrum
    zero-padded and scaled by (M/N)^3 (the inverse transform carries 1/M^3);
    the N/2 Nyquist planes are dropped."""
    N = u[0].shape[0]
    h = N // 2
    ti = np.r_[0:h, M - h + 1:M]      # target index of wavenumbers 0..h-1 and -(h-1)..-1
    si = np.r_[0:h, h + 1:N]
    out = []
    for c in u:
        o = np.zeros((M, M, M // 2 + 1), np.complex128)
        o[np.ix_(ti, ti, np.arange(h))] = (M / N) ** 3 * c[np.ix_(si, si, np.arange(h))]
        out.append(o)
    return out


def copy_peak_gbs(reps=5, kinds=("copy", "scale")):
    """Bandwidth of a plain 1 GiB device-to-device copy (read + write bytes /
    time): torch's copy_ and an elementwise b = 1.0 * a; the faster is the
    reference rate."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    ops = {"copy": lambda: b.copy_(a), "scale": lambda: torch.mul(a, 1.0, out=b)}
    rates = {}
    for k in kinds:
        ops[k]()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ops[k]()
        e1.record()
        torch.cuda.synchronize()
        rates[k] = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    best = min(rates, key=rates.get)
    ms = rates[best]
    return {"bytes_read": n * 8, "bytes_written": n * 8, "avg_ms": ms, "GBps": 2 * n * 8 / (ms * 1e-3) / 1e9,
            "kernel": best, "GBps_by_kernel": {k: 2 * n * 8 / (v * 1e-3) / 1e9 for k, v in rates.items()}}


# the access pattern of each pass's memory traffic, as measured by pattern_ceiling:
# (pattern, fields read, fields written) of its nearest copy kernel
PASS_PATTERN = {"k_x_inv_grad": ("x_pencil", 3, 6), "k_y_inv_grad": ("y_pencil", 6, 6),
                "k_z_product": ("z_rows_truncated", 9, 3), "k_axis_c2c": ("y_pencil", 3, 3),
                "k_x_fwd_epi": ("x_pencil", 6, 3)}


def pattern_ceilings(N, precisions=("fp64", "fp32"), reps=20):
    """Copy-only kernels with the passes' tiles (ns3dperiodic_amd/pattern_ceiling,
    built with the library): the bandwidth each access pattern allows on this GPU.
    Returns the list of measured patterns, or None if the program is not built."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ns3dperiodic_amd", "pattern_ceiling")
    if not os.path.exists(exe):
        print(f"[bench] {exe} is not built (python -m ns3dperiodic_amd.build): no pattern ceilings",
              file=sys.stderr, flush=True)
        return None
    out = []
    for p in precisions:
        r = subprocess.run([exe, str(N), p, str(reps)], capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            raise RuntimeError(f"pattern_ceiling {N} {p} failed: {r.stderr.strip()}")
        out += [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    return out


def ceiling_of(ceilings, kernel, precision):
    """The measured ceiling of a pass's access pattern (bulk precision of its scratch)."""
    if not ceilings or kernel not in PASS_PATTERN:
        return None
    pat, nin, nout = PASS_PATTERN[kernel]
    bulk = "fp64" if precision == "fp64" else "fp32"
    for c in ceilings:
        if (c["pattern"], c["read_fields"], c["write_fields"], c["precision"]) == (pat, nin, nout, bulk):
            return c
    return None


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "allowed": allowed}


def cpu_baseline(N, u0, Re, dt, steps, threads):
    """The oracle (FP64 C restatement, OpenMP) on the same workload: one
    untimed warm-up step, then `steps` steps timed one by one; the median
    step time is the baseline.  Returns the summary and the oracle's state
    after the warm-up step and after all steps (for the parity check)."""
    from oracle.oracle import Oracle
    o = Oracle(N, Re=Re, dt=dt, nthreads=threads)
    zero = [np.zeros(o.cshape, np.complex128)] * 3
    u1 = o.rk3(u0, zero, steps=1)                      # warm-up (page-in, thread pool)
    u = u1
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        u = o.rk3(u, zero, steps=1)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    cpu = host_cpu()
    out = {"value": 1.0 / med, "unit": "RK3 steps/s", "cores": threads, "kind": "port",
           "sample": f"median of {steps} RK3 steps (after 1 warm-up step) of the same {N}^3 FP64 workload, "
                     f"oracle C restatement (OpenMP, own mixed-radix FFT, not FFTW), {sum(times):.1f} s timed",
           "median_s_per_step": med, "min_s_per_step": float(min(times)), "max_s_per_step": float(max(times)),
           "threads": threads, "host_cpu_model": cpu["model"], "host_nproc": cpu["nproc"],
           "host_cpus_allowed": cpu["allowed"]}
    return out, u1, u


def gpu_parity(N, u0, Re, dt, steps, u1_ref, un_ref):
    """The GPU FP64 path from the same initial bits as the oracle: relative L2
    distance after 1 step and after 1 + steps steps (the bar is 1e-12 per step)."""
    import torch
    from ns3dperiodic_amd import NS3D
    g = NS3D(N, Re=Re, dt=dt, precision="fp64")
    u = [torch.from_numpy(c).to("cuda:0") for c in u0]
    g.RK3_SSP(u, None, steps=1)
    torch.cuda.synchronize()

    def rel(a, b):
        num = sum(float(np.sum(np.abs(x.cpu().numpy() - y) ** 2)) for x, y in zip(a, b))
        return math.sqrt(num / sum(float(np.sum(np.abs(y) ** 2)) for y in b))
    e1 = rel(u, u1_ref)
    g.RK3_SSP(u, None, steps=steps)
    torch.cuda.synchronize()
    en = rel(u, un_ref)
    g.close()
    del u
    torch.cuda.empty_cache()
    return {"steps_compared": [1, 1 + steps], "rel_l2_after_1_step": e1, f"rel_l2_after_{1 + steps}_steps": en,
            "bar_per_step": 1e-12, "pass": bool(e1 <= 1e-12 and en <= 1e-12 * (1 + steps))}


def run_step_graph(K=20, reps=10, branches=12, seed=1):
    """Many small fields of uneven sizes advanced by one captured HIP graph
    (StepGraph: K RK3 steps of every field per replay, fields dealt to
    ``branches`` concurrent graph branches) against the same contexts stepped
    eagerly, K steps each, on one stream and on ``branches`` streams with the
    graph's assignment of fields to branches.  FP64, Kolmogorov forcing f_x = 0.1 sin 2y on every field."""
    import torch
    from ns3dperiodic_amd import NS3D, StepGraph, fields
    shapes = [(32, 32, 32)] * 8 + [(48, 64, 32)] * 4 + [(64, 64, 64)] * 4 + [(16, 16, 16)] * 8
    items, refs = [], []
    for i, sh in enumerate(shapes):
        u0 = fields.random_solenoidal(*sh, seed=seed + i, kind="k4exp", k0=4.0, energy=0.5)
        f = [torch.from_numpy(c).cuda() for c in fields.kolmogorov_forcing_sin(*sh, n=2, amp=0.1)]
        u = [torch.from_numpy(c).cuda() for c in u0]
        items.append((NS3D(*sh, Re=200.0, dt=2e-3), u, f))
        refs.append((NS3D(*sh, Re=200.0, dt=2e-3), [c.clone() for c in u], f))
    sg = StepGraph(items, steps=K, branches=branches)
    streams = [torch.cuda.Stream() for _ in range(branches)]
    # untimed warm-up: one replay, and the same K steps eagerly (states stay in step)
    sg.replay()
    for e, r, f in refs:
        e.set_stream(streams[0])
        e.RK3_SSP(r, f, steps=K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sg.replay()
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    t_eager = {}
    for ns in (1, branches):
        for i, (e, _, _) in enumerate(refs):            # the graph's branch assignment
            e.set_stream(streams[sg.branch_of[i] if ns > 1 else 0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for e, r, f in refs:
                e.RK3_SSP(r, f, steps=K)
        torch.cuda.synchronize()
        t_eager[ns] = time.perf_counter() - t0
    # the eager states took two timed legs, the graph's one: replay as many again
    for _ in range(reps):
        sg.replay()
    torch.cuda.synchronize()
    same = all(torch.equal(a, b) for (_, u, _), (_, r, _) in zip(items, refs) for a, b in zip(u, r))
    for g, _, _ in items + refs:
        g.close()
    n = len(shapes) * K * reps
    return {"workload": f"{len(shapes)} FP64 fields (8 x 16^3, 8 x 32^3, 4 x 48x64x32, 4 x 64^3), "
                        f"{K} RK3 steps per graph launch, {reps} launches, {branches} graph branches",
            "field_steps_per_s_graph": n / t_graph, "field_steps_per_s_eager_1_stream": n / t_eager[1],
            "field_steps_per_s_eager_branch_streams": n / t_eager[branches],
            "speedup_vs_eager_1_stream": t_eager[1] / t_graph,
            "speedup_vs_eager_branch_streams": t_eager[branches] / t_graph,
            "bitwise_same_as_eager": bool(same)}


def best_step_ms(N, precision, kind, members=8, steps=20):
    """ms of GPU time per request-step at N when the GPU does nothing else:
    integrate steps as a batch of `members` resident requests
    (ns3d_rk3_ssp_batch_members, graph-replayed), tangent-map steps (J.v
    requests, which run alone) and return-map steps (one RK3 step per loop turn
    with a probe) one request at a time."""
    import torch
    from ns3dperiodic_amd import NS3D, fields
    g = NS3D(N, precision=precision, stream=torch.cuda.Stream(), ipq=False, mask_nyquist=True)
    u0 = [torch.from_numpy(c).cuda() for c in fields.random_solenoidal(N, seed=3)]
    f = [torch.from_numpy(fields.kolmogorov_forcing_sin(N, n=4)[0]).cuda(), None, None]
    s = g.stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        if kind == "integrate":
            ub = torch.stack([torch.stack(u0)] * members).contiguous()
            dt = torch.full((members,), 1e-4, dtype=torch.float64, device="cuda")
            re = torch.full((members,), 200.0, dtype=torch.float64, device="cuda")
            left = torch.full((members,), steps, dtype=torch.int32, device="cuda")
            g.RK3_SSP_batch_members(ub, f, dt, re, left, steps, graph=True)
            e0.record(s)
            g.RK3_SSP_batch_members(ub, f, dt, re, left, steps, graph=True)
            e1.record(s)
            n = members * steps
        else:
            g.set_params(1e-4, 200.0)
            u = [c.clone() for c in u0]
            v = [1e-3 * c for c in u0]
            e0.record(s)
            for _ in range(steps):
                if kind == "jv":
                    g.RK3_SSP_UV(u, v, f)
                else:
                    g.RK3_SSP(u, f)
                    g.probe(u, N // 8, N // 4, N // 2)
            e1.record(s)
            n = steps
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    g.close()
    return ms


def run_serving(n_req, rate, seed, precision, max_active, chunk, rm_fraction=0.15):
    """Served mixed-size stream (Poisson arrivals, N in {32..192}, Re ~ U(50,500),
    T ~ U(0.5,5); a quarter J.v requests with T=1 and a share of return-map
    requests) through the continuous-batching server, against one-at-a-time FP64
    on the same requests."""
    import dataclasses
    import torch
    from ns3dperiodic_amd import serve
    t0 = time.time()
    reqs = serve.poisson_stream(n_req, rate, seed=seed, rm_fraction=rm_fraction)
    io = {r.rid: serve.pin_io(*serve.request_inputs(r)) for r in reqs}
    print(f"[serve] {n_req} requests generated in {time.time() - t0:.1f} s; total steps "
          f"{sum(r.nsteps for r in reqs)}", file=sys.stderr, flush=True)
    # warm-up: every (N, kind) once with 1-step horizons, for both paths (contexts, graphs, code objects)
    warm, wio = [], {}
    for i, (N, kind) in enumerate(sorted({(r.N, r.kind) for r in reqs})):
        src = next(r for r in reqs if (r.N, r.kind) == (N, kind))
        q = dataclasses.replace(src, rid=-1 - i, arrival=0.0, nsteps=chunk + 1, dt=1e-4)
        warm.append(q)
        wio[q.rid] = io[src.rid]
    srv = serve.Server(precision=precision, max_active=max_active, chunk=chunk)
    srv.run(warm, wio)
    ctxs = {}
    serve.one_at_a_time(warm, wio, precision="fp64", ctxs=ctxs)
    torch.cuda.synchronize()
    print("[serve] warm-up done", file=sys.stderr, flush=True)
    res = srv.run(reqs, io)
    served = {k: [c.copy() for c in r.u] for k, r in res.items()}
    s_srv = serve.summarize(list(res.values()))
    print(f"[serve] served {precision}: {s_srv}", file=sys.stderr, flush=True)
    base = serve.one_at_a_time(reqs, io, precision="fp64", ctxs=ctxs)
    s_base = serve.summarize(list(base.values()))
    print(f"[serve] one-at-a-time fp64: {s_base}", file=sys.stderr, flush=True)
    # per grid size: requests, steps and one-at-a-time FP64 service seconds (done - start)
    by_n = {}
    for r in reqs:
        d = by_n.setdefault(r.N, {"requests": 0, "steps": 0, "fp64_service_s": 0.0})
        d["requests"] += 1
        d["steps"] += r.nsteps
        d["fp64_service_s"] += base[r.rid].done - base[r.rid].admitted
    for d in by_n.values():
        d["fp64_ms_per_step"] = 1e3 * d["fp64_service_s"] / max(d["steps"], 1)
    errs, nonfinite = [], []
    kinds = {k: sum(r.kind == k for r in reqs) for k in ("integrate", "jv", "return_map")}
    rm_found = sum(1 for r in reqs if r.kind == "return_map" and res[r.rid].info and res[r.rid].info["found"])
    for r in reqs:
        if r.kind == "return_map":
            continue    # FP32 and FP64 may stop the search at different steps
        a, b = served[r.rid], base[r.rid].u
        if not all(np.isfinite(x).all() for x in a + b):
            nonfinite.append(r.rid)
            print(f"[serve] request {r.rid} ({r.kind} N={r.N} Re={r.Re:.0f} T={r.T:.2f} steps={r.nsteps}) "
                  f"is not finite", file=sys.stderr, flush=True)
            continue
        errs.append(math.sqrt(sum(float(np.sum(np.abs(x - y) ** 2)) for x, y in zip(a, b))
                              / sum(float(np.sum(np.abs(y) ** 2)) for y in b)))
    srv.close()
    # the same stream served in FP64 (bitwise the one-at-a-time FP64 results)
    srv64 = serve.Server(precision="fp64", max_active=max_active, chunk=chunk)
    srv64.run(warm, wio)
    torch.cuda.synchronize()
    res64 = srv64.run(reqs, io)
    s_srv64 = serve.summarize(list(res64.values()))
    same64 = all(all(np.array_equal(a, b) for a, b in zip(res64[r.rid].u, base[r.rid].u)) for r in reqs)
    print(f"[serve] served fp64: {s_srv64} (bitwise one-at-a-time: {same64})", file=sys.stderr, flush=True)
    srv64.close()
    for g in ctxs.values():
        g.close()
    # work-conserving bound: one-at-a-time FP64 service over the GPU time the same
    # steps take when the GPU does nothing else (integrate requests batched)
    fp64_service = sum(base[r.rid].done - base[r.rid].admitted for r in reqs)
    best = {}
    for p in sorted({precision, "fp64"}):
        t = {}
        for N, kind in sorted({(r.N, r.kind) for r in reqs}):
            t[(N, kind)] = best_step_ms(N, p, kind)
        best[p] = sum(r.nsteps * t[(r.N, r.kind)] for r in reqs) * 1e-3
    print(f"[serve] fp64 one-at-a-time service {fp64_service:.1f} s; best GPU time {best}", file=sys.stderr, flush=True)
    return {"requests": n_req, "rate_per_s": rate, "seed": seed, "precision": precision,
            "batching": "integrate requests of one grid share one launch per pass (ns3d_rk3_ssp_batch_members, "
                        "per-member dt/Re); J.v and return-map requests run on slots of their own",
            "served_fp64": {"solves_per_s": s_srv64["solves_per_s"], "p50_latency_s": s_srv64["p50_latency_s"],
                            "p99_latency_s": s_srv64["p99_latency_s"], "span_s": s_srv64["span_s"],
                            "speedup_solves_per_s": s_srv64["solves_per_s"] / s_base["solves_per_s"],
                            "bitwise_same_as_one_at_a_time_fp64": bool(same64)},
            "work_conserving_bound": {
                "fp64_one_at_a_time_service_s": fp64_service,
                "best_gpu_s": best,
                "bound_speedup": {names_of(p): fp64_service / b for p, b in best.items()},
                "definition": "sum of one-at-a-time FP64 service times / sum over requests of steps x the GPU time "
                              "per step of its grid and kind when the GPU runs nothing else (integrate steps in a "
                              "batch of 8)"},
            "max_active": max_active, "graph_chunk_steps": chunk,
            "total_steps": sum(r.nsteps for r in reqs),
            "solves_per_s": s_srv["solves_per_s"], "p50_latency_s": s_srv["p50_latency_s"],
            "p99_latency_s": s_srv["p99_latency_s"], "span_s": s_srv["span_s"],
            "hbm_peak_bytes": srv.ledger.peak,
            "baseline_fp64_one_at_a_time": s_base,
            "speedup_solves_per_s": s_srv["solves_per_s"] / s_base["solves_per_s"],
            "max_rel_l2_vs_fp64": max(errs) if errs else None,
            "median_rel_l2_vs_fp64": float(np.median(errs)) if errs else None, "nonfinite": nonfinite,
            "by_grid": {str(k): v for k, v in sorted(by_n.items())},
            "kinds": kinds, "return_map_crossings_found": rm_found}


def names_of(precision):
    return {"fp64": "fp64", "fp32": "fp32_bulk", "fp32s": "fp32_state", "fp32i": "fp32_incr"}[precision]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32", "fp32s", "fp32i"], default="fp64")
    ap.add_argument("--secondary", default="fp32,fp32s,fp32i",
                    help="comma list of further precisions to time beside the primary (fp64, fp32, fp32s, fp32i) "
                         "or none")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="oracle steps timed one by one after a warm-up step (the CPU baseline is their median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests in the served mixed-size stream (0: skip the serving leg)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--serve-active", type=int, default=8, help="resident request slots")
    ap.add_argument("--serve-chunk", type=int, default=16, help="RK3 steps per captured graph")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (3 times): the reference kernel for HBM counters")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="RK3 steps after which the FP32 modes are compared with FP64 (0: skip)")
    ap.add_argument("--accuracy-steps-512", type=int, default=1000,
                    help="steps of the 512^3 FP32-vs-FP64 accuracy leg (0: skip)")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the main timing: skip the FP32 accuracy, 512^3, passive-scalar, tangent-map, "
                         "batched J.v and copy-peak legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        c = copy_peak_gbs(reps=3, kinds=("copy",))
        print(json.dumps({"copy_bytes_read": c["bytes_read"], "copy_bytes_written": c["bytes_written"],
                          "avg_ms": c["avg_ms"], "GBps": c["GBps"]}))
        return

    from ns3dperiodic_amd import fields

    N = args.N
    Re = 500.0                      # nu = 2e-3
    u0 = fields.random_solenoidal(N, seed=args.seed, kind="k4exp", k0=4.0, energy=0.5)
    umax = max(float(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))).max()) for c in u0)
    dx = 2 * math.pi / N
    dt = float("%.1e" % (0.2 * dx / umax))  # fixed dt = 0.2 dx / max|u0|, rounded
    print(f"[bench] N={N} Re={Re} dt={dt:g} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)

    prim = run_gpu(N, args.precision, args.steps, args.warmup, u0, Re, dt)
    print(f"[bench] {args.precision}: {prim['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    secs = []
    for p in [] if args.secondary == "none" else args.secondary.split(","):
        if p == args.precision:
            continue
        secs.append(run_gpu(N, p, args.steps, args.warmup, u0, Re, dt))
        print(f"[bench] {p}: {secs[-1]['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)

    accuracy = None
    if args.accuracy_steps > 0 and not args.no_extra:
        accuracy = run_accuracy(N, args.accuracy_steps, u0, Re, dt)
        print(f"[bench] rel L2 vs fp64 after {args.accuracy_steps} steps: "
              f"{ {k: v for k, v in accuracy.items() if k != 'steps'} }", file=sys.stderr, flush=True)

    extra = None
    if not args.no_extra:
        extra = {"copy_peak": copy_peak_gbs()}
        print(f"[bench] copy peak {extra['copy_peak']['GBps']:.0f} GB/s", file=sys.stderr, flush=True)
        th0 = fields.random_scalar(N, seed=args.seed)
        sc = run_scalar(N, max(args.steps // 4, 5), 2, u0, th0, Re, dt)
        extra["scalar"] = dict(sc, workload=f"{N}^3 FP64 velocity + WENO5 passive scalar "
                                            "(E_theta ~ k^2 exp(-(k/4)^2), unit variance)")
        print(f"[bench] fp64 + scalar: {sc['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
        extra["tangent"] = [dict(run_tangent(N, p, 10, u0, Re, dt), vs_plain_step=None) for p in ("fp64", "fp32")]
        for t in extra["tangent"]:
            plain = prim if t["precision"] == args.precision else next(
                (x for x in secs if x["precision"] == t["precision"]), None)
            t["vs_plain_step"] = t["ms_per_step"] / plain["ms_per_step"] if plain else None
            print(f"[bench] tangent map {t['precision']}: {t['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
        extra["config2"] = run_config2(seed=args.seed)
        print(f"[bench] config 2 (128^3 Kolmogorov, 2000 steps): "
              f"{ {k: v for k, v in extra['config2'].items() if k.startswith('fp32')} }", file=sys.stderr, flush=True)
        extra["config4"] = run_config4_krylov(seed=args.seed)
        print(f"[bench] config 4 Krylov: arnoldi {extra['config4']['arnoldi']['matvecs']} matvecs in "
              f"{extra['config4']['arnoldi']['seconds']:.1f} s, newton "
              f"{ {k: (v['converged'], v['iterations'], round(v['seconds'], 2)) for k, v in extra['config4']['newton'].items()} }",
              file=sys.stderr, flush=True)
        extra["fd_jv_batch"] = run_fd_jv_batch(seed=args.seed)
        print(f"[bench] batched FD J.v: {extra['fd_jv_batch']['jv_per_s_batched']:.1f} J.v/s "
              f"({extra['fd_jv_batch']['speedup']:.2f}x one at a time)", file=sys.stderr, flush=True)
        extra["step_graph"] = run_step_graph(seed=args.seed)
        print(f"[bench] uneven fields in one graph: {extra['step_graph']['field_steps_per_s_graph']:.0f} field-steps/s "
              f"({extra['step_graph']['speedup_vs_eager_1_stream']:.2f}x eager on one stream, "
              f"{extra['step_graph']['speedup_vs_eager_branch_streams']:.2f}x on as many streams as branches)", file=sys.stderr, flush=True)
        if N == 256:
            M = 512
            u512 = pad_spectrum(u0, M)
            dt512 = float("%.1e" % (dt * N / M))
            big = {"workload": f"the {N}^3 field zero-padded to {M}^3, same spectrum, dt = {dt512:g}", "dt": dt512}
            for p in ["fp64", "fp32", "fp32s", "fp32i"]:
                r = run_gpu(M, p, 5, 1, u512, Re, dt512)
                big[names_of(p)] = {"steps_per_s": r["steps_per_s"], "ms_per_step": r["ms_per_step"],
                                    "step_alg_bytes": r["step_alg_bytes"],
                                    "step_roofline_frac": r["step_GBps"] / HBM_PEAK_GBS, "finite": r["finite"],
                                    "passes_avg_ms": {q["pass"]: q["avg_ms"] for q in r["passes"]}}
                print(f"[bench] {M}^3 {p}: {r['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
            if args.accuracy_steps_512 > 0:
                big["fp32_accuracy"] = run_accuracy_512(u512, Re, dt512, args.accuracy_steps_512)
                print(f"[bench] {M}^3 rel L2 vs fp64 after {args.accuracy_steps_512} steps: "
                      f"{big['fp32_accuracy']}", file=sys.stderr, flush=True)
            del u512
            extra["n512"] = big

    dom = max(prim["passes"], key=lambda p: p["total_ms"])
    traffic = profiled_traffic(args.precision, dom["kernel"]) if N == 256 else None
    roofline = {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": "profiles/rk3_256_passes_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                                  "scaled by a plain-copy calibration)" if traffic else None,
                "alg_bytes_per_launch": dom["alg_bytes"], "avg_launch_ms": dom["avg_ms"]}
    ceilings = None if args.no_extra else pattern_ceilings(N)
    if ceilings:
        print(f"[bench] pattern ceilings: { {(c['pattern'], c['precision'], c['read_fields'], c['write_fields']): round(c['GBps']) for c in ceilings} }",
              file=sys.stderr, flush=True)
        ce = ceiling_of(ceilings, dom["kernel"], args.precision)
        if ce:
            actual = traffic / (dom["avg_ms"] * 1e-3) / 1e9 if traffic else None
            roofline["pattern_ceiling"] = {
                "pattern": f"{ce['pattern']} {ce['read_fields']}->{ce['write_fields']} ({ce['precision']})",
                "GBps": ce["GBps"], "actual_traffic_GBps": actual,
                "frac": actual / ce["GBps"] if actual else None,
                "source": "ns3dperiodic_amd/pattern_ceiling: copy-only kernel with the pass's tiles, same run"}
        for q in prim["passes"]:
            c = ceiling_of(ceilings, q["kernel"], args.precision)
            q["pattern_ceiling_GBps"] = c["GBps"] if c else None

    cpu = parity = None
    if not args.no_cpu:
        # every core this process may use, unless OMP_NUM_THREADS sets the host's share
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host_cpu()["allowed"]
        cpu, u1_ref, un_ref = cpu_baseline(N, u0, Re, dt, args.cpu_steps, threads)
        print(f"[bench] cpu: {cpu['value']:.4f} steps/s ({threads} threads)", file=sys.stderr, flush=True)
        parity = gpu_parity(N, u0, Re, dt, args.cpu_steps, u1_ref, un_ref)
        del u1_ref, un_ref
        print(f"[bench] parity vs oracle: {parity}", file=sys.stderr, flush=True)

    serving = None
    if args.serve_requests > 0:
        serving = run_serving(args.serve_requests, args.serve_rate, args.seed, args.serve_precision,
                              args.serve_active, args.serve_chunk)

    out = {
        "metric": "RK3 steps/s at 256^3 (FP64; FP32-bulk beside it), % of HBM peak",
        "value": prim["steps_per_s"],
        "unit": "RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": prim["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 bulk (f64 state)", "fp32s": "f32 (f32 state)",
                  "fp32i": "f32 bulk and stage increments (f64 state)"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"decaying turbulence {N}^3, nu=2e-3, k0=4 spectrum, reference mode",
                   "N": N, "Re": Re, "dt": dt, "global_batch": 1, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "step_roofline": {"alg_bytes": prim["step_alg_bytes"], "achieved": prim["step_GBps"],
                          "frac": prim["step_GBps"] / HBM_PEAK_GBS},
        "passes": prim["passes"],
        "finite": prim["finite"],
        "serving": serving,
        "fp32_accuracy": accuracy,
        "pattern_ceilings": ceilings,
    }
    if extra is not None:
        cp = extra["copy_peak"]["GBps"]
        out["copy_peak"] = dict(extra["copy_peak"], step_frac_of_copy=prim["step_GBps"] / cp,
                                dominant_frac_of_copy=dom["GBps"] / cp)
        out["scalar"] = extra["scalar"]
        out["fd_jv_batch"] = extra["fd_jv_batch"]
        out["config2"] = extra["config2"]
        out["config4_krylov"] = extra["config4"]
        out["step_graph"] = extra["step_graph"]
        out["tangent_map"] = {"workload": f"{N}^3 RK3_SSP_UV steps (base + one perturbation)",
                              "legs": extra["tangent"]}
        if "n512" in extra:
            out["n512"] = extra["n512"]
    for sec in secs:
        out[names_of(sec["precision"])] = {
            "steps_per_s": sec["steps_per_s"], "ms_per_step": sec["ms_per_step"],
            "speedup_vs_primary": sec["steps_per_s"] / prim["steps_per_s"],
            "step_roofline_frac": sec["step_GBps"] / HBM_PEAK_GBS,
            "passes": sec["passes"], "finite": sec["finite"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
