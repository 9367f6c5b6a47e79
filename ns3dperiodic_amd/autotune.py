"""Autotuner for the tile shapes and FFT plans of the fused RK3 passes.

The kernels take their per-pass shapes as compile-time constants (points per
thread E, pencils per block P, z-pass plan and block size, LDS swizzle, two
transforms in step), so
tuning is a build-and-time search over library variants:

    python -m ns3dperiodic_amd.autotune build [--out DIR] [-j 6]   # CPU: one variant per candidate
    python -m ns3dperiodic_amd.autotune time  [--out DIR] [--N 256] [--N3 192]   # GPU: time every variant
    python -m ns3dperiodic_amd.autotune select [--out DIR]            # write csrc/tuned.hpp
    python -m ns3dperiodic_amd.build                                  # rebuild with the choice

``build`` compiles the baseline (the current tuned.hpp, or the source
defaults) and, for every parameter, one variant per alternative value with
only that parameter changed.  ``time`` runs each variant's RK3 step on the
GPU (HIP events, steady state) in the precisions the parameter affects and
records ms/step: parameters of the power-of-two lengths at N (256), those of
the radix-3 lengths from 192 up (NS3D_R3_*) at N3 (192).  ``select`` keeps, for each parameter, the value with the
lowest time when it beats the baseline by more than ``--margin`` (parameters
of different passes act on different kernels, so their gains add), and
writes ``csrc/tuned.hpp``, which ns3d_api.hip includes before the defaults.
Variant libraries and timings stay in DIR (default ns3dperiodic_amd/_tune).
"""
import argparse
import json
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "csrc", "tuned.hpp")
DEFAULT_OUT = os.path.join(HERE, "_tune")

# parameter -> (source default, candidates, precisions whose step it changes,
#               lengths it applies to: "pow2" timed at N, "r3" (radix-3, >= 192) at N3)
SPACE = {
    "NS3D_EXINV64": (4, [4, 8], ["fp64"], "pow2"),
    "NS3D_EYINV64": (4, [4, 8], ["fp64"], "pow2"),
    "NS3D_EXFWD64": (8, [4, 8], ["fp64"], "pow2"),
    "NS3D_PXINV64": (8, [4, 8], ["fp64"], "pow2"),
    "NS3D_PYINV64": (8, [4, 8], ["fp64"], "pow2"),
    "NS3D_PZT64": (128, [64, 128, 256], ["fp64"], "pow2"),
    "NS3D_ZSWZ": (8, [0, 8, 16], ["fp64"], "pow2"),
    "NS3D_ZILP": (1, [0, 1], ["fp64"], "pow2"),
    "NS3D_XFWD_ILP": (2, [0, 1, 2], ["fp64", "fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_EXINV32": (8, [4, 8], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_EYINV32": (4, [4, 8], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_PXINV32": (8, [4, 8, 16], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_PYINV32": (8, [4, 8, 16], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_PYFWD32": (16, [8, 16], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_PXFWD32": (8, [4, 8, 16], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_PZT32": (128, [64, 128, 256], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_ZPLAN32": (8, [4, 8], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_ZSWZ32": (16, [0, 16], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_EXFWD32": (4, [4, 8], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_ZILP32": (0, [0, 1], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_ZWAVE": (1, [0, 1], ["fp64", "fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_ZTWREG": (1, [0, 1], ["fp64"], "pow2"),
    "NS3D_ZTWREG32": (0, [0, 1], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_WAVES_Z32": (2, [2, 3], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_WAVES_XFWD32": (2, [2, 3, 4], ["fp32", "fp32s", "fp32i"], "pow2"),
    "NS3D_R3_PXINV64": (4, [4, 8], ["fp64"], "r3"),
    "NS3D_R3_PYINV64": (4, [4, 8], ["fp64"], "r3"),
    "NS3D_R3_PZT64": (256, [128, 256], ["fp64"], "r3"),
    "NS3D_R3_ZSWZ": (0, [0, 16], ["fp64"], "r3"),
    "NS3D_R3_PXFWD32": (16, [8, 16], ["fp32", "fp32s", "fp32i"], "r3"),
    "NS3D_R3_PYINV32": (4, [4, 8], ["fp32", "fp32s", "fp32i"], "r3"),
}


def read_tuned(path=TUNED):
    """Values set by an existing tuned.hpp (name -> int)."""
    vals = {}
    if os.path.exists(path):
        for line in open(path):
            p = line.split()
            if len(p) >= 3 and p[0] == "#define" and p[1] in SPACE:
                vals[p[1]] = int(p[2])
    return vals


def baseline():
    cur = {k: v[0] for k, v in SPACE.items()}
    cur.update(read_tuned())
    return cur


def variants(only=None):
    """(name, overrides) for the baseline and every single-parameter change
    (of the parameters in `only`, if given)."""
    base = baseline()
    out = [("base", {})]
    for k, (_, cands, _, _) in SPACE.items():
        if only and k not in only:
            continue
        for c in cands:
            if c != base[k]:
                out.append((f"{k}={c}", {k: c}))
    return base, out


def _build_one(out_dir, name, overrides, base):
    from . import build as B
    vals = dict(base)
    vals.update(overrides)
    flags = [f"-D{k}={v}" for k, v in sorted(vals.items())]
    lib = os.path.join(out_dir, f"lib_{name.replace('=', '_')}.so")
    src = os.path.join(HERE, "csrc", "ns3d_api.hip")
    cmd = [B.hipcc(), *B.HIPCC_FLAGS, "-DNS3D_NO_TUNED_HEADER", *flags, src, "-o", lib]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"variant {name} failed to build:\n{r.stderr[-2000:]}")
    return name, lib, time.time() - t0


def cmd_build(args):
    os.makedirs(args.out, exist_ok=True)
    base, vs = variants(args.only.split(",") if args.only else None)
    print(f"[autotune] baseline {base}", flush=True)
    print(f"[autotune] building {len(vs)} variants with {args.jobs} jobs", flush=True)
    plan = []
    with ThreadPoolExecutor(args.jobs) as ex:
        futs = [ex.submit(_build_one, args.out, n, o, base) for n, o in vs]
        for (n, o), f in zip(vs, futs):
            name, lib, dt = f.result()
            print(f"[autotune] built {name} ({dt:.0f} s)", flush=True)
            plan.append({"name": name, "overrides": o, "lib": os.path.basename(lib)})
    json.dump({"baseline": base, "variants": plan}, open(os.path.join(args.out, "plan.json"), "w"), indent=1)


def _time_lib(lib, N, precision, steps, warmup, u0, Re, dt):
    import torch
    from . import _lib as L
    from .solver import NS3D
    L._lib = None
    L.LIB_PATH = lib
    g = NS3D(N, Re=Re, dt=dt, precision=precision)
    u = [torch.from_numpy(c).to("cuda:0") for c in u0]
    g.RK3_SSP(u, None, steps=warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    e0.record(s)
    g.RK3_SSP(u, None, steps=steps)
    e1.record(s)
    torch.cuda.synchronize()
    ok = bool(torch.isfinite(u[0]).all().item())
    g.close()
    return e0.elapsed_time(e1) / steps, ok


def cmd_time(args):
    import numpy as np
    from . import fields
    plan = json.load(open(os.path.join(args.out, "plan.json")))
    sizes = {"pow2": args.N, "r3": args.N3}
    work = {}
    for cls, N in sizes.items():
        u0 = fields.random_solenoidal(N, seed=1, kind="k4exp", k0=4.0, energy=0.5)
        umax = max(float(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))).max()) for c in u0)
        work[cls] = (N, u0, float("%.1e" % (0.2 * (2 * np.pi / N) / umax)))
    res = {}
    for rep in range(args.reps):
        for v in plan["variants"]:
            if v["overrides"]:
                _, _, precs, cls = SPACE[next(iter(v["overrides"]))]
                runs = [(cls, precs)]
            else:   # the baseline at both sizes (only the classes the plan's variants need)
                need = {SPACE[next(iter(x["overrides"]))][3] for x in plan["variants"] if x["overrides"]}
                runs = [(cls, ["fp64", "fp32", "fp32s", "fp32i"]) for cls in sizes if cls in need or not need]
            for cls, precs in runs:
                N, u0, dt = work[cls]
                for p in precs:
                    ms, ok = _time_lib(os.path.join(args.out, v["lib"]), N, p, args.steps, 3, u0, 500.0, dt)
                    res.setdefault(f"{v['name']}@{N}|{p}", []).append(ms if ok else float("inf"))
                    print(f"[autotune] {v['name']:20s} N={N:3d} {p:6s} {ms:.4f} ms/step{'' if ok else ' NON-FINITE'}",
                          flush=True)
    import torch
    props = torch.cuda.get_device_properties(0)
    out = {"N": args.N, "N3": args.N3, "dt": {cls: w[2] for cls, w in work.items()}, "steps": args.steps,
           "gpu": torch.cuda.get_device_name(0), "arch": getattr(props, "gcnArchName", "").split(":")[0],
           "ms_per_step": {k: min(x) for k, x in res.items()}}
    json.dump(out, open(os.path.join(args.out, "times.json"), "w"), indent=1)


def cmd_select(args):
    plan = json.load(open(os.path.join(args.out, "plan.json")))
    t = json.load(open(os.path.join(args.out, "times.json")))
    ms = t["ms_per_step"]
    base = plan["baseline"]
    chosen, notes = dict(base), []
    planned = {next(iter(v["overrides"])) for v in plan["variants"] if v["overrides"]}
    for k, (_, cands, precs, cls) in SPACE.items():
        N = t["N"] if cls == "pow2" else t["N3"]
        if k not in planned:        # not timed in this run: the current value stays
            notes.append(f"{k} = {base[k]} (not timed in the last run)")
            continue

        def cost(name):
            return sum(ms[f"{name}@{N}|{p}"] for p in precs)
        best, bcost = base[k], cost("base")
        for c in cands:
            if c == base[k]:
                continue
            cc = cost(f"{k}={c}")
            if cc < bcost * (1.0 - args.margin):
                best, bcost = c, cc
        chosen[k] = best
        notes.append(f"{k} = {best} ({'changed' if best != base[k] else 'kept'}; "
                     f"{', '.join(precs)} step at {N}^3 {bcost:.4f} ms summed)")
    lines = ["// tuned.hpp -- written by ns3dperiodic_amd/autotune.py (select); do not edit by hand.",
             f"// Timed on {t.get('gpu', 'the GPU')} ({t.get('arch') or 'gfx950'}): RK3 steps at {t['N']}^3 "
             f"(power-of-two parameters) and {t['N3']}^3 (NS3D_R3_*); "
             f"a value replaces the baseline only when its step is faster by more than {args.margin:.1%}.",
             "#pragma once"]
    for n in notes:
        lines.append("// " + n)
    for k in SPACE:
        lines.append(f"#define {k} {chosen[k]}")
    open(TUNED, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser(prog="python -m ns3dperiodic_amd.autotune")
    ap.add_argument("phase", choices=["build", "time", "select"])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("-j", "--jobs", type=int, default=min(6, os.cpu_count() or 1))
    ap.add_argument("--N", type=int, default=256, help="grid for the power-of-two parameters")
    ap.add_argument("--N3", type=int, default=192, help="grid for the radix-3 parameters (NS3D_R3_*)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--margin", type=float, default=0.005)
    ap.add_argument("--only", default="", help="comma list of parameters to vary (build); default all")
    args = ap.parse_args()
    {"build": cmd_build, "time": cmd_time, "select": cmd_select}[args.phase](args)


if __name__ == "__main__":
    sys.exit(main())
