"""Autotuner host logic: the variant list (baseline plus one change per
candidate) and the selection rule (a value replaces the baseline only when
the summed step time of the precisions it affects drops by more than the
margin), on synthetic timings."""
import argparse
import json

from ns3dperiodic_amd import autotune


def test_variants_cover_every_candidate_once():
    base, vs = autotune.variants()
    names = [n for n, _ in vs]
    assert names[0] == "base" and len(set(names)) == len(names)
    n_alt = sum(len([c for c in cands if c != base[k]]) for k, (_, cands, _, _) in autotune.SPACE.items())
    assert len(vs) == 1 + n_alt
    for n, o in vs[1:]:
        (k, v), = o.items()
        assert n == f"{k}={v}" and v != base[k]


def test_select_rule(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "TUNED", str(tmp_path / "tuned.hpp"))
    base, vs = autotune.variants()
    plan = {"baseline": base, "variants": [{"name": n, "overrides": o, "lib": f"lib_{i}.so"}
                                           for i, (n, o) in enumerate(vs)]}
    json.dump(plan, open(tmp_path / "plan.json", "w"))
    ms = {}
    sizes = {"pow2": 256, "r3": 192}
    for N in sizes.values():
        for p, t in (("fp64", 4.70), ("fp32", 3.06), ("fp32s", 2.75), ("fp32i", 2.90)):
            ms[f"base@{N}|{p}"] = t
    for n, o in vs[1:]:
        (k, v), = o.items()
        N = sizes[autotune.SPACE[k][3]]
        for p in autotune.SPACE[k][2]:
            ms[f"{n}@{N}|{p}"] = ms[f"base@{N}|{p}"] * 1.02          # every change slower ...
    alt = next(c for c in autotune.SPACE["NS3D_PZT64"][1] if c != base["NS3D_PZT64"])
    ms[f"NS3D_PZT64={alt}@256|fp64"] = 4.60                          # ... but one (2% faster: taken)
    alt32 = next(c for c in autotune.SPACE["NS3D_PXFWD32"][1] if c != base["NS3D_PXFWD32"])
    ms[f"NS3D_PXFWD32={alt32}@256|fp32"] = 3.05                      # 0.3% on the fp32+fp32s+fp32i sum: kept out
    ms[f"NS3D_PXFWD32={alt32}@256|fp32s"] = 2.74
    ms[f"NS3D_PXFWD32={alt32}@256|fp32i"] = 2.89
    alt3 = next(c for c in autotune.SPACE["NS3D_R3_PYINV64"][1] if c != base["NS3D_R3_PYINV64"])
    ms[f"NS3D_R3_PYINV64={alt3}@192|fp64"] = 4.50                    # radix-3 group, timed at 192^3
    json.dump({"N": 256, "N3": 192, "dt": {"pow2": 1.9e-3, "r3": 2.5e-3}, "steps": 20, "gpu": "test",
               "ms_per_step": ms}, open(tmp_path / "times.json", "w"))
    autotune.cmd_select(argparse.Namespace(out=str(tmp_path), margin=0.005))
    got = autotune.read_tuned(str(tmp_path / "tuned.hpp"))
    assert got["NS3D_PZT64"] == alt
    assert got["NS3D_R3_PYINV64"] == alt3
    assert got["NS3D_PXFWD32"] == base["NS3D_PXFWD32"]
    changed = ("NS3D_PZT64", "NS3D_R3_PYINV64")
    assert {k: v for k, v in got.items() if k not in changed} == {k: v for k, v in base.items() if k not in changed}
