"""The access-pattern bandwidth benchmark that bench.py reports beside the
roofline (ns3dperiodic_amd/pattern_ceiling, built with the library): it runs,
prints one JSON line per pattern and measures plausible rates."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ns3dperiodic_amd", "pattern_ceiling")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pattern_ceiling_runs(precision):
    print(f"\n[case] pattern_ceiling 128 {precision}", flush=True)
    assert os.path.exists(EXE), "build the package first (python -m ns3dperiodic_amd.build)"
    r = subprocess.run([EXE, "128", precision, "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    pats = {(x["pattern"], x["read_fields"], x["write_fields"]) for x in rows}
    assert {("flat", 3, 3), ("x_pencil", 6, 3), ("x_pencil", 3, 6), ("y_pencil", 3, 3), ("y_pencil", 6, 6),
            ("z_rows_truncated", 9, 3)} <= pats
    mz = 128 // 2 + 1
    esz = 16 if precision == "fp64" else 8
    for x in rows:
        assert x["precision"] == precision and x["N"] == 128
        assert x["bytes"] > 0 and x["ms"] > 0 and 50 < x["GBps"] < 20000
    flat = next(x for x in rows if x["pattern"] == "flat")
    assert flat["bytes"] == 6 * 128 * 128 * mz * esz


def test_pattern_ceiling_rejects_bad_arguments():
    r = subprocess.run([EXE, "100", "fp64"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    r = subprocess.run([EXE, "128", "fp16"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
