"""The oracle against the committed golden fixtures (independent numpy restatement)."""
import glob
import os

import numpy as np
import pytest

from tests.conftest import rel_l2

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "golden_*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_matches_golden(oracle_mod, path):
    d = np.load(path)  # allow_pickle stays False
    shape = tuple(int(v) for v in d["shape"])
    o = oracle_mod.Oracle(*shape, Re=float(d["Re"]), dt=float(d["dt"]))
    u, v, f = list(d["u"]), list(d["v"]), list(d["f"])
    np.testing.assert_array_equal(o.tables()["mask"], d["mask"])
    assert rel_l2(o.nonlinear(u), list(d["Q"])) < 1e-13
    assert rel_l2(o.bilinear(u, v), list(d["B"])) < 1e-13
    assert rel_l2(o.single_time_step(u, f), list(d["stage"])) < 1e-13
    assert rel_l2(o.rk3(u, f), list(d["rk3"])) < 1e-13
    assert rel_l2(o.rk3_uv_rhs(u, v), list(d["jv"])) < 1e-13
    assert rel_l2(o.divergence(u), d["div"]) < 1e-14


def test_fixtures_present():
    assert len(FIXTURES) >= 2
