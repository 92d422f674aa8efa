"""The oracle's outcomes on the reference's own L-BFGS vectors under
FP32-sized evaluation noise (tests/golden/lbfgs_branch_outcomes.json, written
by tests/golden/make_lbfgs_branch_outcomes.py), which the device L-BFGS
tests, the C++ host test and smoke() accept besides the gtest's own result.

The fixture is recomputed here (the noise streams are seeded, so the
computation is deterministic), and its most frequent outcome of each vector
must be the gtest's expected final state: the alternatives are the same
algorithm taking the other branch of its line search, not a looser bound.
"""
import json
import os
import sys

import numpy as np

from conftest import GOLDEN, ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def test_branch_outcomes_are_reproduced_and_contain_the_gtest_result():
    import make_lbfgs_branch_outcomes as M
    with open(os.path.join(GOLDEN, "lbfgs_branch_outcomes.json")) as fh:
        ref = json.load(fh)
    with open(os.path.join(GOLDEN, "gtest_vectors.json")) as fh:
        gt = json.load(fh)["lbfgs"]["cases"]
    got = M.measure()
    assert len(got) == len(ref["cases"]) == len(gt)
    for g, r, c in zip(got, ref["cases"], gt):
        assert [o["count"] for o in g["outcomes"]] == [o["count"] for o in r["outcomes"]]
        assert sum(o["count"] for o in r["outcomes"]) == ref["seeds"]
        for a, b in zip(g["outcomes"], r["outcomes"]):
            np.testing.assert_allclose(a["energy"], b["energy"], rtol=1e-12)
            np.testing.assert_allclose(a["out_torsions"], b["out_torsions"], rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(a["out_xyz"], b["out_xyz"], rtol=1e-12, atol=1e-12)
        top = r["outcomes"][0]
        assert abs(top["energy"] - c["energy"]) <= 1e-4 * abs(c["energy"])
        np.testing.assert_allclose(top["out_torsions"], c["out_torsions"], atol=1e-3)
        np.testing.assert_allclose(top["out_xyz"], c["out_xyz"], atol=1e-3)
        # every alternative is a distinct final state of the same run length
        es = [o["energy"] for o in r["outcomes"]]
        assert all(abs(x - y) > 1e-4 * abs(y) for i, x in enumerate(es) for y in es[:i])
