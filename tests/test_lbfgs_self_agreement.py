"""The FP64 oracle's L-BFGS agreement with itself under evaluation noise,
against the committed rates (tests/golden/lbfgs_self_agreement.json, written
by tests/golden/make_lbfgs_self_agreement.py).

These rates are the yardstick of the device L-BFGS parity tests: at the
reference full step, runs from the same start perturbed by one part in 10^7
reach the same energy (1e-3 kcal/mol) for only a few percent of the starts,
while the first iterations agree everywhere.  The computation is
deterministic (every start's noise stream is seeded), so a change to the
oracle's optimiser shows up here as a changed rate.
"""
import json
import os
import sys

import numpy as np

from conftest import GOLDEN, ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def test_oracle_self_agreement_rates_are_reproduced():
    import make_lbfgs_self_agreement as M
    with open(os.path.join(GOLDEN, "lbfgs_self_agreement.json")) as fh:
        ref = json.load(fh)
    got = M.measure()
    np.testing.assert_allclose(got["full_step_energies"], ref["full_step_energies"], rtol=1e-12,
                               atol=1e-12)
    for part in ("prefix", "full_step", "local_step"):
        assert json.dumps(got[part], sort_keys=True) == json.dumps(ref[part], sort_keys=True), part
    # the shape the device tests rely on: early iterations agree at every
    # start, the full runs almost never
    assert ref["prefix"]["1"]["agree_noise_1e-05"] == 1.0
    assert ref["prefix"]["2"]["agree_noise_1e-05"] == 1.0
    assert ref["full_step"]["agree_abs_1e-3_noise_1e-07"] < 0.25
