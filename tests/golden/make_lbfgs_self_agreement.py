"""Writes lbfgs_self_agreement.json: how often the FP64 oracle's L-BFGS
(optimizer.cpp:429-560) reaches the same energy from the same start when its
evaluations are perturbed by one part in 10^7 or 10^5 (the oracle's
evaluation-noise mode), on a small docking system: truncated runs of 1-20
iterations, the reference full step (history 100, 240 iterations,
|g|inf <= 0.5) and the search's local step (history 8, 300 iterations,
|g|inf <= 1e-3).  Oracle only; run from the repository root:

    python tests/golden/make_lbfgs_self_agreement.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "protenix-dock_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

OUT = os.path.join(HERE, "lbfgs_self_agreement.json")
PREFIXES = (1, 2, 3, 5, 10, 20)
NOISE = (1e-7, 1e-5)
REL = 1e-5


def system():
    import pyoracle as O
    from pxdock_amd import lib, synth
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    ligs = [synth.make_ligand(600 + i, n_heavy=14 + 2 * i, n_torsions=2 + i % 6,
                              n_rings=i % 3) for i in range(12)]
    box = lib.decide_grid_dims(rec.pocket_center, [14.0] * 3, 0.5)
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    og = O.grid_build(rec, box, types, nthreads=8)
    systems = [O.System(rec, m, og) for m in ligs]
    return rec, ligs, systems


def measure():
    import lbfgs_agreement as A
    from pxdock_amd import ffi
    rec, ligs, systems = system()
    pose_lig, xyz = A.pocket_starts(rec, ligs, 4, seed=17)
    tors = np.zeros((len(pose_lig), rec.leaves.n_leaves))
    runs = {"prefix": {str(k): ffi.lbfgs_params(k) for k in PREFIXES},
            "full_step": ffi.lbfgs_params(240),
            "local_step": ffi.lbfgs_params(300, history=8, tol_grad=1e-3)}
    out = {"starts": int(len(pose_lig)), "rel_tolerance": REL, "abs_tolerance": 1e-3,
           "noise": list(NOISE)}

    def rates(prm):
        e0, c0, n0 = A.oracle_runs(systems, pose_lig, xyz, tors, prm, nthreads=8)
        r = {"converged_frac": float(c0.mean()), "mean_nevals": float(n0.mean())}
        for nz in NOISE:
            e1, _, _ = A.oracle_runs(systems, pose_lig, xyz, tors, prm, noise=nz, seed=1,
                                     nthreads=8)
            r[f"agree_noise_{nz:g}"] = A.agree(e1, e0, REL)
            r[f"agree_abs_1e-3_noise_{nz:g}"] = A.agree(e1, e0, 0.0)
        return r, e0
    out["prefix"] = {}
    for k, prm in runs["prefix"].items():
        out["prefix"][k], _ = rates(prm)
    out["full_step"], e_full = rates(runs["full_step"])
    out["local_step"], _ = rates(runs["local_step"])
    out["full_step_energies"] = [float(v) for v in e_full]
    return out


if __name__ == "__main__":
    res = measure()
    with open(OUT, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps({k: v for k, v in res.items() if k != "full_step_energies"}, indent=1))
