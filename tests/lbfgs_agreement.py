"""Same-start agreement of L-BFGS runs, and the oracle's own agreement under
a perturbation of its evaluations.

L-BFGS with a strong-Wolfe line search (optimizer.cpp:429-560) is chaotic on
these energy surfaces: a change of one part in 10^7 in the evaluations moves
the line search onto another branch after a few iterations, and the runs end
in different places.  So a same-start comparison of the FP32 device
optimiser with the FP64 oracle is judged against the oracle compared with
itself under a perturbation of the size of the device's evaluation error:
the oracle's evaluation-noise mode multiplies every energy and gradient
component the optimiser sees by (1 + rel u), u uniform in [-1, 1), from a
stream seeded per start.  Shared by the CPU test of the committed rates and
the GPU parity tests.
"""
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import pyoracle as O
from pxdock_amd import synth


def threads():
    return max(1, os.cpu_count() or 1)


def pocket_starts(rec, ligs, per, seed, spread=1.5, big_per=None):
    """per starts of every ligand (big_per for the last one, if given): centre
    N(pocket centre, spread^2) per axis, uniform orientation, torsions
    uniform in [-pi, pi), rounded to FP32 as the device holds them."""
    rng = np.random.default_rng(seed)
    pose_lig, xyz = [], []
    for l, lig in enumerate(ligs):
        n = big_per if (big_per is not None and l == len(ligs) - 1) else per
        for _ in range(n):
            c = rec.pocket_center + rng.normal(0.0, spread, 3)
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            x = synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)
            xyz.append(np.asarray(np.asarray(x, np.float32), np.float64))
            pose_lig.append(l)
    return np.array(pose_lig, np.int32), xyz


def oracle_runs(systems, pose_lig, xyz, tors, prm, noise=0.0, seed=0, nthreads=None,
                poses=False):
    """FP64 oracle L-BFGS from every start; start i's noise stream is seeded
    seed * 1000003 + i.  Returns (energy, converged, nevals) arrays, and the
    final ligand coordinates when poses is set."""
    def one(i):
        O.lib().or_set_eval_noise(noise, seed * 1000003 + i)
        try:
            return systems[pose_lig[i]].lbfgs(xyz[i], tors[i], prm)
        finally:
            O.lib().or_set_eval_noise(0.0, 0)
    with ThreadPoolExecutor(nthreads or threads()) as ex:
        out = list(ex.map(one, range(len(pose_lig))))
    res = (np.array([o["energy"] for o in out]), np.array([o["converged"] for o in out]),
           np.array([o["nevals"] for o in out]))
    return res + ([o["xyz"] for o in out],) if poses else res


def rmsd(a, b):
    return float(np.sqrt(((np.asarray(a) - np.asarray(b)) ** 2).sum(axis=1).mean()))


def agree(a, b, rel, atol=1e-3):
    """Fraction of starts whose energies agree within max(atol, rel |b|)."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.mean(np.abs(a - b) <= np.maximum(atol, rel * np.abs(b))))


def binomial_margin(p, n, z=4.0):
    """z standard errors of the difference of two independent rates p over n
    starts each."""
    return z * math.sqrt(max(p * (1.0 - p), 1.0 / n) * 2.0 / n)
