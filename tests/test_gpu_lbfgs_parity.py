"""Persistent batched L-BFGS on the device against the FP64 oracle, on the
config-0 system: the 2,000-heavy-atom receptor with 8 rotatable hydroxyls, a
61^3 grid (22.5 A at 0.375 A), ligands of the screening generator (10-60
heavy atoms) plus one of 92 atoms, whose 284 variables take the widest
vector class.

Same-start final energies cannot be the test on their own: the oracle
perturbed by one part in 10^7 reaches the same final energy (1e-3 kcal/mol)
from the same start for only a few percent of the starts (the committed
rates in tests/golden/lbfgs_self_agreement.json, and the rates measured
here).  What the device must show instead:

  * the trajectory: after k iterations (k = 1 .. 20) it agrees with the
    oracle at least as often as the oracle agrees with itself under a
    1e-5 relative perturbation of its evaluations (the size of the FP32
    evaluation error, within the 1e-4 per-evaluation tolerance), less four
    standard errors;
  * the optimiser state: the plain form (FP64 host optimiser on the same
    device evaluations) and the batched kernel (FP32 vectors, FP64 scalars
    and curvature products) give the same energies over the first two
    iterations;
  * the full runs, at the reference full step (history 100, 240 iterations,
    |g|inf <= 0.5) and at the search's local step (history 8, 300
    iterations, |g|inf <= 1e-3): the distribution of final energies is the
    oracle's (two-sample Kolmogorov-Smirnov), and so are the converged
    fraction and the mean evaluation count;
  * restarted next to a minimum of the brute-force energy (where the
    gradient is the energy's own, so minima exist), the full step lands on
    the oracle's pose (RMSD <= 0.05 A) and on its energy within 1e-3
    kcal/mol as often as the oracle does under the same perturbation.
"""
import math

import numpy as np
import pytest
from scipy import stats

import lbfgs_agreement as A
import pyoracle as O
from pxdock_amd import ffi, lib, synth

pytestmark = pytest.mark.gpu

REL = 1e-5  # FP32 evaluation error scale used for the perturbed oracle


def _screen_ligand(i):
    rng = np.random.default_rng(10000 + i + 7919)
    nh = int(rng.integers(10, 61))
    nt = int(rng.integers(0, 16))
    nr = int(rng.integers(0, 4))
    return synth.make_ligand(10000 + i, n_heavy=nh, n_torsions=nt, n_rings=nr)


@pytest.fixture(scope="module")
def c0():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(1)
    box = lib.decide_grid_dims(rec.pocket_center, [22.5] * 3, 0.375)
    ligs = [_screen_ligand(i) for i in range(48)]
    big = synth.make_ligand(777, n_heavy=75, n_torsions=12, n_rings=2)
    assert 3 * big.n_atoms + rec.leaves.n_leaves > 256  # the widest vector class
    ligs.append(big)
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types)
    og = O.Grid(box, G.types, G.download().astype(np.float64), 1e6, rec.leaves.movable())
    return dict(rec=rec, box=box, ligs=ligs, R=R, G=G, B=lib.LigandBatch(ligs),
                systems=[O.System(rec, m, og) for m in ligs],
                bf_systems=[O.System(rec, m, None) for m in ligs])


def _device(c0, pose_lig, xyz, prm, plain=False, grid=True):
    nl = c0["rec"].leaves.n_leaves
    P = lib.PoseSet(c0["B"], pose_lig, xyz, nl)
    fn = lib.lbfgs_plain if plain else lib.lbfgs
    e, ne, cv = fn(c0["R"], c0["G"] if grid else None, P, prm)
    return e.cpu().numpy(), cv.cpu().numpy().astype(bool), ne.cpu().numpy(), P


def test_gpu_lbfgs_trajectory_tracks_the_oracle(c0):
    pose_lig, xyz = A.pocket_starts(c0["rec"], c0["ligs"], 2, seed=101, big_per=8)
    n = len(pose_lig)
    tors = np.zeros((n, c0["rec"].leaves.n_leaves))
    print(f"\n[lbfgs prefix] {n} starts", flush=True)
    for k in (1, 2, 4, 10, 20):
        prm = ffi.lbfgs_params(k)
        e, _, _, _ = _device(c0, pose_lig, xyz, prm)
        e0, _, _ = A.oracle_runs(c0["systems"], pose_lig, xyz, tors, prm)
        e1, _, _ = A.oracle_runs(c0["systems"], pose_lig, xyz, tors, prm, noise=REL, seed=1)
        dev, self_ = A.agree(e, e0, REL), A.agree(e1, e0, REL)
        print(f"   iterations {k:3d}: device vs oracle {dev:.3f}, oracle vs perturbed oracle "
              f"{self_:.3f}", flush=True)
        assert dev >= self_ - A.binomial_margin(self_, n), (k, dev, self_)
        if k <= 2:
            ep, _, _, _ = _device(c0, pose_lig, xyz, prm, plain=True)
            assert A.agree(e, ep, 1e-6) >= 0.98, k


@pytest.mark.parametrize("step", ["full", "local"])
def test_gpu_lbfgs_final_energies_have_the_oracle_distribution(c0, step):
    prm = (ffi.lbfgs_params(240) if step == "full"
           else ffi.lbfgs_params(300, history=8, tol_grad=1e-3))
    pose_lig, xyz = A.pocket_starts(c0["rec"], c0["ligs"], 4, seed=202, big_per=16)
    n = len(pose_lig)
    tors = np.zeros((n, c0["rec"].leaves.n_leaves))
    e, cv, ne, _ = _device(c0, pose_lig, xyz, prm)
    e0, c0_, n0 = A.oracle_runs(c0["systems"], pose_lig, xyz, tors, prm)
    e1, _, _ = A.oracle_runs(c0["systems"], pose_lig, xyz, tors, prm, noise=1e-7, seed=2)
    ks = stats.ks_2samp(e, e0).pvalue
    print(f"\n[lbfgs {step}] {n} starts: KS p {ks:.3g}; converged {cv.mean():.3f} vs "
          f"{c0_.mean():.3f}; evaluations {ne.mean():.1f} vs {n0.mean():.1f}; same final "
          f"energy (1e-3) device {A.agree(e, e0, 0.0):.3f}, oracle under 1e-7 noise "
          f"{A.agree(e1, e0, 0.0):.3f}", flush=True)
    assert np.isfinite(e).all()
    assert ks > 1e-3
    assert abs(cv.mean() - c0_.mean()) <= A.binomial_margin(c0_.mean(), n)
    assert abs(ne.mean() - n0.mean()) <= 0.1 * n0.mean()
    big = pose_lig == len(c0["ligs"]) - 1
    # the widest class: finite, within the evaluation budget, and its
    # energies drawn from the oracle's distribution for that ligand
    assert (ne[big] >= 1).all() and (ne[big] <= prm.max_eval + prm.max_ls + 1).all()
    assert stats.ks_2samp(e[big], e0[big]).pvalue > 1e-3


def test_gpu_lbfgs_returns_to_the_oracle_minimum(c0):
    """Brute-force energy (no maps): minima are found by a long device run
    (history 8, 3,000 iterations), then every start's ligand is moved 0.01 A
    per coordinate (receptor torsions kept) and optimised again at the
    reference full step on both sides.  Both must land on the same pose
    (RMSD <= 0.05 A; the oracle's runs under a 1e-5 perturbation stay within
    0.01 A of each other), with the same energy within 1e-3 kcal/mol as often
    as the perturbed oracle does."""
    ligs = c0["ligs"]
    sel = list(range(0, 48, 3))
    pose_lig, xyz = A.pocket_starts(c0["rec"], [ligs[i] for i in sel], 4, seed=303)
    pose_lig = np.array([sel[i] for i in pose_lig], np.int32)
    n = len(pose_lig)
    long = ffi.lbfgs_params(3000, history=8, tol_grad=1e-3, tol_change=1e-12)
    _, _, _, P = _device(c0, pose_lig, xyz, long, grid=False)
    nl = c0["rec"].leaves.n_leaves
    mins = [np.asarray(P.xyz_of(i), np.float64) for i in range(n)]
    mtors = P.tors.cpu().numpy()[:, :nl].astype(np.float64)
    gmax = []
    for i in range(n):
        _, g, tg = c0["bf_systems"][pose_lig[i]].energy(mins[i], mtors[i])
        gmax.append(max(np.abs(g).max(), np.abs(tg).max(initial=0.0)))
    rng = np.random.default_rng(4)
    starts = [np.asarray(np.asarray(m + rng.normal(0.0, 0.01, m.shape), np.float32), np.float64)
              for m in mins]
    prm = ffi.lbfgs_params(240)
    Pd = lib.PoseSet(c0["B"], pose_lig, starts, nl, tors=mtors)
    e, _, _ = lib.lbfgs(c0["R"], None, Pd, prm)
    e = e.cpu().numpy()
    e0, _, _, x0 = A.oracle_runs(c0["bf_systems"], pose_lig, starts, mtors, prm, poses=True)
    e1, _, _, x1 = A.oracle_runs(c0["bf_systems"], pose_lig, starts, mtors, prm, noise=REL,
                                 seed=3, poses=True)
    dev_pose = np.mean([A.rmsd(Pd.xyz_of(i), x0[i]) <= 0.05 for i in range(n)])
    self_pose = np.mean([A.rmsd(x1[i], x0[i]) <= 0.05 for i in range(n)])
    dev, self_ = A.agree(e, e0, 0.0), A.agree(e1, e0, 0.0)
    print(f"\n[lbfgs restart] {n} starts at device minima (|g|inf median "
          f"{np.median(gmax):.2g}): same pose (0.05 A) device {dev_pose:.3f}, oracle under "
          f"{REL:g} noise {self_pose:.3f}; same energy (1e-3) device {dev:.3f}, oracle {self_:.3f}",
          flush=True)
    # the long device runs end at minima of the FP64 energy to within what FP32
    # coordinates resolve (median |g|inf ~0.06), well inside the full step's 0.5
    assert np.median(gmax) <= 0.2
    assert dev_pose >= 0.9 and self_pose >= 0.9
    assert dev >= self_ - A.binomial_margin(self_, n, z=3.0)
