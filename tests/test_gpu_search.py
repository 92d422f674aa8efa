"""Device Monte-Carlo search, clustering and pose parameterisation (through
the C ABI) against the FP64 oracle.

The walk is driven by the same Philox4x32-10 stream on both sides, so:
  * with the optimiser off (relax/full max_iter = 0) a walker's trajectory
    is decided by its draws and by pscore comparisons only: GPU and oracle
    walk the same path unless an FP32-vs-FP64 pscore lands within rounding
    of a Metropolis or box threshold (allowed for at most 1 walker in 8);
    poses agree to 2e-3 A, pscores to 1e-4 relative;
  * walker 0 runs exactly one full L-BFGS: it equals the batched L-BFGS
    entry point from the same start (same device code);
  * with the optimiser on, walks are chaotic in FP32 vs FP64, so a full
    search is checked through what it reports about itself: the best pose's
    energy and pscore re-evaluated by the score and pscore entry points, the
    box, the evaluation counts, and the clustering against the oracle's
    pose_cluster run on the device results;
  * splitting the step loop over more or fewer launches changes nothing.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import ffi, lib, scoring, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sw():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    ligs = [synth.make_ligand(44, n_heavy=16, n_torsions=4, n_rings=1),
            synth.make_ligand(45, n_heavy=24, n_torsions=6, n_rings=2)]
    box = lib.decide_grid_dims(rec.pocket_center, [16.0] * 3, 0.5)
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types)
    og = O.Grid(box, G.types, G.download().astype(np.float64), 1e6, rec.leaves.movable())
    x0 = [np.asarray(np.asarray(synth.ligand_apply(m.tree, m.xyz[0], rec.pocket_center, np.eye(3),
                                                   np.zeros(m.tree.n_torsions)), np.float32),
                     np.float64) for m in ligs]
    return dict(rec=rec, ligs=ligs, box=box, R=R, G=G, og=og, x0=x0,
                sp=scoring.load_pose_selection(), B=lib.LigandBatch(ligs),
                systems=[O.System(rec, m, og) for m in ligs])


def _no_opt(seed, steps, spl=10):
    prm = ffi.mc_params(seed=seed, max_nsteps=steps, steps_per_launch=spl)
    prm.relax = ffi.lbfgs_params(0)
    prm.full = ffi.lbfgs_params(0)
    return prm


def _in_box(x, box):
    lo, hi = np.array(box.init_xyz), np.array(box.oppo_xyz)
    return bool(((x > lo) & (x < hi)).all())


def _compare_with_oracle(sw, prm, walkers):
    print(f"\n[mc] seed={prm.seed} steps={prm.max_nsteps} walkers={walkers}", flush=True)
    res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], walkers)
    bp, be, ne = (res.best_pscore.cpu().numpy(), res.best_energy.cpu().numpy(),
                  res.nevals.cpu().numpy())
    bt = res.best_tors.cpu().numpy()
    nl = sw["rec"].leaves.n_leaves
    ok, rows = 0, []
    for l, lig in enumerate(sw["ligs"]):
        for k in range(walkers):
            w = l * walkers + k
            o = O.mc_walker(sw["systems"][l], lig, sw["sp"], prm, sw["box"], k, l, sw["x0"][l])
            x = res.walker_xyz(w)
            dx = float(np.abs(x - o["xyz"]).max())
            same = (dx <= 2e-3 and abs(bp[w] - o["pscore"]) <= 1e-4 * max(abs(o["pscore"]), 1.0)
                    and ne[w] == o["nevals"]
                    and np.abs(bt[w, :nl] - o["tors"]).max(initial=0.0) <= 1e-4)
            ok += same
            rows.append((l, k, bp[w], o["pscore"], int(ne[w]), o["nevals"], dx))
    for r in rows:
        print("   ", r)
    print(f"   agree {ok}/{len(rows)}")
    return ok, len(rows), res


def test_gpu_randomize_matches_oracle(sw):
    ok, n, res = _compare_with_oracle(sw, _no_opt(11, 0), 8)
    assert ok >= n - n // 8, (ok, n)
    for w in range(n):
        if w % 8:
            assert _in_box(res.walker_xyz(w), sw["box"])


def test_gpu_mc_walk_matches_oracle_without_optimiser(sw):
    ok, n, _ = _compare_with_oracle(sw, _no_opt(5, 20), 8)
    assert ok >= n - n // 8, (ok, n)


def test_gpu_mc_launch_split_invariance(sw):
    """The walk is a function of the draws only: 1, 3 or all steps per launch
    give the same bits."""
    outs = []
    for spl in (1, 3, 20):
        print(f"\n[mc] steps_per_launch={spl}", flush=True)
        prm = ffi.mc_params(seed=2, max_nsteps=6, relax_iters=10, full_iters=30,
                            steps_per_launch=spl)
        r = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], 4)
        outs.append((r.best_xyz.cpu().numpy(), r.best_pscore.cpu().numpy(),
                     r.nevals.cpu().numpy()))
    for o in outs[1:]:
        np.testing.assert_array_equal(o[0], outs[0][0])
        np.testing.assert_array_equal(o[1], outs[0][1])
        np.testing.assert_array_equal(o[2], outs[0][2])


def test_gpu_walker0_equals_batched_lbfgs(sw):
    prm = ffi.mc_params(seed=9, max_nsteps=5, relax_iters=10, full_iters=60)
    print("\n[mc] walker 0 vs pxd_lbfgs_batch", flush=True)
    res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], 1)
    nl = sw["rec"].leaves.n_leaves
    P = lib.PoseSet(sw["B"], [0, 1], sw["x0"], nl)
    e, ne, _ = lib.lbfgs(sw["R"], sw["G"], P, prm.full)
    e, ne = e.cpu().numpy(), ne.cpu().numpy()
    np.testing.assert_allclose(res.best_energy.cpu().numpy(), e, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(res.nevals.cpu().numpy(), ne)
    for p in range(2):
        np.testing.assert_allclose(res.walker_xyz(p), P.xyz_of(p), atol=1e-6)
    # the pscore of walker 0 is the pscore entry point at that pose
    s, _ = lib.pscore(sw["R"], P, sw["sp"], with_terms=False)
    np.testing.assert_allclose(res.best_pscore.cpu().numpy(), s.cpu().numpy(), rtol=1e-9)


def test_gpu_full_search_self_consistent_and_clustered(sw):
    walkers, modes = 8, 4
    prm = ffi.mc_params(seed=17, max_nsteps=12, relax_iters=20, full_iters=60)
    print("\n[mc] full search + clustering", flush=True)
    res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], walkers,
                        num_modes=modes, min_rmsd=1.0)
    bp, be, ne = (res.best_pscore.cpu().numpy(), res.best_energy.cpu().numpy(),
                  res.nevals.cpu().numpy())
    bt = res.best_tors.cpu().numpy()
    nl = sw["rec"].leaves.n_leaves
    W = len(sw["ligs"]) * walkers
    pose_lig = [w // walkers for w in range(W)]
    P = lib.PoseSet(sw["B"], pose_lig, [res.walker_xyz(w) for w in range(W)], nl, tors=bt)
    s, _ = lib.pscore(sw["R"], P, sw["sp"], with_terms=False)
    e, _, _ = lib.score_grad(sw["R"], sw["G"], P)
    s, e = s.cpu().numpy(), e.cpu().numpy()
    # the same walkers stopped after their first full optimisation
    p0 = ffi.mc_params(seed=17, max_nsteps=0, relax_iters=20, full_iters=60)
    first = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], p0, sw["box"], walkers)
    out_of_box = 0
    for w in range(W):
        assert ne[w] > 0
        assert abs(bp[w] - s[w]) <= 1e-6 * max(abs(s[w]), 1.0), (w, bp[w], s[w])
        assert abs(be[w] - e[w]) <= 1e-5 * max(abs(e[w]), 1.0), (w, be[w], e[w])
        # a best pose may lie outside the box: a relaxed start that left it
        # stops the walker (searcher.cpp:130, 182-184), and a refined pose is
        # kept on pscore alone with no box check (searcher.cpp:167-174)
        out_of_box += not _in_box(res.walker_xyz(w), sw["box"])
    print(f"   best poses outside the box: {out_of_box}/{W}")
    # walkers never end above their first optimised pose
    assert (bp <= first.best_pscore.cpu().numpy() + 1e-12).all()
    # clustering vs the oracle's pose_cluster on the same candidates
    ci, cc = res.cluster_idx.cpu().numpy(), res.cluster_count.cpu().numpy()
    for l, lig in enumerate(sw["ligs"]):
        xs = np.stack([res.walker_xyz(l * walkers + k) for k in range(walkers)]).astype(np.float64)
        out = np.zeros(modes, np.int32)
        cnt = O.lib().or_pose_cluster(walkers, lig.n_atoms, O.dp(np.ascontiguousarray(xs)),
                                      O.dp(np.ascontiguousarray(bp[l * walkers:(l + 1) * walkers])),
                                      O.ip(np.ascontiguousarray(ne[l * walkers:(l + 1) * walkers])),
                                      modes, 1.0, O.ip(out))
        assert cc[l] == cnt
        np.testing.assert_array_equal(ci[l, :cnt], out[:cnt])
        assert (ci[l, cnt:] == -1).all()


def test_gpu_apply_poses_matches_oracle(sw):
    rng = np.random.default_rng(4)
    pose_lig, xin, cen, rot, tors = [], [], [], [], []
    for p in range(12):
        l = p % 2
        lig = sw["ligs"][l]
        pose_lig.append(l)
        xin.append(np.asarray(np.asarray(lig.xyz[0], np.float32), np.float64))
        cen.append(np.asarray(rng.normal(size=3) * 5, np.float32).astype(np.float64))
        rot.append(np.asarray(synth.quat_to_rot(synth.random_quaternion(rng)), np.float32)
                   .astype(np.float64))
        tors.append(np.asarray(rng.uniform(-math.pi, math.pi, lig.tree.n_torsions), np.float32)
                    .astype(np.float64))
    print("\n[apply] 12 poses", flush=True)
    out, off = lib.apply_poses(sw["B"], pose_lig, xin, np.array(cen), np.array(rot), tors)
    for p in range(12):
        lig = sw["ligs"][pose_lig[p]]
        ref = O.ligand_apply(lig.tree, xin[p], cen[p], rot[p].reshape(3, 3).ravel(), tors[p])
        got = out[off[p]:off[p] + lig.n_atoms]
        assert np.abs(got - ref).max() <= 1e-4, (p, np.abs(got - ref).max())


def test_gpu_search_distributions_over_seeds_match_oracle(sw):
    """With the optimiser on, FP32 and FP64 walks part ways after a few steps,
    so the search is compared as a stochastic process: over 6 seeds x 4
    walkers per ligand, the best pscores found and the RMSDs of the walkers'
    best poses to the ligand's best pose must come from the same
    distributions on the GPU and in the oracle (two-sample Kolmogorov-Smirnov
    test not rejecting at 0.1 %), and the best pscores found over all seeds
    must agree within 1 kcal/mol (the minimum of 24 chaotic walks is an
    extreme order statistic: changing only FP32 rounding in the score has
    moved it by 0.6 on this system while the KS tests stayed far from
    rejecting)."""
    from concurrent.futures import ThreadPoolExecutor

    from scipy.stats import ks_2samp
    seeds, walkers, steps = range(1, 7), 4, 8
    gpu = {l: [] for l in range(len(sw["ligs"]))}
    jobs = []
    for seed in seeds:
        prm = ffi.mc_params(seed=seed, max_nsteps=steps, relax_iters=10, full_iters=30)
        print(f"\n[mc] distribution seed={seed}", flush=True)
        res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"],
                            walkers)
        bp = res.best_pscore.cpu().numpy()
        for l in range(len(sw["ligs"])):
            for k in range(walkers):
                w = l * walkers + k
                gpu[l].append((float(bp[w]), res.walker_xyz(w)))
                jobs.append((l, k, prm))
    with ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(lambda j: (j[0], O.mc_walker(sw["systems"][j[0]], sw["ligs"][j[0]],
                                                        sw["sp"], j[2], sw["box"], j[1], j[0],
                                                        sw["x0"][j[0]])), jobs))
    ora = {l: [] for l in gpu}
    for l, o in outs:
        ora[l].append((o["pscore"], o["xyz"]))

    def rmsd_to_best(rows):
        best = min(rows, key=lambda r: r[0])[1]
        return [float(np.sqrt(((r[1] - best) ** 2).sum(1).mean())) for r in rows]

    for l in gpu:
        gs = np.array([r[0] for r in gpu[l]])
        os_ = np.array([r[0] for r in ora[l]])
        gr, orr = rmsd_to_best(gpu[l]), rmsd_to_best(ora[l])
        p_s = ks_2samp(gs, os_).pvalue
        p_r = ks_2samp(gr, orr).pvalue
        print(f"   ligand {l}: best pscore gpu {gs.min():.3f} oracle {os_.min():.3f}; "
              f"median gpu {np.median(gs):.3f} oracle {np.median(os_):.3f}; "
              f"KS p(score) {p_s:.3g} p(rmsd) {p_r:.3g}")
        assert p_s > 1e-3 and p_r > 1e-3
        assert abs(gs.min() - os_.min()) <= 1.0


def test_gpu_gaussian_walk_matches_oracle_without_optimiser(sw):
    """The full-search mutation (pxd_mc_params.mutation 1: translation
    N(0, 1 A^2) per axis, rotation vector N(0, 0.3^2 rad^2), one torsion
    uniform), step by step against the oracle on the shared stream."""
    prm = _no_opt(23, 20)
    prm.mutation = 1
    prm.trans_sigma, prm.rot_sigma = 1.0, 0.3
    ok, n, _ = _compare_with_oracle(sw, prm, 8)
    assert ok >= n - n // 8, (ok, n)


def _search_params(seed, steps):
    """The full docking search's chain parameters: Gaussian mutation, local
    L-BFGS with history 8, at most 300 iterations, |g|inf <= 1e-3."""
    prm = ffi.mc_params(seed=seed, max_nsteps=steps, steps_per_launch=5)
    prm.mutation = 1
    prm.trans_sigma, prm.rot_sigma = 1.0, 0.3
    prm.relax = ffi.lbfgs_params(300, history=8, tol_grad=1e-3)
    prm.full = ffi.lbfgs_params(300, history=8, tol_grad=1e-3)
    return prm


def test_gpu_full_search_parameters_match_oracle_distribution(sw):
    """The full-search chain (Gaussian mutation, L-BFGS history 8 / 300
    iterations / 1e-3) over 4 seeds x 6 walkers per ligand: the best pscores
    and the evaluation counts per walker come from the oracle's distributions
    (two-sample KS not rejecting at 0.1 %), the top-9 clustering equals the
    oracle's pose_cluster on the device candidates."""
    from concurrent.futures import ThreadPoolExecutor

    from scipy.stats import ks_2samp
    walkers, steps = 6, 6
    gpu_ps, gpu_ne, jobs = {0: [], 1: []}, {0: [], 1: []}, []
    for seed in range(31, 35):
        prm = _search_params(seed, steps)
        print(f"\n[mc] full-search parameters seed={seed}", flush=True)
        res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], walkers,
                            num_modes=9, min_rmsd=0.5)
        bp, ne = res.best_pscore.cpu().numpy(), res.nevals.cpu().numpy()
        ci, cc = res.cluster_idx.cpu().numpy(), res.cluster_count.cpu().numpy()
        for l, lig in enumerate(sw["ligs"]):
            for k in range(1, walkers):  # walker 0 takes no Monte-Carlo step
                gpu_ps[l].append(float(bp[l * walkers + k]))
                gpu_ne[l].append(int(ne[l * walkers + k]))
                jobs.append((l, k, prm))
            xs = np.stack([res.walker_xyz(l * walkers + k) for k in range(walkers)]).astype(np.float64)
            out = np.zeros(9, np.int32)
            cnt = O.lib().or_pose_cluster(walkers, lig.n_atoms, O.dp(np.ascontiguousarray(xs)),
                                          O.dp(np.ascontiguousarray(bp[l * walkers:(l + 1) * walkers])),
                                          O.ip(np.ascontiguousarray(ne[l * walkers:(l + 1) * walkers])),
                                          9, 0.5, O.ip(out))
            assert cc[l] == cnt
            np.testing.assert_array_equal(ci[l, :cnt], out[:cnt])
    with ThreadPoolExecutor(16) as ex:
        outs = list(ex.map(lambda j: (j[0], O.mc_walker(sw["systems"][j[0]], sw["ligs"][j[0]],
                                                        sw["sp"], j[2], sw["box"], j[1], j[0],
                                                        sw["x0"][j[0]])), jobs))
    for l in gpu_ps:
        ops = np.array([o["pscore"] for m, o in outs if m == l])
        one = np.array([o["nevals"] for m, o in outs if m == l])
        gps, gne = np.array(gpu_ps[l]), np.array(gpu_ne[l])
        p_s, p_n = ks_2samp(gps, ops).pvalue, ks_2samp(gne, one).pvalue
        print(f"   ligand {l}: best pscore median gpu {np.median(gps):.3f} oracle "
              f"{np.median(ops):.3f} (KS p {p_s:.3g}); evaluations median gpu {np.median(gne):.0f} "
              f"oracle {np.median(one):.0f} (KS p {p_n:.3g})", flush=True)
        assert p_s > 1e-3 and p_n > 1e-3


def test_gpu_energy_prune_matches_oracle_without_optimiser(sw):
    """searcher.cpp:141-157 on the device: two walkers per ligand (walker 0
    takes no Monte-Carlo step, so the ligand minimum is walker 1's alone and
    the walk is deterministic), optimiser off, step by step against the
    oracle; the threshold is low enough for the oracle to prune steps."""
    prm = _no_opt(41, 24)
    prm.mmenergy_threshold = 25.0
    res = lib.mc_search(sw["R"], sw["G"], sw["B"], sw["x0"], sw["sp"], prm, sw["box"], 2)
    bp, ne = res.best_pscore.cpu().numpy(), res.nevals.cpu().numpy()
    pruned, ok = 0, 0
    for l, lig in enumerate(sw["ligs"]):
        emin = np.array([1e8])
        o = O.mc_walker(sw["systems"][l], lig, sw["sp"], prm, sw["box"], 1, l, sw["x0"][l],
                        emin=emin)
        w = 2 * l + 1
        dx = float(np.abs(res.walker_xyz(w) - o["xyz"]).max())
        print(f"\n[mc prune] ligand {l}: pruned {o['pruned']} of 24; pscore gpu {bp[w]:.4f} "
              f"oracle {o['pscore']:.4f}; evaluations {ne[w]} / {o['nevals']}; dx {dx:.2e}",
              flush=True)
        pruned += o["pruned"]
        ok += dx <= 2e-3 and ne[w] == o["nevals"] and \
            abs(bp[w] - o["pscore"]) <= 1e-4 * max(abs(o["pscore"]), 1.0)
    assert pruned > 0
    assert ok == len(sw["ligs"])


def test_gpu_own_conformer_starts_match_oracle_without_optimiser(sw):
    """searcher.cpp:107-113: ligand l brings three conformers; walkers 1 and 2
    start from their own (no torsion draws), walkers 3 and 4 from conformer
    0 with random torsions; every walker against the oracle started the same
    way, optimiser off."""
    rng = np.random.default_rng(6)
    confs = []
    for l, lig in enumerate(sw["ligs"]):
        cl = [sw["x0"][l]]
        for _ in range(2):
            t = rng.uniform(-np.pi, np.pi, lig.tree.n_torsions)
            x = synth.ligand_apply(lig.tree, sw["x0"][l], sw["rec"].pocket_center, np.eye(3), t)
            cl.append(np.asarray(np.asarray(x, np.float32), np.float64))
        confs.append(cl)
    walkers = 5
    prm = _no_opt(43, 6)
    res = lib.mc_search(sw["R"], sw["G"], sw["B"], None, sw["sp"], prm, sw["box"], walkers,
                        conformers=confs)
    bp, ne = res.best_pscore.cpu().numpy(), res.nevals.cpu().numpy()
    ok, n = 0, 0
    for l, lig in enumerate(sw["ligs"]):
        for k in range(walkers):
            own = 0 < k < len(confs[l])
            o = O.mc_walker(sw["systems"][l], lig, sw["sp"], prm, sw["box"], k, l,
                            confs[l][k] if own else confs[l][0], own_conformer=own)
            w = l * walkers + k
            dx = float(np.abs(res.walker_xyz(w) - o["xyz"]).max())
            n += 1
            ok += dx <= 2e-3 and ne[w] == o["nevals"] and \
                abs(bp[w] - o["pscore"]) <= 1e-4 * max(abs(o["pscore"]), 1.0)
    print(f"\n[mc conformers] agree {ok}/{n}", flush=True)
    assert ok >= n - n // 8, (ok, n)
