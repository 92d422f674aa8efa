"""HIP path (libpxdock.so through its C ABI) against the FP64 oracle.

Tolerances (FP32 device arithmetic vs FP64 reference semantics):
  * energies: |dE| <= 1e-4 * max(|E|, 1) (the 1e-4 relative target; the
    max(., 1) keeps near-zero totals from demanding more than 1e-4 absolute);
  * gradients: norm-wise ||dg|| <= 1e-4 * max(||g||, 1);
  * grid maps: |dv| <= 1e-4 * |v| + 1e-4 per channel.
"""
import json
import math

import numpy as np
import pytest

import pyoracle as O
from conftest import golden_path
from pxdock_amd import data, ffi, lib, scoring, synth

pytestmark = pytest.mark.gpu

ETOL = 1e-4


def _close_e(a, b, rel=ETOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b) <= rel * np.maximum(np.abs(b), 1.0)


def _grad_ok(g, ref, rel=ETOL):
    return np.linalg.norm(g - ref) <= rel * max(np.linalg.norm(ref), 1.0)


@pytest.fixture(scope="module")
def world():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(1)
    box = lib.decide_grid_dims(rec.pocket_center, [22.5] * 3, 0.375)
    ligs = [synth.make_ligand(1000 + i, n_heavy=h, n_torsions=t, n_rings=r)
            for i, (h, t, r) in enumerate([(30, 6, 1), (12, 2, 0), (45, 10, 2), (58, 14, 3)])]
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types)
    return dict(rec=rec, box=box, ligs=ligs, types=types, R=R, G=G)


def _poses(world, n_per, seed, spread=None):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(world["box"].init_xyz), np.array(world["box"].oppo_xyz)
    pose_lig, xyz = [], []
    for l, lig in enumerate(world["ligs"]):
        for _ in range(n_per):
            c = rng.uniform(lo, hi) if spread is None else world["rec"].pocket_center + rng.normal(size=3) * spread
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)))
            pose_lig.append(l)
    tors = _f32(rng.uniform(-math.pi, math.pi, (len(xyz), world["rec"].leaves.n_leaves)))
    return pose_lig, xyz, tors


def _f32(a):
    """The device works on FP32 poses: hand the oracle the same rounded values."""
    return np.asarray(a, np.float32).astype(np.float64)


# ---------------------------------------------------------------- reference gtest vectors on GPU

def test_gpu_gtest_receptor_torsions(golden):
    """interaction test_interaction.cpp:280-454 through the HIP score kernel (no grid)."""
    g = golden["system_torsions"]
    rec = data.parse_receptor(golden_path(g["receptor_json"]))
    lig = data.parse_ligand(golden_path(g["ligand_json"]))
    R = lib.Receptor(rec)
    B = lib.LigandBatch([lig])
    for case in g["cases"]:
        t = np.radians(case["torsions_deg"])[None]
        P = lib.PoseSet(B, [0], [lig.xyz[0]], 3, tors=t)
        e, _, tg = lib.score_grad(R, None, P)
        e, tg = e.cpu().numpy()[0], tg.cpu().numpy()[0]
        assert _close_e(e, case["energy"]) or abs(e - case["energy"]) < case["energy_tol"]
        np.testing.assert_allclose(tg, case["tors_grad"], atol=max(case["tors_grad_tol"], 1e-4))


@pytest.mark.parametrize("idx", [0, 1])
def test_gpu_gtest_lbfgs(golden, idx):
    """lbfgs_step(1,8) and lbfgs_step(2,12) (test_optimizer.cpp:177-221) on the GPU."""
    g = golden["lbfgs"]
    case = g["cases"][idx]
    rec = data.parse_receptor(golden_path(g["receptor_json"]))
    lig = data.parse_ligand(golden_path(g["ligand_json"]))
    R = lib.Receptor(rec)
    B = lib.LigandBatch([lig])
    P = lib.PoseSet(B, [0], [lig.xyz[0]], 3, tors=np.array(case["torsions"], float)[None])
    prm = ffi.lbfgs_params(case["max_iter"], max_eval=case["max_eval"])
    e, ne, cv = lib.lbfgs(R, None, P, prm)
    e = e.cpu().numpy()[0]
    tors, xyz = P.tors.cpu().numpy()[0], P.xyz_of(0)
    print("gpu", e, "nevals", int(ne.cpu()[0]), "tors", tors, "xyz", xyz)
    if abs(e - case["energy"]) <= 1e-4 * abs(case["energy"]):
        np.testing.assert_allclose(tors, case["out_torsions"], atol=1e-3)
        np.testing.assert_allclose(xyz, case["out_xyz"], atol=1e-3)
        return
    # Both vectors sit on a branch of the line search (optimizer.cpp:217-231:
    # the sign of d1^2 - g1 g2 in cubic_interpolate picks the cubic minimum
    # or the bracket midpoint).  Relative evaluation noise of 1e-7 (FP32
    # level) sends the FP64 oracle down either branch
    # (lbfgs_branch_outcomes.json, recomputed by test_lbfgs_branch_outcomes.py);
    # the GPU result must be one of those oracle outcomes.
    with open(golden_path("lbfgs_branch_outcomes.json")) as fh:
        outs = json.load(fh)["cases"][idx]["outcomes"]
    match = [o for o in outs if abs(o["energy"] - e) <= 1e-4 * abs(e)]
    assert match, (e, [round(o["energy"], 4) for o in outs])
    np.testing.assert_allclose(tors, match[0]["out_torsions"], atol=1e-3)
    np.testing.assert_allclose(xyz, match[0]["out_xyz"], atol=1e-3)


# ---------------------------------------------------------------- grid build

def test_gpu_grid_build_matches_oracle(world):
    """K1 maps at sampled points vs the oracle's FP64 per-point sum (grid.cpp:164-194)."""
    maps = world["G"].download()
    box = world["box"]
    ny, nz = box.npoints[1], box.npoints[2]
    rng = np.random.default_rng(7)
    P = box.n_points()
    for p in np.concatenate([rng.integers(0, P, 300), [0, P - 1]]):
        i, j, k = p // (ny * nz), (p // nz) % ny, p % nz
        pt = np.array(box.init_xyz) + box.spacing * np.array([i, j, k])
        ref = O.grid_point(world["rec"], pt, world["G"].types)
        got = maps[:, p, :]
        big = np.abs(ref) > 3e37
        ok = np.abs(got - ref) <= 1e-4 * np.abs(ref) + 1e-4
        assert (ok | big).all(), (p, got, ref)


# ---------------------------------------------------------------- score + gradient

def _oracle_grid(world):
    maps = world["G"].download().astype(np.float64)
    movable = world["rec"].leaves.movable()
    return O.Grid(world["box"], world["G"].types, maps, 1e6, movable)


def test_gpu_score_grad_matches_oracle_with_maps(world):
    """K2 on config-0-style poses (uniform in the box) vs the oracle reading the same maps."""
    og = _oracle_grid(world)
    pose_lig, xyz, tors = _poses(world, 16, 11)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e, g, tg = lib.score_grad(world["R"], world["G"], P)
    e, tg = e.cpu().numpy(), tg.cpu().numpy()
    systems = [O.System(world["rec"], lig, og) for lig in world["ligs"]]
    bad = []
    for p in range(P.n_poses):
        re, rg, rtg = systems[pose_lig[p]].energy(xyz[p], tors[p])
        gg = P.xyz_of(p, g)
        full = np.concatenate([gg.ravel(), tg[p]])
        rfull = np.concatenate([rg.ravel(), rtg])
        if not (_close_e(e[p], re) and _grad_ok(full, rfull)):
            bad.append((p, e[p], re, np.linalg.norm(full - rfull), np.linalg.norm(rfull)))
    assert not bad, bad[:5]


def test_gpu_pair_schedules_at_size_boundaries(world):
    """Ligands of 3, 5, 33, 63, 64 atoms take the ring pair schedule (each
    pair once, reactions in a travelling accumulator); 65, 66, 96, 100, 127
    and 128 atoms take the two-slot ring (atoms lane and lane + 64, the
    accumulators moving across the slot boundary and wrapping at n).
    Energies and gradients against the oracle on the same maps, poses near
    the pocket (many close pairs)."""
    picks = [(5001, 2, 3), (5013, 3, 5), (5007, 28, 33), (5004, 54, 63), (5018, 54, 64),
             (5000, 54, 65), (5002, 54, 66), (5005, 75, 96), (5008, 81, 100),
             (5005, 102, 127), (5002, 103, 128)]
    ligs = []
    for seed, nh, n in picks:
        m = synth.make_ligand(seed, n_heavy=nh, n_torsions=min(3, max(0, nh - 3)), n_rings=0)
        assert m.n_atoms == n, (seed, nh, m.n_atoms)
        ligs.append(m)
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    G = lib.Grid(world["R"], world["box"], types)
    og = O.Grid(world["box"], G.types, G.download().astype(np.float64), 1e6,
                world["rec"].leaves.movable())
    rng = np.random.default_rng(77)
    pose_lig, xyz = [], []
    for l, lig in enumerate(ligs):
        for _ in range(3):
            c = world["rec"].pocket_center + rng.normal(size=3) * 2.0
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)))
            pose_lig.append(l)
    tors = _f32(rng.uniform(-math.pi, math.pi, (len(xyz), world["rec"].leaves.n_leaves)))
    B = lib.LigandBatch(ligs)
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e, g, tg = lib.score_grad(world["R"], G, P)
    e, tg = e.cpu().numpy(), tg.cpu().numpy()
    systems = [O.System(world["rec"], lig, og) for lig in ligs]
    bad = []
    for p in range(P.n_poses):
        re, rg, rtg = systems[pose_lig[p]].energy(xyz[p], tors[p])
        full = np.concatenate([P.xyz_of(p, g).ravel(), tg[p]])
        rfull = np.concatenate([rg.ravel(), rtg])
        if not (_close_e(e[p], re) and _grad_ok(full, rfull)):
            bad.append((ligs[pose_lig[p]].n_atoms, e[p], re, np.linalg.norm(full - rfull)))
    assert not bad, bad


def test_gpu_pair_list_form_when_masks_cannot_hold_the_list(world):
    """A pair list that names a pair twice cannot be held by partner masks:
    the ligand keeps the list form (each listed entry evaluated, gradients by
    LDS atomics).  The oracle walks the same list."""
    m = synth.make_ligand(5007, n_heavy=28, n_torsions=3, n_rings=0)
    m.pairs = np.concatenate([m.pairs, m.pairs[:5]])
    ligs = [m]
    types = sorted(set(int(t) for t in m.vdw_type))
    G = lib.Grid(world["R"], world["box"], types)
    og = O.Grid(world["box"], G.types, G.download().astype(np.float64), 1e6,
                world["rec"].leaves.movable())
    rng = np.random.default_rng(78)
    xyz = []
    for _ in range(4):
        c = world["rec"].pocket_center + rng.normal(size=3) * 2.0
        R = synth.quat_to_rot(synth.random_quaternion(rng))
        t = rng.uniform(-math.pi, math.pi, m.tree.n_torsions)
        xyz.append(_f32(synth.ligand_apply(m.tree, m.xyz[0], c, R, t)))
    tors = _f32(rng.uniform(-math.pi, math.pi, (len(xyz), world["rec"].leaves.n_leaves)))
    P = lib.PoseSet(lib.LigandBatch(ligs), [0] * len(xyz), xyz, 8, tors=tors)
    e, g, tg = lib.score_grad(world["R"], G, P)
    e, tg = e.cpu().numpy(), tg.cpu().numpy()
    sysm = O.System(world["rec"], m, og)
    for p in range(P.n_poses):
        re, rg, rtg = sysm.energy(xyz[p], tors[p])
        full = np.concatenate([P.xyz_of(p, g).ravel(), tg[p]])
        rfull = np.concatenate([rg.ravel(), rtg])
        assert _close_e(e[p], re), (p, e[p], re)
        assert _grad_ok(full, rfull), (p, np.linalg.norm(full - rfull))


def test_gpu_score_grad_brute_force_matches_oracle(world):
    """No grid: inter term by explicit pairs over every receptor atom (interaction.cpp:241-254)."""
    pose_lig, xyz, tors = _poses(world, 4, 12, spread=2.0)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e, g, tg = lib.score_grad(world["R"], None, P)
    e, tg = e.cpu().numpy(), tg.cpu().numpy()
    systems = [O.System(world["rec"], lig, None) for lig in world["ligs"]]
    for p in range(P.n_poses):
        re, rg, rtg = systems[pose_lig[p]].energy(xyz[p], tors[p])
        full = np.concatenate([P.xyz_of(p, g).ravel(), tg[p]])
        rfull = np.concatenate([rg.ravel(), rtg])
        assert _close_e(e[p], re), (p, e[p], re)
        assert _grad_ok(full, rfull), (p, np.linalg.norm(full - rfull), np.linalg.norm(rfull))


def test_gpu_grid_end_to_end_small_box():
    """Grid built on the GPU vs grid built by the oracle, through scoring (coarse box)."""
    rec = synth.make_receptor(2, n_heavy=600, radius=18.0, pocket_offset=8.0, n_ring6=4,
                              n_ring5=2, n_leaves=4)
    lig = synth.make_ligand(5, n_heavy=20, n_torsions=4, n_rings=1)
    box = lib.decide_grid_dims(rec.pocket_center, [12.0] * 3, 0.75)
    types = sorted(set(int(t) for t in lig.vdw_type))
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types)
    og = O.grid_build(rec, box, types)
    s = O.System(rec, lig, og)
    rng = np.random.default_rng(3)
    xyz, tors = [], []
    for _ in range(24):
        c = rec.pocket_center + rng.normal(size=3)
        xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c,
                                           synth.quat_to_rot(synth.random_quaternion(rng)),
                                           rng.uniform(-3, 3, lig.tree.n_torsions))))
        tors.append(_f32(rng.uniform(-3, 3, rec.leaves.n_leaves)))
    B = lib.LigandBatch([lig])
    P = lib.PoseSet(B, [0] * 24, xyz, rec.leaves.n_leaves, tors=np.array(tors))
    e, _, _ = lib.score_grad(R, G, P)
    e = e.cpu().numpy()
    for p in range(24):
        re, _, _ = s.energy(xyz[p], tors[p])
        assert _close_e(e[p], re), (p, e[p], re)


# ---------------------------------------------------------------- L-BFGS
# Same-start parity of the batched L-BFGS is tested in test_gpu_lbfgs_parity.py.


# ---------------------------------------------------------------- pscore

def test_gpu_pscore_matches_oracle(world):
    sp = scoring.load_pose_selection()
    pose_lig, xyz, tors = _poses(world, 8, 31, spread=2.0)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    s, t = lib.pscore(world["R"], P, sp)
    s, t = s.cpu().numpy(), t.cpu().numpy()
    for p in range(P.n_poses):
        rx = O.receptor_apply(world["rec"].leaves, world["rec"].xyz, tors[p])
        rs, rt = O.pscore(sp, world["rec"], rx, world["ligs"][pose_lig[p]], xyz[p])
        np.testing.assert_allclose(t[p], rt, rtol=1e-4, atol=1e-4)
        assert abs(s[p] - rs) <= 1e-4 * max(abs(rs), 1.0)


# ---------------------------------------------------------------- receptor leaf tables

def test_gpu_leaf_tables_match_direct_and_oracle(world):
    """Leaf energy tables (FP64 Hermite) vs direct pair-by-pair evaluation and
    vs the FP64 oracle on the same poses, including torsions far outside
    [-pi, pi] (the tables are periodic)."""
    nl, nodes, ee, ed = world["R"].table_info()
    print(f"\nleaf tables: {nl} leaves x {nodes} nodes, midpoint errors {ee:.2e} / {ed:.2e}")
    assert nl == 8 and nodes >= 2048 and ee <= 1e-8 and ed <= 1e-6
    pose_lig, xyz, tors = _poses(world, 6, 41, spread=2.0)
    tors = _f32(tors * 5.0)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e1, g1, t1 = [a.cpu().numpy() for a in lib.score_grad(world["R"], world["G"], P)]
    world["R"].set_leaf_tables(False)
    try:
        e0, g0, t0 = [a.cpu().numpy() for a in lib.score_grad(world["R"], world["G"], P)]
    finally:
        world["R"].set_leaf_tables(True)
    og = _oracle_grid(world)
    systems = [O.System(world["rec"], lig, og) for lig in world["ligs"]]
    for p in range(P.n_poses):
        re, rg, rt = systems[pose_lig[p]].energy(xyz[p], tors[p])
        assert _close_e(e1[p], re), (p, e1[p], e0[p], re)
        assert abs(e1[p] - e0[p]) <= 1e-5 * max(abs(e0[p]), 1.0), (p, e1[p], e0[p])
        o = P.pose_atom_off[p]
        n = world["ligs"][pose_lig[p]].n_atoms
        full1 = np.concatenate([g1.reshape(-1, 3)[o:o + n].ravel(), t1[p]])
        fullr = np.concatenate([rg.ravel(), rt])
        assert _grad_ok(full1, fullr), (p, np.abs(full1 - fullr).max())
        np.testing.assert_allclose(t1[p], t0[p], rtol=1e-4, atol=1e-3)


def test_gpu_leaf_tables_huge_torsions_stay_in_range(world):
    """Receptor torsions far outside any revolution (|theta| up to 3e38, the
    FP32 range, and either side of the 2^20 rad switch to the exact fmod
    reduction) and non-finite ones: the leaf-table index is reduced and
    clamped into the table, so finite angles give finite energies and
    gradients, and NaN / infinite angles give a NaN energy, without a fault."""
    pose_lig, xyz, tors = _poses(world, 1, 77, spread=2.0)
    n = len(pose_lig)
    big = np.array([1e17, -1e17, 1e30, -1e30, 3e38, -3e38, 2.0 ** 21 + 0.5, -(2.0 ** 20) - 3.0])
    tors = np.tile(big, (n, 1))[:, :8]
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e, g, t = [a.cpu().numpy() for a in lib.score_grad(world["R"], world["G"], P)]
    assert np.isfinite(e).all() and np.isfinite(g).all() and np.isfinite(t).all()
    bad = tors.copy()
    bad[:, 0] = np.nan
    bad[:, 1] = np.inf
    Pb = lib.PoseSet(B, pose_lig, xyz, 8, tors=bad)
    eb = lib.score_grad(world["R"], world["G"], Pb)[0].cpu().numpy()
    assert np.isnan(eb).all()


# ---------------------------------------------------------------- rigid-body + torsion variables

def test_gpu_dof_grad_matches_oracle(world):
    """Force, torque and torsion torques (pxd_dof_grad_batch) vs the oracle's
    chain rule (pinned by finite differences in test_dof_oracle.py)."""
    og = _oracle_grid(world)
    pose_lig, xyz, tors = _poses(world, 4, 51, spread=2.0)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    _, g, _ = lib.score_grad(world["R"], world["G"], P)
    dg = lib.dof_grad(P, g).cpu().numpy()
    gc = g.cpu().numpy().reshape(-1, 3)
    for p in range(P.n_poses):
        lig = world["ligs"][pose_lig[p]]
        o = P.pose_atom_off[p]
        gp = gc[o:o + lig.n_atoms].astype(np.float64)
        cen = xyz[p].mean(axis=0)
        ref = np.zeros(6 + lig.tree.n_torsions)
        tc = lig.tree.to_c()
        O.lib().or_dof_grad(O.byref(tc), lig.n_atoms, O.dp(O.f64(xyz[p])), O.dp(O.f64(gp)),
                            O.dp(O.f64(cen)), O.dp(ref))
        got = dg[p, :len(ref)]
        assert np.linalg.norm(got - ref) <= 1e-4 * max(np.linalg.norm(ref), 1.0), \
            (p, np.abs(got - ref).max())
        assert (dg[p, len(ref):] == 0).all()


def _rotvec(w):
    th = np.linalg.norm(w)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def _rigid_starts(world, seed, per=4):
    """Rigid placements of each ligand's own (clash-free) conformer: the
    torsion variables cannot undo internal clashes of random torsions."""
    rng = np.random.default_rng(seed)
    pose_lig, xyz = [], []
    for l, lig in enumerate(world["ligs"]):
        for _ in range(per):
            c = world["rec"].pocket_center + rng.normal(size=3) * 1.5
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R,
                                               np.zeros(lig.tree.n_torsions))))
            pose_lig.append(l)
    tors = _f32(rng.uniform(-0.5, 0.5, (len(xyz), world["rec"].leaves.n_leaves)))
    return pose_lig, xyz, tors


@pytest.mark.parametrize("iters,need", [(3, 0.9), (300, 0.75)])
def test_gpu_lbfgs_rigid_torsion_matches_oracle(world, iters, need):
    """L-BFGS over translation, rotation vector, ligand and receptor torsions
    (dof_mode 1) vs the oracle from the same starts, with a plain strong-Wolfe
    line search (gtd_threshold off: the reference's outright acceptance of
    steps with |g.d| > 1e5, optimizer.cpp:258-263, meets the large torques of
    rotation variables and sends both optimisers far off).  From rigid
    placements, after 3 iterations the runs have not parted (90 % agree to
    1e-4 relative).  Run to convergence (m = 8, 300 iterations,
    |g|inf <= 1e-3) from small perturbations of minima (0.2 A, 0.05 rad,
    0.05 rad per torsion), 75 % return to the same minimum within 1e-3
    kcal/mol (from clashing placements FP32 and FP64 runs settle in
    different minima).  Every run descends and keeps every bond length."""
    og = _oracle_grid(world)
    pose_lig, xyz, tors = _rigid_starts(world, 61)
    systems = [O.System(world["rec"], lig, og) for lig in world["ligs"]]
    if iters > 10:  # starts near minima of the oracle's own runs
        rng = np.random.default_rng(62)
        p0 = ffi.lbfgs_params(300, history=8, tol_grad=1e-3, dof_mode=1)
        p0.gtd_threshold = 1e300
        for p in range(len(xyz)):
            lig = world["ligs"][pose_lig[p]]
            m = systems[pose_lig[p]].lbfgs(xyz[p], tors[p], p0)
            c = m["xyz"].mean(axis=0) + rng.normal(size=3) * 0.2
            w = rng.normal(size=3) * 0.05
            xyz[p] = _f32(synth.ligand_apply(lig.tree, m["xyz"], c, _rotvec(w),
                                             rng.normal(size=lig.tree.n_torsions) * 0.05))
            tors[p] = _f32(m["tors"])
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    e0, _, _ = lib.score_grad(world["R"], world["G"], P)
    prm = ffi.lbfgs_params(iters, history=8, tol_grad=1e-3, dof_mode=1)
    prm.gtd_threshold = 1e300
    e, ne, cv = lib.lbfgs(world["R"], world["G"], P, prm)
    e, ne, e0, cv = e.cpu().numpy(), ne.cpu().numpy(), e0.cpu().numpy(), cv.cpu().numpy()
    agree, rows = 0, []
    for p in range(P.n_poses):
        lig = world["ligs"][pose_lig[p]]
        out = systems[pose_lig[p]].lbfgs(xyz[p], tors[p], prm)
        tol = 1e-4 * max(abs(out["energy"]), 1.0) if iters < 10 else \
            max(1e-3, 1e-4 * abs(out["energy"]))
        agree += abs(e[p] - out["energy"]) <= tol
        rows.append((p, e0[p], e[p], out["energy"], int(ne[p]), out["nevals"], int(cv[p])))
        assert np.isfinite(e[p]) and e[p] <= e0[p] + 1e-6 * abs(e0[p])
        x1 = P.xyz_of(p)
        ij = np.asarray(lig.bond_ij)
        b0 = np.linalg.norm(xyz[p][ij[:, 0]] - xyz[p][ij[:, 1]], axis=1)
        b1 = np.linalg.norm(x1[ij[:, 0]] - x1[ij[:, 1]], axis=1)
        assert np.abs(b1 - b0).max() <= 2e-3, np.abs(b1 - b0).max()
    print(f"\nrigid+torsion L-BFGS, {iters} iterations "
          "(pose, E0, E_gpu, E_oracle, nevals_gpu, nevals_oracle, converged_gpu):")
    for r in rows:
        print("  ", r)
    print(f"agree {agree}/{P.n_poses}")
    assert agree >= need * P.n_poses, (agree, P.n_poses)


def test_gpu_map_layouts_give_identical_results(world):
    """Map layouts 1 (cell corners together) and 2 (component planes) read the
    same values in the same order as layout 0, and 2 or 4 poses per workgroup
    evaluate each pose as one: energies and gradients are bit-identical.
    The pose count (23) is not a multiple of 2 or 4, so the last block holds
    waves without a pose."""
    pose_lig, xyz, tors = _poses(world, 6, 71)
    pose_lig, xyz, tors = pose_lig[:-1], xyz[:-1], tors[:-1]
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    assert P.n_poses % 2 != 0
    e0, g0, t0 = [a.cpu().numpy() for a in lib.score_grad(world["R"], world["G"], P)]
    try:
        for layout in (1, 2):
            world["G"].set_layout(layout)
            for ppb in (1, 2, 4):
                lib.set_tuning(8, 0, 8, ppb)
                e1, g1, t1 = [a.cpu().numpy() for a in lib.score_grad(world["R"], world["G"], P)]
                np.testing.assert_array_equal(e1, e0, err_msg=f"layout {layout} ppb {ppb}")
                np.testing.assert_array_equal(g1, g0, err_msg=f"layout {layout} ppb {ppb}")
                np.testing.assert_array_equal(t1, t0, err_msg=f"layout {layout} ppb {ppb}")
    finally:
        lib.set_tuning()
        world["G"].set_layout(0)


def test_gpu_grid_build_variants_identical_and_autotune(world):
    """4, 8 or 16 type maps per pass build bit-identical maps; pxd_autotune
    times every variant and keeps a valid choice."""
    maps = []
    try:
        for tpp in (4, 8, 16):
            lib.set_tuning(tpp, 0, 8)
            maps.append(lib.Grid(world["R"], world["box"], world["types"]).download())
    finally:
        lib.set_tuning()
    np.testing.assert_array_equal(maps[0], maps[1])
    np.testing.assert_array_equal(maps[0], maps[2])
    pose_lig, xyz, tors = _poses(world, 16, 81, spread=2.0)
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
    x_before = P.xyz.cpu().numpy().copy()
    try:
        t = lib.autotune(world["R"], world["box"], world["types"], P)
        d = t.as_dict()
        print("\nautotune:", d)
        assert d["grid_types_per_pass"] in (4, 8, 16) and d["map_layout"] in (0, 1, 2)
        assert d["lbfgs_slots_per_cu"] in (2, 4, 8, 16)
        assert d["score_poses_per_block"] in (1, 2, 4)
        assert all(v > 0 for v in d["grid_ms"].values())
        assert all(v > 0 for v in d["score_ms"].values())
        assert all(v > 0 for v in d["score_ppb_ms"].values())
        # every choice is the fastest of its measured variants
        assert d["score_ppb_ms"][d["score_poses_per_block"]] == min(d["score_ppb_ms"].values())
        assert d["score_ms"][d["map_layout"]] == min(d["score_ms"].values())
        assert all(v > 0 for v in d["lbfgs_ms"].values())
        np.testing.assert_array_equal(P.xyz.cpu().numpy(), x_before)  # inputs untouched
    finally:
        lib.set_tuning()


def test_gpu_lbfgs_stats_count_the_work(world):
    """pxd_lbfgs_last_stats: evaluations equal the per-pose counts; map bytes
    are 256 B per ligand atom and evaluation; iterations are at most the
    evaluations after each pose's first; the history traffic is a whole
    number of (8 n + 8)-byte entries, none with one iteration, and at most
    2 m + 1 entries per iteration."""
    l = 0  # poses of one ligand: one n for all of them
    lig = world["ligs"][l]
    pose_lig, xyz, tors = _poses(world, 12, 93, spread=2.0)
    keep = [i for i, pl in enumerate(pose_lig) if pl == l]
    pose_lig = [pose_lig[i] for i in keep]
    xyz = [xyz[i] for i in keep]
    tors = tors[keep]
    B = lib.LigandBatch(world["ligs"])
    n = 3 * lig.n_atoms + 8
    entry = 8 * n + 8
    for max_iter, m in ((1, 100), (30, 4), (60, 100)):
        P = lib.PoseSet(B, pose_lig, xyz, 8, tors=tors)
        e, ne, cv = lib.lbfgs(world["R"], world["G"], P, ffi.lbfgs_params(max_iter, history=m))
        st = lib.lbfgs_last_stats()
        ne = ne.cpu().numpy().astype(np.int64)
        print(f"\nmax_iter {max_iter} history {m}: {st}")
        assert st["evaluations"] == int(ne.sum())
        assert st["map_bytes"] == int(ne.sum()) * 256 * lig.n_atoms
        assert 0 <= st["iterations"] <= int(ne.sum()) - len(ne)
        assert st["history_bytes"] % entry == 0
        assert st["history_bytes"] <= st["iterations"] * (2 * m + 1) * entry
        if max_iter == 1:
            assert st["history_bytes"] == 0
        else:
            assert st["history_bytes"] > 0 and st["iterations"] > len(ne)


def test_gpu_lbfgs_large_batch_keeps_its_work_slots(world):
    """600,000 poses of one ligand (one size class, 2 L-BFGS iterations each):
    the work slots shared out among the size classes must not collapse (the
    share is computed in FP64: slots x poses x width passes 2^31 here), so
    the batch finishes within seconds, every pose evaluated, and equal
    starts give equal results."""
    import time
    import torch
    lig = world["ligs"][0]
    pose_lig, xyz, tors = _poses(world, 1, 97, spread=2.0)
    n = 600_000
    B = lib.LigandBatch(world["ligs"])
    P = lib.PoseSet(B, np.zeros(n, np.int32), [xyz[0]] * n, 8, tors=np.repeat(tors[:1], n, 0))
    torch.cuda.synchronize()
    t0 = time.time()
    e, ne, cv = lib.lbfgs(world["R"], world["G"], P, ffi.lbfgs_params(2))
    torch.cuda.synchronize()
    dt = time.time() - t0
    st = lib.lbfgs_last_stats()
    e, ne = e.cpu().numpy(), ne.cpu().numpy()
    print(f"\n{n} poses of {lig.n_atoms} atoms: {dt:.2f} s, {st}")
    assert dt < 10.0
    assert st["evaluations"] == int(ne.sum()) and (ne >= 1).all()
    assert np.all(e == e[0]) and np.all(ne == ne[0])
