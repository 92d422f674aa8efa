"""The large workload and the differentiation workload on the device against
the FP64 oracle.

Large workload: a synthetic receptor of 8,000 heavy atoms (9,294 with polar
H; the 2,000-atom recipe with the sphere radius scaled to keep the density),
a 40 A box at 0.375 A (109^3 points) built on the GPU, and screening ligands
of 10-60 heavy atoms with 0-20 torsions.  The maps sum ~9,300 receptor atoms
per point with heavy cancellation next to atoms, so the sampled points
include points 1.0-2.5 A from receptor atoms.

Differentiation workload: pose scores on the 2,000-heavy-atom receptor with
ligands of 20-50 heavy atoms, their coordinate and parameter derivatives
against FP64 central differences of the oracle's pscore at h = 1e-4.  A
coordinate or parameter whose difference quotient at h and at h / 2 disagree
straddles a ramp or clamp breakpoint (the pscore leaves are piecewise) and
is left out; at most a tenth may be.

Tolerances as in test_gpu_parity.py: energies 1e-4 max(|E|, 1), gradients
norm-wise 1e-4 max(||g||, 1), maps 1e-4 |v| + 1e-4; derivatives of the
score 1e-3 relative norm-wise.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import ffi, lib, scoring, synth, train

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def _large_ligand(i):
    rng = np.random.default_rng(50000 + i + 7919)
    nh = int(rng.integers(10, 61))
    nt = int(rng.integers(0, 21))
    nr = int(rng.integers(0, 4))
    return synth.make_ligand(50000 + i, n_heavy=nh, n_torsions=nt, n_rings=nr)


@pytest.fixture(scope="module")
def large():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(2, n_heavy=8000, radius=47.6, pocket_offset=24.0, n_ring6=80,
                              n_ring5=24)
    box = lib.decide_grid_dims(rec.pocket_center, [40.0] * 3, 0.375)
    assert box.npts == (109, 109, 109)
    ligs = [_large_ligand(i) for i in range(12)]
    types = sorted(set(int(t) for m in ligs for t in m.vdw_type))
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, types)
    return dict(rec=rec, box=box, ligs=ligs, R=R, G=G)


def test_gpu_large_grid_build_matches_oracle(large):
    box, rec = large["box"], large["rec"]
    maps = large["G"].download()
    ny, nz = box.npoints[1], box.npoints[2]
    P = box.n_points()
    rng = np.random.default_rng(8)
    pts = list(rng.integers(0, P, 200)) + [0, P - 1]
    # points next to receptor atoms inside the box, where the sums cancel most
    lo, hi = np.array(box.init_xyz), np.array(box.oppo_xyz)
    inside = np.where(((rec.xyz > lo + 3) & (rec.xyz < hi - 3)).all(axis=1))[0]
    for a in rng.choice(inside, 100, replace=False):
        q = rec.xyz[a] + synth._rand_unit(rng) * rng.uniform(1.0, 2.5)
        ijk = np.clip(np.rint((q - lo) / box.spacing).astype(int), 0, np.array(box.npoints) - 1)
        pts.append(int((ijk[0] * ny + ijk[1]) * nz + ijk[2]))
    print(f"\n[large grid] {len(pts)} points of {P}, {rec.n_atoms} receptor atoms", flush=True)
    bad = []
    for p in pts:
        i, j, k = p // (ny * nz), (p // nz) % ny, p % nz
        pt = lo + box.spacing * np.array([i, j, k])
        ref = O.grid_point(rec, pt, large["G"].types)
        got = maps[:, p, :]
        big = np.abs(ref) > 3e37
        ok = (np.abs(got - ref) <= 1e-4 * np.abs(ref) + 1e-4) | big
        if not ok.all():
            bad.append((p, float(np.abs(got - ref)[~ok].max())))
    assert not bad, bad[:5]


def test_gpu_large_score_grad_matches_oracle(large):
    rec, box = large["rec"], large["box"]
    og = O.Grid(box, large["G"].types, large["G"].download().astype(np.float64), 1e6,
                rec.leaves.movable())
    rng = np.random.default_rng(12)
    lo, hi = np.array(box.init_xyz), np.array(box.oppo_xyz)
    pose_lig, xyz = [], []
    for l, lig in enumerate(large["ligs"]):
        for k in range(6):
            c = rng.uniform(lo, hi) if k % 2 else rec.pocket_center + rng.normal(0, 2.0, 3)
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)))
            pose_lig.append(l)
    nl = rec.leaves.n_leaves
    tors = _f32(rng.uniform(-math.pi, math.pi, (len(xyz), nl)))
    B = lib.LigandBatch(large["ligs"])
    P = lib.PoseSet(B, pose_lig, xyz, nl, tors=tors)
    e, g, tg = lib.score_grad(large["R"], large["G"], P)
    e, tg = e.cpu().numpy(), tg.cpu().numpy()
    systems = [O.System(rec, lig, og) for lig in large["ligs"]]
    bad = []
    for p in range(P.n_poses):
        re, rg, rtg = systems[pose_lig[p]].energy(xyz[p], tors[p])
        full = np.concatenate([P.xyz_of(p, g).ravel(), tg[p]])
        rfull = np.concatenate([rg.ravel(), rtg])
        if not (abs(e[p] - re) <= 1e-4 * max(abs(re), 1.0) and
                np.linalg.norm(full - rfull) <= 1e-4 * max(np.linalg.norm(rfull), 1.0)):
            bad.append((p, e[p], re, np.linalg.norm(full - rfull), np.linalg.norm(rfull)))
    print(f"\n[large score] {P.n_poses} poses, {len(bad)} outside tolerance", flush=True)
    assert not bad, bad[:5]


# ---------------------------------------------------------------- differentiation workload

def _train_ligand(i):
    rng = np.random.default_rng(40000 + i + 7919)
    nh = int(rng.integers(20, 51))
    nt = int(rng.integers(0, 16))
    nr = int(rng.integers(0, 4))
    return synth.make_ligand(40000 + i, n_heavy=nh, n_torsions=nt, n_rings=nr)


@pytest.fixture(scope="module")
def c3():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(1)
    ligs = [_train_ligand(i) for i in range(4)]
    rng = np.random.default_rng(41)
    pose_lig, xyz = [], []
    for l, lig in enumerate(ligs):
        for _ in range(2):
            c = rec.pocket_center + rng.uniform(-6.0, 6.0, 3)
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)))
            pose_lig.append(l)
    nl = rec.leaves.n_leaves
    B = lib.LigandBatch(ligs)
    P = lib.PoseSet(B, pose_lig, xyz, nl)
    return dict(rec=rec, ligs=ligs, pose_lig=pose_lig, xyz=xyz, R=lib.Receptor(rec), B=B, P=P,
                sp=scoring.load_pose_selection())


def _score(c3, p, x, sp):
    return O.pscore(sp, c3["rec"], c3["rec"].xyz, c3["ligs"][c3["pose_lig"][p]], x)[0]


def _fd(f, v, i, h):
    vp, vm = v.copy(), v.copy()
    vp.flat[i] += h
    vm.flat[i] -= h
    return (f(vp) - f(vm)) / (2.0 * h)


def _fd_checked(f, v, idx, h=1e-4):
    """Central differences at h; entries whose quotient at h / 2 differs by
    more than 1e-3 of the larger are marked (a breakpoint lies within h)."""
    d1 = np.array([_fd(f, v, i, h) for i in idx])
    d2 = np.array([_fd(f, v, i, h / 2) for i in idx])
    smooth = np.abs(d1 - d2) <= 1e-3 * np.maximum(np.abs(d1), 1.0) + 1e-6
    return d2, smooth


def test_gpu_score_derivatives_on_the_training_workload(c3):
    s, dx, dp = lib.pscore_grad(c3["R"], c3["P"], c3["sp"])
    s, dx, dp = s.cpu().numpy(), dx.cpu().numpy().reshape(-1, 3), dp.cpu().numpy()
    theta = train.params_to_vector(c3["sp"])
    types = sorted(set(int(t) for m in c3["ligs"] for t in m.vdw_type))
    pidx = list(range(ffi.NUM_TERMS)) + [ffi.NUM_TERMS + t for t in types] + \
        [ffi.NUM_TERMS + ffi.NUM_VDW_TYPES + t for t in types]
    rng = np.random.default_rng(3)
    n_skip = n_all = 0
    for p in range(c3["P"].n_poses):
        x = c3["xyz"][p]
        ref = _score(c3, p, x, c3["sp"])
        assert abs(s[p] - ref) <= 1e-4 * max(abs(ref), 1.0)
        # coordinates: 8 atoms x 3 per pose
        atoms = rng.choice(len(x), min(8, len(x)), replace=False)
        cidx = [3 * a + d for a in atoms for d in range(3)]
        fd, ok = _fd_checked(lambda v: _score(c3, p, v.reshape(-1, 3), c3["sp"]), x.ravel(), cidx)
        o = c3["P"].pose_atom_off[p]
        got = dx[o:o + len(x)].ravel()[cidx]
        err = np.linalg.norm((got - fd)[ok])
        assert err <= 1e-3 * max(np.linalg.norm(fd[ok]), 1.0), (p, err)
        # parameters: the 9 weights and the scales of the types in use
        fdp, okp = _fd_checked(
            lambda v: _score(c3, p, x, train.vector_to_params(v, c3["sp"])), theta, pidx)
        errp = np.linalg.norm((dp[p, pidx] - fdp)[okp])
        assert errp <= 1e-3 * max(np.linalg.norm(fdp[okp]), 1.0), (p, errp)
        n_skip += int((~ok).sum() + (~okp).sum())
        n_all += len(ok) + len(okp)
    print(f"\n[train fd] {c3['P'].n_poses} poses, {n_all} derivatives, {n_skip} at breakpoints",
          flush=True)
    assert n_skip <= 0.1 * n_all
