"""pxd_bscore_batch on the device against the FP64 oracle's bscore (itself
equal to the reference's compiled scorer, tests/test_bscore.py), on the
2,000-heavy-atom receptor with its 8 rotatable hydroxyls at random torsions
and screening ligands (10-60 heavy atoms) near the pocket.

Tolerances: the leaf terms are the pscore kernel's (FP32 per-pair
arithmetic): 1e-4 max(|term|, 1) each; the three GBSA energies are FP64 from
the same FP64 inputs (the pose coordinates are FP32 values, given to the
oracle exactly): 1e-9 relative; the score within the sum of both."""
import math

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import ffi, lib, scoring, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(1)
    ligs = []
    for i in range(6):
        rng = np.random.default_rng(60000 + i)
        ligs.append(synth.make_ligand(60000 + i, n_heavy=int(rng.integers(10, 61)),
                                      n_torsions=int(rng.integers(0, 12)),
                                      n_rings=int(rng.integers(0, 4))))
    rng = np.random.default_rng(61)
    pose_lig, xyz = [], []
    for l, lig in enumerate(ligs):
        for _ in range(2):
            c = rec.pocket_center + rng.normal(0.0, 1.5, 3)
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            x = synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)
            xyz.append(np.asarray(np.asarray(x, np.float32), np.float64))
            pose_lig.append(l)
    nl = rec.leaves.n_leaves
    tors = np.asarray(rng.uniform(-math.pi, math.pi, (len(xyz), nl)).astype(np.float32),
                      np.float64)
    B = lib.LigandBatch(ligs)
    P = lib.PoseSet(B, pose_lig, xyz, nl, tors=tors)
    return dict(rec=rec, ligs=ligs, pose_lig=pose_lig, xyz=xyz, tors=tors, R=lib.Receptor(rec),
                P=P, bp=scoring.load_affinity_ranking())


def test_gpu_bscore_matches_the_oracle(world):
    s, t = lib.bscore(world["R"], world["P"], world["bp"])
    s, t = s.cpu().numpy(), t.cpu().numpy()
    rec = world["rec"]
    rec_gb = O.gbsa_atom_params(rec)
    print(f"\n[bscore] {len(s)} poses, receptor {rec.n_atoms} atoms", flush=True)
    for p in range(len(s)):
        lig = world["ligs"][world["pose_lig"][p]]
        rx = O.receptor_apply(rec.leaves, rec.xyz, world["tors"][p])
        ref, rt = O.bscore(world["bp"], rec, rx, lig, world["xyz"][p], rec_gb=rec_gb)
        leaf_tol = 1e-4 * np.maximum(np.abs(rt[:ffi.NUM_TERMS]), 1.0)
        assert (np.abs(t[p, :ffi.NUM_TERMS] - rt[:ffi.NUM_TERMS]) <= leaf_tol).all(), p
        assert np.allclose(t[p, ffi.NUM_TERMS:], rt[ffi.NUM_TERMS:], rtol=1e-9, atol=0.0), \
            (p, t[p, ffi.NUM_TERMS:], rt[ffi.NUM_TERMS:])
        coefs = scoring.params_to_vector(world["bp"].leaves)[:ffi.NUM_TERMS]
        tol = float(np.abs(coefs) @ leaf_tol) + 1e-9 * float(np.abs(rt[ffi.NUM_TERMS:]).sum()) * 3
        assert abs(s[p] - ref) <= tol, (p, s[p], ref)


def test_gpu_bscore_is_deterministic_and_gbsa_off_is_the_leaves(world):
    s1, t1 = lib.bscore(world["R"], world["P"], world["bp"])
    s2, t2 = lib.bscore(world["R"], world["P"], world["bp"])
    assert s1.cpu().numpy().tobytes() == s2.cpu().numpy().tobytes()
    assert t1.cpu().numpy().tobytes() == t2.cpu().numpy().tobytes()
    off = ffi.PxdBscoreParams()
    off.leaves = world["bp"].leaves
    off.gbsa_coef = 0.0
    s0, t0 = lib.bscore(world["R"], world["P"], off)
    ps, _ = lib.pscore(world["R"], world["P"], world["bp"].leaves)
    assert np.array_equal(s0.cpu().numpy(), ps.cpu().numpy())
    assert (t0.cpu().numpy()[:, ffi.NUM_TERMS:] == 0).all()
