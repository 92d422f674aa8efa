"""Host-side logic: reference JSON round trips, the scoring-config loader,
the synthetic generator and the numpy pose parameterisation."""
import json
import math

import numpy as np
import pytest

import pyoracle as O
from conftest import golden_path
from pxdock_amd import data, ffi, scoring, synth


def test_receptor_json_roundtrip_and_leaves():
    rec = data.parse_receptor(golden_path("receptor_12.json"))
    assert rec.leaves.n_leaves == 3
    # movable = fragment atoms minus the axis atom (data.cpp:359-361)
    assert rec.leaves.atoms.tolist() == [3, 6, 9, 10, 11]
    r2 = data.parse_receptor(json.loads(json.dumps(data.receptor_json(rec))))
    np.testing.assert_allclose(r2.leaves.axis, rec.leaves.axis)


def test_ligand_json_tree_matches_reference_traversal():
    lig = data.parse_ligand(golden_path("ligand_27.json"))
    t = lig.tree
    assert t.n_frags == 6 and t.n_torsions == 5
    assert t.tors_frag.tolist() == [1, 2, 3, 4, 5]  # pre-order 0->1->{2,3}, 0->4->5
    l2 = data.parse_ligand(json.loads(json.dumps(data.ligand_json(lig))))
    assert (l2.tree.frag_bond == t.frag_bond).all()


def test_bad_vdw_type_is_rejected():
    obj = json.load(open(golden_path("ligand_3.json")))
    obj["ffdata"]["FF_vdW_paraidx"] = [1019, 1099, 1016]
    with pytest.raises(data.ConfError):
        data.parse_ligand(obj)


def test_receptor_atom_count_mismatch_is_rejected():
    obj = json.load(open(golden_path("receptor_12.json")))
    obj["geometry"]["num_atoms"] = 11
    with pytest.raises(data.ConfError):
        data.parse_receptor(obj)


def test_default_scoring_config():
    sp = scoring.load_pose_selection()
    assert sp.hbond_coef == pytest.approx(1.0142641067504883)
    assert sp.vdw_cutoff == pytest.approx(1.3484539985656738)
    assert sp.hbond_inter[0][4] > 0  # dist_center present
    assert len(scoring.params_to_vector(sp)) == 9 + 2 * 99


@pytest.mark.parametrize("seed", range(6))
def test_synthetic_ligand_is_valid(seed):
    lig = synth.make_ligand(seed, n_heavy=10 + 9 * seed, n_torsions=3 * seed, n_rings=seed % 4)
    n = lig.n_atoms
    assert n <= ffi.MAX_LIGAND_ATOMS
    d = np.linalg.norm(lig.xyz[0][lig.bond_ij[:, 0]] - lig.xyz[0][lig.bond_ij[:, 1]], axis=1)
    assert (d > 0.9).all() and (d < 1.6).all()
    # pairs exclude 1-2 / 1-3 and are unique
    ex = {tuple(sorted(b)) for b in lig.bond_ij.tolist()} | \
         {tuple(sorted((a[0], a[2]))) for a in lig.angle_ijk.tolist()}
    allp = [tuple(p) for p in lig.pairs.tolist()] + [tuple(p) for p in lig.pairs14.tolist()]
    assert len(set(allp)) == len(allp) and not (set(allp) & ex)
    assert abs(lig.charge.sum()) < 1e-9
    # torsions rotate rigid fragments: bond lengths are invariant
    rng = np.random.default_rng(seed)
    x = synth.ligand_apply(lig.tree, lig.xyz[0], rng.normal(size=3),
                           synth.quat_to_rot(synth.random_quaternion(rng)),
                           rng.uniform(-math.pi, math.pi, lig.tree.n_torsions))
    d2 = np.linalg.norm(x[lig.bond_ij[:, 0]] - x[lig.bond_ij[:, 1]], axis=1)
    np.testing.assert_allclose(d, d2, atol=1e-9)


def test_numpy_apply_equals_oracle_apply():
    lig = synth.make_ligand(99, n_heavy=40, n_torsions=9, n_rings=2)
    rng = np.random.default_rng(1)
    for _ in range(5):
        c = rng.normal(size=3)
        R = synth.quat_to_rot(synth.random_quaternion(rng))
        t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
        a = synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)
        b = O.ligand_apply(lig.tree, lig.xyz[0], c, R.ravel(), t)
        np.testing.assert_allclose(a, b, atol=1e-12)


def test_synthetic_receptor_layout():
    rec = synth.make_receptor(1)
    heavy = (rec.atomic_numbers > 1).sum()
    assert heavy == 2000 and rec.leaves.n_leaves == 8
    # every non-bonded receptor pair touches a movable atom (receptor.py:230-265)
    mov = set(rec.leaves.atoms.tolist())
    assert all((a in mov or b in mov) for a, b in rec.pairs.tolist())
    assert np.linalg.norm(rec.xyz - rec.pocket_center, axis=1).min() > 4.0
