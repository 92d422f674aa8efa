"""The reference-side adaptors (protenix-dock_amd/cpp/pxd_reference_adaptors.hpp)
against the Python reader.

oracle/_ref/ref_adaptors is compiled by oracle/Makefile against the
reference's own standard-library headers (system.h, data.h).  It fills the
reference's structs (force_field_params, fragment_tree, rotatable_leaf) from
a receptor or ligand JSON the way data.cpp does, converts them with the
adaptors, and prints the pxd descriptors; they must equal what
pxdock_amd/data.py builds from the same file.  Skipped where the reference
tree was not available to build it.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from pxdock_amd import data, synth

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_adaptors")


@pytest.fixture(autouse=True)
def _built():
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/ref_adaptors not built")


def _run(kind, path):
    out = subprocess.run([EXE, kind, path], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def _flat(a):
    return np.asarray(a).ravel().tolist()


def _check_mol(got, m):
    assert got["n_atoms"] == m.n_atoms
    if m.is_ligand:
        assert got["vdw_type"] == _flat(m.vdw_type)
    else:
        np.testing.assert_array_equal(got["sigma"], _flat(m.sigma))
        np.testing.assert_array_equal(got["epsilon"], _flat(m.epsilon))
    np.testing.assert_array_equal(got["charge"], _flat(m.charge))
    for k_c, k_py in [("bond_ij", "bond_ij"), ("angle_ijk", "angle_ijk"), ("dih_ijkl", "dih_ijkl"),
                      ("pairs14", "pairs14"), ("pairs", "pairs"), ("hydrophobic", "hydrophobic"),
                      ("cation", "cation"), ("anion", "anion"), ("ring5", "ring5"),
                      ("ring6", "ring6"), ("don_charged", "don_charged"),
                      ("don_neutral", "don_neutral"), ("acc_charged", "acc_charged"),
                      ("acc_neutral", "acc_neutral")]:
        assert got[k_c] == _flat(getattr(m, k_py)), k_c
    for k in ("bond_k", "bond_r0", "angle_k", "angle_t0"):
        np.testing.assert_allclose(got[k], _flat(getattr(m, k)), rtol=1e-15, atol=0, err_msg=k)
    # dihedral series: equal up to the first periodicity 0 (interaction.cpp:88)
    n = np.asarray(got["dih_n"]).reshape(-1, 6)
    live = np.cumprod(n != 0, axis=1).astype(bool)
    np.testing.assert_array_equal(n, np.where(live, m.dih_n, 0))
    for k in ("dih_phase", "dih_k"):
        np.testing.assert_allclose(np.asarray(got[k]).reshape(-1, 6),
                                   np.where(live, getattr(m, k), 0.0), rtol=1e-15, atol=0)
    if m.is_ligand:
        np.testing.assert_array_equal(got["tstrain_coef"], _flat(m.tstrain_coef))
        assert got["tstrain_off"] == _flat(m.tstrain_off)
        assert got["tstrain_ijkl"] == _flat(m.tstrain_ijkl)
        assert got["n_rotatable"] == len(m.rotatable_bonds)


def _check_tree(got, t):
    assert got["perm"] == _flat(t.perm)
    assert got["frag_start"] == _flat(t.frag_start)
    assert got["frag_end"] == _flat(t.frag_end)
    assert got["frag_parent"] == _flat(t.frag_parent)
    # a root's bond slots are unused
    bond = np.asarray(got["frag_bond"]).reshape(-1, 2)
    nonroot = np.asarray(t.frag_parent) >= 0
    np.testing.assert_array_equal(bond[nonroot], np.asarray(t.frag_bond)[nonroot])
    assert got["tors_frag"] == _flat(t.tors_frag)


def test_golden_receptor_adapts_like_the_python_reader():
    path = os.path.join(GOLDEN, "receptor_12.json")
    got = _run("receptor", path)
    m = data.parse_receptor(path)
    _check_mol(got["mol"], m)
    L = m.leaves
    np.testing.assert_allclose(got["leaves"]["origin"], _flat(L.origin), rtol=0, atol=0)
    np.testing.assert_allclose(got["leaves"]["axis"], _flat(L.axis), rtol=1e-15, atol=1e-16)
    assert got["leaves"]["atom_off"] == _flat(L.atom_off)
    assert got["leaves"]["atoms"] == _flat(L.atoms)


def test_golden_ligand_tree_adapts_like_the_python_reader():
    path = os.path.join(GOLDEN, "ligand_27.json")
    got = _run("ligand", path)
    _check_tree(got["tree"], data.parse_ligand(path).tree)


@pytest.mark.parametrize("seed,heavy,tors,rings", [(71, 20, 4, 1), (72, 45, 11, 3),
                                                     (73, 60, 15, 2)])
def test_synthetic_ligands_adapt_like_the_python_reader(tmp_path, seed, heavy, tors, rings):
    lig = synth.make_ligand(seed, n_heavy=heavy, n_torsions=tors, n_rings=rings)
    path = tmp_path / "lig.json"
    path.write_text(json.dumps(data.ligand_json(lig)))
    got = _run("ligand", str(path))
    m = data.parse_ligand(str(path))
    _check_mol(got["mol"], m)
    _check_tree(got["tree"], m.tree)


def test_synthetic_receptor_adapts_like_the_python_reader(tmp_path):
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    path = tmp_path / "rec.json"
    path.write_text(json.dumps(data.receptor_json(rec)))
    got = _run("receptor", str(path))
    m = data.parse_receptor(str(path))
    _check_mol(got["mol"], m)
    assert got["leaves"]["atoms"] == _flat(m.leaves.atoms)
    np.testing.assert_allclose(got["leaves"]["axis"], _flat(m.leaves.axis), rtol=1e-14, atol=1e-15)
