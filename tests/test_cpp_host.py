"""The C++ host layer (protenix-dock_amd/cpp/pxdock.hpp): its host-only cases
(reference JSON readers, receptor torsions against the gtest vector), and
its receptor / ligand readers against the Python readers on the same files."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden_path
from pxdock_amd import data

BIN = os.path.join(ROOT, "protenix-dock_amd", "cpp", "_build", "test_host")


@pytest.fixture(scope="module")
def binary():
    subprocess.run(["make", "-s", "-C", os.path.dirname(os.path.dirname(BIN))], check=True)
    assert os.path.exists(BIN)
    return BIN


def test_cpp_host_cpu_cases(binary):
    out = subprocess.run([binary, GOLDEN, "--cpu"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failures" in out.stdout


@pytest.mark.parametrize("rec,lig", [("receptor_12.json", "ligand_3.json"),
                                     ("receptor_12.json", "ligand_27.json")])
def test_cpp_readers_match_python_readers(binary, rec, lig):
    out = subprocess.run([binary, GOLDEN, "--dump", golden_path(rec), golden_path(lig)],
                         capture_output=True, text=True, check=True)
    d = json.loads(out.stdout)
    r = data.parse_receptor(golden_path(rec))
    m = data.parse_ligand(golden_path(lig))
    assert d["rec_atoms"] == r.n_atoms and d["rec_leaves"] == r.leaves.n_leaves
    assert d["rec_bonds"] == len(r.bond_ij) and d["rec_pairs"] == len(r.pairs)
    assert d["lig_atoms"] == m.xyz.shape[1] and d["lig_poses"] == len(m.xyz)
    assert d["lig_torsions"] == m.tree.n_torsions and d["lig_frags"] == m.tree.n_frags
    assert d["lig_dihedrals"] == len(m.dih_ijkl)
    assert d["lig_pairs14"] == len(m.pairs14) and d["lig_pairs"] == len(m.pairs)
    assert d["tors_frag"] == list(m.tree.tors_frag) and d["perm"] == list(m.tree.perm)
    cs = sum(float(v) * (i + 1) + float(k) for i, (v, k) in
             enumerate(zip(np.ravel(m.dih_phase), np.ravel(m.dih_k)))) + float(np.sum(m.angle_t0))
    assert d["lig_checksum"] == pytest.approx(cs, rel=1e-14, abs=1e-14)
