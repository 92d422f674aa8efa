"""The C-ABI library loads on a machine without a GPU and exports every entry
point that include/*.h declares (no compute calls here)."""
import ctypes as C
import glob
import os
import re

import pytest

from conftest import ROOT
from pxdock_amd import lib


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(pxd_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing


def test_error_channel_and_host_helpers():
    L = lib.load()
    assert L.pxd_version().decode().startswith("pxdock")
    n = C.c_int32(-1)
    assert L.pxd_device_count(C.byref(n)) == 0 and n.value >= 0
    # argument validation happens before any device work
    assert L.pxd_receptor_create(None, None, None, None) == 1
    assert b"null" in L.pxd_last_error()
    b = lib.decide_grid_dims([0, 0, 0], [22.5] * 3, 0.375)
    assert b.npts == (61, 61, 61)
    b = lib.decide_grid_dims([1, 2, 3], [40.0] * 3, 0.375)
    assert b.npts == (109, 109, 109)  # 107 voxels rounded up to even (grid.cpp:78)
    assert abs(b.init_xyz[0] - (1 - 108 * 0.375 / 2)) < 1e-12


def test_tuning_defaults_and_validation():
    t = lib.get_tuning()
    assert (t.grid_types_per_pass, t.map_layout, t.lbfgs_slots_per_cu,
            t.score_poses_per_block) == (8, 0, 8, 1)
    for bad in ({"grid_types_per_pass": 5}, {"map_layout": 3}, {"map_layout": -1},
                {"lbfgs_slots_per_cu": 0}, {"score_poses_per_block": 3},
                {"score_poses_per_block": 8}):
        with pytest.raises(lib.PxdError):
            lib.set_tuning(**bad)
    lib.set_tuning(16, 2, 4, 4)
    t = lib.get_tuning()
    assert (t.grid_types_per_pass, t.map_layout, t.lbfgs_slots_per_cu,
            t.score_poses_per_block) == (16, 2, 4, 4)
    lib.set_tuning()


@pytest.mark.parametrize("field,value,msg", [
    ("tors_frag", 0, b"not a child fragment"),      # the root has no torsion
    ("tors_frag", 99, b"not a child fragment"),
    ("frag_bond", 10_000, b"bond atom out of range"),
])
def test_fragment_tree_is_validated_before_device_work(field, value, msg):
    from pxdock_amd import synth
    lig = synth.make_ligand(3, n_heavy=20, n_torsions=4, n_rings=1)
    getattr(lig.tree, field)[-1] = value
    with pytest.raises(lib.PxdError) as e:
        lib.LigandBatch([lig])
    assert msg in str(e.value).encode()


def _tree_ligand():
    from pxdock_amd import synth
    lig = synth.make_ligand(3, n_heavy=20, n_torsions=4, n_rings=1)
    assert lig.tree.n_torsions >= 2
    return lig


def _bad_trees():
    """(name, mutate(tree), expected message) for trees the ABI must reject."""
    import numpy as np

    def perm_dup(t):
        t.perm[1] = t.perm[0]

    def perm_range(t):
        t.perm[0] = len(t.perm)

    def drop_torsion(t):
        t.tors_frag = t.tors_frag[:-1].copy()

    def dup_torsion(t):
        t.tors_frag[-1] = t.tors_frag[0]

    def child_first(t):
        # a grandchild listed before its parent: pick a fragment whose parent
        # is not a root and move it to the front
        order = list(t.tors_frag)
        for k, f in enumerate(order):
            if t.frag_parent[f] >= 0 and t.frag_parent[t.frag_parent[f]] >= 0:
                order.insert(0, order.pop(k))
                t.tors_frag = np.array(order, dtype=t.tors_frag.dtype)
                return
        # no grandchild: make the last child fragment hang off the first child
        # and list it first
        a, b = order[0], order[-1]
        t.frag_parent[b] = a
        t.frag_bond[b][0] = t.frag_start[a]
        order.remove(b)
        t.tors_frag = np.array([b] + order, dtype=t.tors_frag.dtype)

    def overlap(t):
        t.frag_end[0] = t.frag_end[0] + 1

    def bond_wrong_side(t):
        f = int(t.tors_frag[0])
        t.frag_bond[f][1] = t.frag_start[int(t.frag_parent[f])]

    return [("perm_dup", perm_dup, b"not a permutation"),
            ("perm_range", perm_range, b"not a permutation"),
            ("drop_torsion", drop_torsion, b"torsions for"),
            ("dup_torsion", dup_torsion, b"listed twice"),
            ("child_first", child_first, b"precedes its parent"),
            ("overlap", overlap, b"overlap"),
            ("bond_wrong_side", bond_wrong_side, b"does not join")]


@pytest.mark.parametrize("name,mutate,msg", _bad_trees(), ids=[b[0] for b in _bad_trees()])
def test_malformed_fragment_trees_are_rejected(name, mutate, msg):
    lig = _tree_ligand()
    try:  # the unmodified tree passes validation (no device here: a HIP error)
        lib.LigandBatch([lig])
    except lib.PxdError as e:
        assert e.status == 3, str(e)
    mutate(lig.tree)
    with pytest.raises(lib.PxdError) as e:
        lib.LigandBatch([lig])
    assert msg in str(e.value).encode(), str(e.value)
