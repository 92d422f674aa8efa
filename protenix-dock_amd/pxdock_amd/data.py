"""Receptor / ligand data model and the reference's JSON wire format.

Readers follow torsional_receptor::parse and free_ligand::parse
(engine/src/core/data.cpp:246-276, 292-364, 410-520); writers emit the same
keys so synthetic inputs are valid reference inputs.  Arrays are numpy; the
`to_c()` methods build the pxdock.h descriptors (ffi.PxdMol, PxdFragTree,
PxdLeaves) consumed by both the GPU library and the oracle.
"""
import ctypes as C
import json
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import ffi
from .ffi import as_f64, as_i32


class ConfError(ValueError):
    """Malformed input data (failed_conf_error, engine/include/bytedock/lib/error.h:38-43)."""


def _deg2rad(x):
    # MATH_TO_RADIAN (math.h:229-230)
    return np.asarray(x, dtype=np.float64) / 180.0 * math.pi


@dataclass
class FragTree:
    """Rigid-fragment tree of a ligand in flat form (pxd_frag_tree)."""
    perm: np.ndarray          # [n] permuted slot -> original atom
    frag_start: np.ndarray
    frag_end: np.ndarray
    frag_parent: np.ndarray   # -1 for roots
    frag_bond: np.ndarray     # [F,2] (parent atom, mine atom) as permuted slots
    tors_frag: np.ndarray     # non-root fragments in the reference pre-order

    @property
    def n_frags(self):
        return len(self.frag_start)

    @property
    def n_torsions(self):
        return len(self.tors_frag)

    def to_c(self):
        keep = [as_i32(self.perm), as_i32(self.frag_start), as_i32(self.frag_end),
                as_i32(self.frag_parent), as_i32(self.frag_bond), as_i32(self.tors_frag)]
        t = ffi.PxdFragTree()
        t.n_frags = self.n_frags
        t.perm = ffi.ptr(keep[0], C.c_int32)
        t.frag_start = ffi.ptr(keep[1], C.c_int32)
        t.frag_end = ffi.ptr(keep[2], C.c_int32)
        t.frag_parent = ffi.ptr(keep[3], C.c_int32)
        t.frag_bond = ffi.ptr(keep[4], C.c_int32)
        t.n_torsions = self.n_torsions
        t.tors_frag = ffi.ptr(keep[5], C.c_int32)
        t._keep = keep
        return t

    def moved_atoms(self, t):
        """Original atom indices moved by torsion t (its fragment's subtree)."""
        f0 = int(self.tors_frag[t])
        out = []
        for f in range(self.n_frags):
            g = f
            while g >= 0 and g != f0:
                g = int(self.frag_parent[g])
            if g == f0:
                out.extend(int(self.perm[p]) for p in range(self.frag_start[f], self.frag_end[f]))
        return sorted(out)


@dataclass
class Leaves:
    """Receptor rotatable leaves (torsional_receptor::parse, data.cpp:348-363)."""
    origin: np.ndarray        # [L,3]
    axis: np.ndarray          # [L,3] unit
    atom_off: np.ndarray      # [L+1]
    atoms: np.ndarray         # movable atoms
    axis_atoms: np.ndarray = None   # [L,2] (rot_bond_1, rot_bond_2)
    fragments: list = None          # raw fragment atom lists (for writing back)
    bond_order: np.ndarray = None   # rot_bond_order per leaf
    keys: list = None               # geometry.data keys per leaf

    @property
    def n_leaves(self):
        return len(self.origin)

    def movable(self):
        return np.unique(self.atoms) if len(self.atoms) else np.zeros(0, np.int32)

    def to_c(self):
        keep = [as_f64(self.origin), as_f64(self.axis), as_i32(self.atom_off), as_i32(self.atoms)]
        L = ffi.PxdLeaves()
        L.n_leaves = self.n_leaves
        L.origin = ffi.ptr(keep[0], C.c_double)
        L.axis = ffi.ptr(keep[1], C.c_double)
        L.atom_off = ffi.ptr(keep[2], C.c_int32)
        L.atoms = ffi.ptr(keep[3], C.c_int32)
        L._keep = keep
        return L


_I2 = ("bond_ij", "pairs14", "pairs", "don_charged", "don_neutral")


@dataclass
class Molecule:
    """force_field_params (engine/include/bytedock/core/system.h:85-109) plus coordinates."""
    xyz: np.ndarray                       # receptor [n,3]; ligand poses [P,n,3]
    charge: np.ndarray
    vdw_type: Optional[np.ndarray] = None  # ligand GAFF2 index
    sigma: Optional[np.ndarray] = None     # receptor
    epsilon: Optional[np.ndarray] = None
    bond_ij: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angle_ijk: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    angle_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angle_t0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    n_impropers: int = 0
    dih_ijkl: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    dih_n: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.int32))
    dih_phase: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    dih_k: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    pairs14: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    hydrophobic: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    cation: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    anion: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ring5: np.ndarray = field(default_factory=lambda: np.zeros((0, 5), np.int32))
    ring6: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.int32))
    don_charged: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    don_neutral: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    acc_charged: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    acc_neutral: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    tstrain_coef: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    tstrain_off: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    tstrain_ijkl: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    rotatable_bonds: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    all_bonds: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    atomic_numbers: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    molecular_weight: float = -1.0
    smiles: str = ""
    tree: Optional[FragTree] = None
    leaves: Optional[Leaves] = None

    @property
    def n_atoms(self):
        return len(self.charge)

    @property
    def is_ligand(self):
        return self.vdw_type is not None

    def to_c(self):
        """pxd_mol descriptor; the returned struct keeps its arrays alive."""
        m = ffi.PxdMol()
        keep = []

        def i32(name, width=None):
            a = as_i32(getattr(self, name), width)
            keep.append(a)
            return ffi.ptr(a, C.c_int32), (len(a) if a.ndim else 0)

        def f64(name):
            v = getattr(self, name)
            if v is None:
                return None
            a = as_f64(v)
            keep.append(a)
            return ffi.ptr(a, C.c_double)

        m.n_atoms = self.n_atoms
        if self.is_ligand:
            m.vdw_type, _ = i32("vdw_type")
        else:
            m.sigma = f64("sigma")
            m.epsilon = f64("epsilon")
        m.charge = f64("charge")
        m.bond_ij, m.n_bonds = i32("bond_ij", 2)
        m.bond_k = f64("bond_k")
        m.bond_r0 = f64("bond_r0")
        m.angle_ijk, m.n_angles = i32("angle_ijk", 3)
        m.angle_k = f64("angle_k")
        m.angle_t0 = f64("angle_t0")
        m.dih_ijkl, m.n_dihedrals = i32("dih_ijkl", 4)
        m.dih_n, _ = i32("dih_n", 6)
        m.dih_phase = f64("dih_phase")
        m.dih_k = f64("dih_k")
        m.pairs14, m.n_pairs14 = i32("pairs14", 2)
        m.pairs, m.n_pairs = i32("pairs", 2)
        m.hydrophobic, m.n_hydrophobic = i32("hydrophobic")
        m.cation, m.n_cation = i32("cation")
        m.anion, m.n_anion = i32("anion")
        m.ring5, m.n_ring5 = i32("ring5", 5)
        m.ring6, m.n_ring6 = i32("ring6", 6)
        m.don_charged, m.n_don_charged = i32("don_charged", 2)
        m.don_neutral, m.n_don_neutral = i32("don_neutral", 2)
        m.acc_charged, m.n_acc_charged = i32("acc_charged")
        m.acc_neutral, m.n_acc_neutral = i32("acc_neutral")
        m.tstrain_coef = f64("tstrain_coef")
        m.n_tstrain = len(self.tstrain_coef)
        m.tstrain_match_off, _ = i32("tstrain_off")
        m.tstrain_ijkl, _ = i32("tstrain_ijkl", 4)
        m.n_rotatable = len(self.rotatable_bonds)
        m._keep = keep
        return m


# --------------------------------------------------------------------------
# JSON reading (data.cpp:34-276)

def _ids(node, width):
    a = np.asarray(node if len(node) else np.zeros((0, width)), dtype=np.int64)
    return a.reshape(-1, width).astype(np.int32)


def _series(rows, n_rows, fill, dtype):
    out = np.full((n_rows, ffi.DIH_TERMS), fill, dtype=dtype)
    for r, vals in enumerate(rows):
        for d, v in enumerate(vals):
            out[r, d] = v
    return out


def _dihedrals(ff, prefix):
    # fill(node, prefix, fourier_dihedral) (data.cpp:83-128)
    ids = _ids(ff[prefix + "_atomidx"], 4)
    n = len(ids)
    per = _series(ff[prefix + "_periodicity"], n, 0, np.int32)
    k = _series(ff[prefix + "_k"], n, 0.0, np.float64)
    ph = _deg2rad(_series(ff[prefix + "_phase"], n, 0.0, np.float64))
    return ids, per, ph, k


def _fill_ff(ff, ligand, mol_kwargs):
    # fill(node, ligand, force_field_params) (data.cpp:246-276)
    kw = mol_kwargs
    kw["bond_ij"] = _ids(ff["FF_Bonds_atomidx"], 2)
    kw["bond_k"] = as_f64(ff["FF_Bonds_k"])
    kw["bond_r0"] = as_f64(ff["FF_Bonds_length"])
    kw["angle_ijk"] = _ids(ff["FF_Angles_atomidx"], 3)
    kw["angle_k"] = as_f64(ff["FF_Angles_k"])
    kw["angle_t0"] = _deg2rad(as_f64(ff["FF_Angles_angle"]))
    imp = _dihedrals(ff, "FF_ImproperTorsions")
    pro = _dihedrals(ff, "FF_ProperTorsions")
    kw["n_impropers"] = len(imp[0])
    kw["dih_ijkl"] = np.concatenate([imp[0], pro[0]])
    kw["dih_n"] = np.concatenate([imp[1], pro[1]])
    kw["dih_phase"] = np.concatenate([imp[2], pro[2]])
    kw["dih_k"] = np.concatenate([imp[3], pro[3]])
    if ligand:
        t = np.asarray(ff["FF_vdW_paraidx"], dtype=np.int64) - ffi.VDW_ID_OFFSET
        for i, v in enumerate(t):
            # nolmaliz_vdw_types (data.cpp:234-244).  The reference lets id 99
            # through (out of its 99-entry table); it is rejected here.
            if v < 0 or v >= ffi.NUM_VDW_TYPES:
                raise ConfError(f"VDW type [{v}] of atom#{i} is illegal!")
        kw["vdw_type"] = t.astype(np.int32)
    else:
        kw["sigma"] = as_f64(ff["FF_vdW_sigma"])
        kw["epsilon"] = as_f64(ff["FF_vdW_epsilon"])
    kw["charge"] = as_f64(ff["partial_charges"])
    kw["pairs14"] = _ids(ff["FF_Nonbonded14_atomidx"], 2)
    kw["pairs"] = _ids(ff["FF_NonbondedAll_atomidx"], 2)
    kw["hydrophobic"] = as_i32(ff["hydrophobic_atomidx"])
    kw["cation"] = as_i32(ff["cation_atomidx"])
    kw["anion"] = as_i32(ff["anion_atomidx"])
    kw["ring5"] = _ids(ff["piring5_atomidx"], 5)
    kw["ring6"] = _ids(ff["piring6_atomidx"], 6)
    kw["don_charged"] = _ids(ff["hbonddon_charged_atomidx"], 2)
    kw["don_neutral"] = _ids(ff["hbonddon_neut_atomidx"], 2)
    kw["acc_charged"] = as_i32(ff["hbondacc_charged_atomidx"])
    kw["acc_neutral"] = as_i32(ff["hbondacc_neut_atomidx"])
    if ligand:
        params = ff.get("tstrain_params_pose_selection", [])
        if len(params):  # data.cpp:204-223
            kw["tstrain_coef"] = as_f64(params, 4)
            matches = ff["tstrain_atomidx_pose_selection"]
            off = [0]
            flat = []
            for m in matches:
                flat.extend(m)
                off.append(len(flat))
            kw["tstrain_off"] = as_i32(off)
            kw["tstrain_ijkl"] = _ids(flat, 4)
        kw["rotatable_bonds"] = _ids(ff.get("rotatable_bond_index", []), 2)
    kw["molecular_weight"] = float(ff.get("molecular_weight", -1.0))
    kw["atomic_numbers"] = as_i32(ff.get("atomic_numbers", []))
    kw["all_bonds"] = _ids(ff.get("bond_index", []), 2)


def _receptor_leaves(geo, xyz):
    # fill(node, rotable_geometry) + leaf precalculation (data.cpp:292-364)
    axis_atoms, frags, orders, keys = [], [], [], []
    for key, batch in geo["data"].items():
        r1, r2 = batch["rot_bond_1"], batch["rot_bond_2"]
        for idx in range(len(r1)):
            axis_atoms.append((int(r1[idx]), int(r2[idx])))
            frags.append([int(a) for a in batch["fragments"][idx]])
            keys.append(key)
        orders.extend(int(o) for o in batch["rot_bond_order"])
    origin, axis, off, atoms = [], [], [0], []
    for (a0, a1), fr in zip(axis_atoms, frags):
        o = xyz[a0]
        v = xyz[a1] - o
        v = v / math.sqrt(float(v @ v))
        origin.append(o)
        axis.append(v)
        atoms.extend(a for a in fr if a != a1)
        off.append(len(atoms))
    n = len(axis_atoms)
    return Leaves(origin=as_f64(origin).reshape(n, 3), axis=as_f64(axis).reshape(n, 3),
                  atom_off=as_i32(off), atoms=as_i32(atoms),
                  axis_atoms=as_i32(axis_atoms).reshape(n, 2), fragments=frags,
                  bond_order=as_i32(orders), keys=keys)


def _ligand_tree(geo):
    # fill(node, fragment_tree) + pre_order_traverse (data.cpp:396-482)
    n = int(geo["num_atoms"])
    split = [int(v) for v in geo["frag_split_index"]]
    nf = len(split) + 1
    bounds = [0] + split + [n]
    start = np.array(bounds[:-1], np.int32)
    end = np.array(bounds[1:], np.int32)
    perm = as_i32(geo["permute_index"])
    if len(perm) != n:
        raise ConfError("permute_index length mismatches num_atoms")
    rev = np.empty(n, np.int32)
    rev[perm] = np.arange(n, dtype=np.int32)
    parent = np.full(nf, -2, np.int32)
    children = {f: [] for f in range(nf)}
    roots = []
    for level in geo["frag_traverse_levels"]:
        for key, node in level.items():
            f = int(key)
            if len(node["parent"]) == 0:
                roots.append(f)
                parent[f] = -1
            for c in node["children"]:
                children[f].append(int(c))
                parent[int(c)] = f
    if (parent == -2).any():
        raise ConfError("fragment tree does not cover every fragment")
    bond = np.zeros((nf, 2), np.int32)
    edges = geo.get("edge_to_bond", {})
    for f in range(nf):
        p = int(parent[f])
        if p < 0:
            continue
        a, b = edges[f"{p}->{f}"]
        bond[f] = (rev[int(a)], rev[int(b)])
    order = []

    def visit(f):
        for c in children[f]:
            order.append(c)
            visit(c)

    for r in roots:
        visit(r)
    return FragTree(perm=perm, frag_start=start, frag_end=end, frag_parent=parent,
                    frag_bond=bond, tors_frag=as_i32(order))


def parse_receptor(obj):
    """torsional_receptor::parse (data.cpp:331-364); obj is a dict or a path."""
    if not isinstance(obj, dict):
        with open(obj) as fh:
            obj = json.load(fh)
    kw = {}
    _fill_ff(obj["ffdata"], False, kw)
    xyz = as_f64(obj["xyz"], 3)
    geo = obj["geometry"]
    if int(geo["num_atoms"]) != len(xyz):
        raise ConfError(f"Atom count mismatches! Wanted [{geo['num_atoms']}] vs. "
                        f"Parsed [{len(xyz)}]")
    kw["leaves"] = _receptor_leaves(geo, xyz)
    return Molecule(xyz=xyz, **kw)


def parse_ligand(obj, geometry=True):
    """free_ligand::parse (data.cpp:496-520); xyz keeps every pose [P,n,3]."""
    if not isinstance(obj, dict):
        with open(obj) as fh:
            obj = json.load(fh)
    kw = {}
    _fill_ff(obj["ffdata"], True, kw)
    poses = np.asarray(obj["xyz"], dtype=np.float64)
    if poses.ndim == 2:
        poses = poses[None]
    kw["smiles"] = obj.get("mapped_smiles", "")
    if geometry:
        kw["tree"] = _ligand_tree(obj["geometry"])
    return Molecule(xyz=poses.reshape(len(poses), -1, 3), **kw)


# --------------------------------------------------------------------------
# JSON writing (same keys as the reader; pxdock/pipeline/prepare_ligand.py:31-42)

def _rows(a):
    return np.asarray(a).tolist()


def _dih_json(m, lo, hi, prefix, out):
    ids = m.dih_ijkl[lo:hi]
    per, ph, k = [], [], []
    for r in range(lo, hi):
        nterm = 0
        while nterm < ffi.DIH_TERMS and m.dih_n[r, nterm] != 0:
            nterm += 1
        per.append([int(v) for v in m.dih_n[r, :nterm]])
        ph.append([float(v) for v in (m.dih_phase[r, :nterm] / math.pi * 180.0)])
        k.append([float(v) for v in m.dih_k[r, :nterm]])
    out[prefix + "_atomidx"] = _rows(ids)
    out[prefix + "_periodicity"] = per
    out[prefix + "_phase"] = ph
    out[prefix + "_k"] = k


def _ff_json(m):
    ff = {
        "FF_Bonds_atomidx": _rows(m.bond_ij), "FF_Bonds_k": _rows(m.bond_k),
        "FF_Bonds_length": _rows(m.bond_r0),
        "FF_Angles_atomidx": _rows(m.angle_ijk), "FF_Angles_k": _rows(m.angle_k),
        "FF_Angles_angle": _rows(np.asarray(m.angle_t0) / math.pi * 180.0),
    }
    _dih_json(m, 0, m.n_impropers, "FF_ImproperTorsions", ff)
    _dih_json(m, m.n_impropers, len(m.dih_ijkl), "FF_ProperTorsions", ff)
    ff.update({
        "FF_Nonbonded14_atomidx": _rows(m.pairs14), "FF_NonbondedAll_atomidx": _rows(m.pairs),
        "hydrophobic_atomidx": _rows(m.hydrophobic), "cation_atomidx": _rows(m.cation),
        "anion_atomidx": _rows(m.anion), "piring5_atomidx": _rows(m.ring5),
        "piring6_atomidx": _rows(m.ring6),
        "hbonddon_charged_atomidx": _rows(m.don_charged),
        "hbonddon_neut_atomidx": _rows(m.don_neutral),
        "hbondacc_charged_atomidx": _rows(m.acc_charged),
        "hbondacc_neut_atomidx": _rows(m.acc_neutral),
        "partial_charges": _rows(m.charge),
        "bond_index": _rows(m.all_bonds), "atomic_numbers": _rows(m.atomic_numbers),
        "rotatable_bond_index": _rows(m.rotatable_bonds),
    })
    if m.is_ligand:
        ff["FF_vdW_paraidx"] = _rows(np.asarray(m.vdw_type) + ffi.VDW_ID_OFFSET)
        ff["tstrain_params_pose_selection"] = _rows(m.tstrain_coef)
        ff["tstrain_atomidx_pose_selection"] = [
            _rows(m.tstrain_ijkl[m.tstrain_off[p]:m.tstrain_off[p + 1]])
            for p in range(len(m.tstrain_coef))]
        ff["molecular_weight"] = float(m.molecular_weight)
    else:
        ff["FF_vdW_sigma"] = _rows(m.sigma)
        ff["FF_vdW_epsilon"] = _rows(m.epsilon)
    return ff


def receptor_json(m):
    L = m.leaves
    data = {}
    for idx in range(L.n_leaves):
        key = L.keys[idx] if L.keys else str(idx)
        b = data.setdefault(key, {"fragments": [], "rot_bond_1": [], "rot_bond_2": [],
                                  "rot_bond_order": []})
        b["fragments"].append([int(a) for a in L.fragments[idx]])
        b["rot_bond_1"].append(int(L.axis_atoms[idx, 0]))
        b["rot_bond_2"].append(int(L.axis_atoms[idx, 1]))
        b["rot_bond_order"].append(int(L.bond_order[idx]) if L.bond_order is not None else idx)
    return {"xyz": _rows(m.xyz), "ffdata": _ff_json(m),
            "geometry": {"num_atoms": int(m.n_atoms), "num_rotatable_bonds": int(L.n_leaves),
                         "data": data}}


def ligand_json(m):
    t = m.tree
    n = int(np.asarray(m.xyz).shape[-2])
    levels = []
    children = {f: [] for f in range(t.n_frags)}
    for f in range(t.n_frags):
        if t.frag_parent[f] >= 0:
            children[int(t.frag_parent[f])].append(f)
    # children in the order the pre-order visits them
    order_pos = {int(f): i for i, f in enumerate(t.tors_frag)}
    for f in children:
        children[f].sort(key=lambda c: order_pos[c])
    current = [f for f in range(t.n_frags) if t.frag_parent[f] < 0]
    while current:
        lvl = {}
        nxt = []
        for f in current:
            p = int(t.frag_parent[f])
            lvl[str(f)] = {"children": children[f], "parent": [] if p < 0 else [p]}
            nxt.extend(children[f])
        levels.append(lvl)
        current = nxt
    edges = {}
    for f in range(t.n_frags):
        p = int(t.frag_parent[f])
        if p < 0:
            continue
        a, b = int(t.perm[t.frag_bond[f, 0]]), int(t.perm[t.frag_bond[f, 1]])
        edges[f"{p}->{f}"] = [a, b]
        edges[f"{f}->{p}"] = [b, a]
    geo = {"num_atoms": n, "permute_index": _rows(t.perm),
           "frag_split_index": _rows(t.frag_end[:-1]), "frag_traverse_levels": levels,
           "edge_to_bond": edges}
    return {"xyz": _rows(m.xyz), "ffdata": _ff_json(m), "mapped_smiles": m.smiles,
            "geometry": geo}
