"""ctypes binding of the FP64 oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as a product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source built -O3 -march=x86-64-v4 (AVX-512, FMA contraction): a
# CPU-baseline variant; the checker is LIB
NATIVE_LIB = os.path.join(HERE, "liboracle_native.so")
REF_LIB = os.path.join(HERE, "_ref", "libpxd_ref.so")

from pxdock_amd import ffi  # noqa: E402  (path set up by the caller)
from pxdock_amd.ffi import f64p, i32p  # noqa: E402

_lib = None
_ref = None
_libs = {}


def use(path):
    """Route every later call through the oracle built at path (LIB or
    NATIVE_LIB).  Only the CPU-baseline timing switches builds."""
    global _lib, LIB
    LIB = path
    _lib = _libs.get(path)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} is not built (make -C oracle)")
        L = C.CDLL(LIB)
        d, i, P = C.c_double, C.c_int32, C.c_void_p
        sig = {
            "or_matmul3": (None, [f64p, f64p, f64p]),
            "or_rotmat_quat": (None, [f64p, f64p]),
            "or_rotmat_axis": (None, [f64p, d, f64p]),
            "or_rotmat_axis_grad": (None, [f64p, d, f64p, f64p]),
            "or_angle_grad": (d, [f64p, f64p, f64p, f64p, f64p]),
            "or_dihedral": (d, [f64p, f64p, f64p, f64p]),
            "or_dihedral_grad": (d, [f64p] * 8),
            "or_rev_linear": (d, [d, d, d]),
            "or_rev_linear_c": (d, [d, d, d, d]),
            "or_coul_pair": (d, [d, d, f64p, d, d, f64p, f64p]),
            "or_vdw_pair": (d, [d, d, d, d, f64p, d, d, f64p, f64p]),
            "or_bond_energy": (d, [P, f64p, f64p]),
            "or_angle_energy": (d, [P, f64p, f64p]),
            "or_dihedral_energy": (d, [P, f64p, f64p]),
            "or_nonbonded_energy": (d, [P, f64p, f64p]),
            "or_intra_energy": (d, [P, f64p, f64p]),
            "or_grid_point": (None, [P, f64p, P, f64p, i32p, i, f64p]),
            "or_grid_build": (None, [P, f64p, P, P, i32p, i, f64p, i]),
            "or_grid_get": (i, [P, i, d, f64p, f64p, f64p]),
            "or_inter_energy": (d, [P, f64p, P, P, f64p, f64p, f64p]),
            "or_system_energy": (d, [P, f64p, f64p, f64p, f64p]),
            "or_lbfgs": (i, [P, P, f64p, f64p, f64p, i32p]),
            "or_score_many": (None, [P, i, i32p, i32p, f64p, f64p, i, f64p, f64p, f64p, i]),
            "or_lbfgs_many": (None, [P, P, i, i32p, i32p, f64p, f64p, i, f64p, i32p, i32p, i]),
            "or_receptor_apply": (None, [P, i, f64p, f64p, f64p]),
            "or_ligand_apply": (None, [P, i, f64p, f64p, f64p, f64p, f64p]),
            "or_pscore": (d, [P, P, f64p, P, f64p, f64p]),
            "or_pscore_vdw": (d, [P, P, f64p, P, f64p]),
            "or_philox4x32": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]),
            "or_pose_cluster": (i, [i, i, f64p, f64p, i32p, i, d, i32p]),
            "or_set_eval_noise": (None, [d, C.c_uint64]),
            "or_mc_walker": (i, [P, P, P, P, P, i, i, f64p, i, f64p, f64p, f64p, f64p, f64p,
                                 i32p, i32p]),
            "or_dof_grad": (None, [P, i, f64p, f64p, f64p, f64p]),
            "or_dof_energy": (d, [P, f64p, f64p, f64p, f64p]),
            "or_mc_mutate": (i, [P, P, i, i, C.POINTER(C.c_uint32), i, f64p]),
            "or_mc_draws": (None, [i, C.c_uint64, i, i, i, f64p]),
            "or_gbsa_atom_params": (None, [i, i32p, i, i32p, f64p, f64p]),
            "or_gbsa_energy": (d, [i, f64p, f64p, f64p, f64p]),
            "or_bscore": (d, [P, P, f64p, f64p, f64p, P, f64p, f64p, f64p, f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        _libs[LIB] = L
    return _lib


def ref():
    """The reference's own scorer/math/random built from its sources, or None."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        L = C.CDLL(REF_LIB)
        d, P = C.c_double, C.c_void_p
        L.ref_vdw_table.argtypes = [f64p, f64p]
        L.ref_coul_pair.restype = d
        L.ref_coul_pair.argtypes = [d, d, f64p, d, d, f64p, f64p]
        L.ref_vdw_pair.restype = d
        L.ref_vdw_pair.argtypes = [d, d, d, d, f64p, d, d, f64p, f64p]
        L.ref_angle_grad.restype = d
        L.ref_angle_grad.argtypes = [f64p] * 5
        L.ref_dihedral_grad.restype = d
        L.ref_dihedral_grad.argtypes = [f64p] * 8
        L.ref_rotmat_axis_grad.argtypes = [f64p, d, f64p, f64p]
        L.ref_rotmat_quat.argtypes = [f64p, f64p]
        L.ref_pscore.restype = d
        L.ref_pscore.argtypes = [P, P, f64p, P, f64p, f64p]
        L.ref_random_draws.argtypes = [C.c_uint, C.c_int, f64p, f64p]
        L.ref_score_params_compare.restype = C.c_int
        L.ref_score_params_compare.argtypes = [C.c_char_p, P, C.c_char_p, C.c_size_t]
        L.ref_pscore_json.restype = d
        L.ref_pscore_json.argtypes = [C.c_char_p, P, f64p, P, f64p]
        L.ref_bscore_json.restype = d
        L.ref_bscore_json.argtypes = [C.c_char_p, P, f64p, i32p, i32p, C.c_int, P, f64p, i32p,
                                      i32p, C.c_int, f64p]
        _ref = L
    return _ref


def dp(a):
    return a.ctypes.data_as(f64p)


def ip(a):
    return a.ctypes.data_as(i32p)


def byref(s):
    return C.cast(C.pointer(s), C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OrGrid(C.Structure):
    _fields_ = [("box", ffi.PxdBox), ("n_types", C.c_int32), ("types", i32p),
                ("maps", f64p), ("slope", C.c_double), ("n_excluded", C.c_int32),
                ("excluded", i32p)]


class OrSystem(C.Structure):
    _fields_ = [("rec", C.c_void_p), ("rec_xyz", f64p), ("leaves", C.c_void_p),
                ("lig", C.c_void_p), ("grid", C.c_void_p), ("tree", C.c_void_p)]


class Grid:
    """FP64 maps built by the oracle (or wrapped from elsewhere)."""

    def __init__(self, box, types, maps, slope, excluded):
        self.box = box
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.maps = f64(maps)
        self.excluded = np.ascontiguousarray(excluded, dtype=np.int32)
        self.c = OrGrid(box, len(self.types), ip(self.types), dp(self.maps), slope,
                        len(self.excluded), ip(self.excluded))


def grid_build(receptor, box, types, slope=1e6, nthreads=None):
    L = lib()
    rc = receptor.to_c()
    xyz = f64(receptor.xyz)
    excl = np.zeros(receptor.n_atoms, np.uint8)
    movable = receptor.leaves.movable() if receptor.leaves is not None else []
    excl[movable] = 1
    types = np.ascontiguousarray(types, dtype=np.int32)
    maps = np.zeros((len(types) + 1, box.n_points(), 4))
    L.or_grid_build(byref(rc), dp(xyz), excl.ctypes.data_as(C.c_void_p), byref(box), ip(types),
                    len(types), dp(maps), nthreads or os.cpu_count())
    return Grid(box, types, maps, slope, movable)


def grid_point(receptor, p, types):
    L = lib()
    rc = receptor.to_c()
    excl = np.zeros(receptor.n_atoms, np.uint8)
    if receptor.leaves is not None:
        excl[receptor.leaves.movable()] = 1
    types = np.ascontiguousarray(types, dtype=np.int32)
    out = np.zeros((len(types) + 1) * 4)
    L.or_grid_point(byref(rc), dp(f64(receptor.xyz)), excl.ctypes.data_as(C.c_void_p),
                    dp(f64(p)), ip(types), len(types), dp(out))
    return out.reshape(-1, 4)


class System:
    """or_system for one (receptor, ligand[, grid]) triple."""

    def __init__(self, receptor, ligand, grid=None):
        self.rc = receptor.to_c()
        self.lc = ligand.to_c()
        self.rxyz = f64(receptor.xyz)
        self.leaves = receptor.leaves.to_c() if receptor.leaves is not None else None
        self.grid = grid
        self.n_leaves = receptor.leaves.n_leaves if receptor.leaves is not None else 0
        self.n_lig = ligand.n_atoms
        self.tc = ligand.tree.to_c() if ligand.tree is not None else None
        self.n_tors = ligand.tree.n_torsions if ligand.tree is not None else 0
        self.c = OrSystem(C.addressof(self.rc), dp(self.rxyz),
                          C.addressof(self.leaves) if self.leaves is not None else None,
                          C.addressof(self.lc), C.addressof(grid.c) if grid else None,
                          C.addressof(self.tc) if self.tc is not None else None)

    def energy(self, lig_xyz, tors=None):
        tors = f64(tors if tors is not None else np.zeros(self.n_leaves))
        lx = f64(lig_xyz).reshape(-1, 3)
        g = np.zeros_like(lx)
        tg = np.zeros(max(self.n_leaves, 1))
        e = lib().or_system_energy(byref(self.c), dp(tors), dp(lx), dp(tg), dp(g))
        return e, g, tg[:self.n_leaves]

    def dof_energy(self, xyz0, x):
        """Energy, gradient and pose of the rigid-body + torsion variables x."""
        xx = f64(x)
        g = np.zeros_like(xx)
        pose = np.zeros((self.n_lig, 3))
        e = lib().or_dof_energy(byref(self.c), dp(f64(xyz0).reshape(-1, 3)), dp(xx), dp(g),
                                dp(pose))
        return e, g, pose

    def lbfgs(self, lig_xyz, tors, params):
        lx = f64(lig_xyz).reshape(-1, 3).copy()
        t = f64(tors if tors is not None else np.zeros(self.n_leaves)).copy()
        if len(t) == 0:
            t = np.zeros(1)
        e = np.zeros(1)
        ne = np.zeros(1, np.int32)
        conv = lib().or_lbfgs(byref(self.c), byref(params), dp(lx), dp(t), dp(e), ip(ne))
        return dict(xyz=lx, tors=t[:self.n_leaves], energy=float(e[0]), nevals=int(ne[0]),
                    converged=bool(conv))


def systems_array(systems):
    arr = (OrSystem * len(systems))(*[s.c for s in systems])
    return arr


def score_many(systems, pose_lig, pose_atom_off, xyz, tors, n_leaves, nthreads):
    arr = systems_array(systems)
    n = len(pose_lig)
    e = np.zeros(n)
    g = np.zeros_like(xyz)
    tg = np.zeros((n, max(n_leaves, 1)))
    t = f64(tors) if tors is not None else None
    lib().or_score_many(C.cast(arr, C.c_void_p), n, ip(pose_lig), ip(pose_atom_off), dp(xyz),
                        dp(t) if t is not None else None, n_leaves, dp(e), dp(g), dp(tg),
                        nthreads)
    return e, g, tg[:, :n_leaves]


def lbfgs_many(systems, params, pose_lig, pose_atom_off, xyz, tors, n_leaves, nthreads):
    arr = systems_array(systems)
    n = len(pose_lig)
    xyz = f64(xyz).copy()
    tors = f64(tors).copy()
    e = np.zeros(n)
    ne = np.zeros(n, np.int32)
    cv = np.zeros(n, np.int32)
    lib().or_lbfgs_many(C.cast(arr, C.c_void_p), byref(params), n, ip(pose_lig),
                        ip(pose_atom_off), dp(xyz), dp(tors), n_leaves, dp(e), ip(ne), ip(cv),
                        nthreads)
    return xyz, tors, e, ne, cv


def pscore(sp, receptor, rec_xyz, ligand, lig_xyz, use_ref=False):
    rc, lc = receptor.to_c(), ligand.to_c()
    terms = np.zeros(ffi.NUM_TERMS)
    L = ref() if use_ref else lib()
    fn = L.ref_pscore if use_ref else L.or_pscore
    s = fn(byref(sp), byref(rc), dp(f64(rec_xyz)), byref(lc), dp(f64(lig_xyz)), dp(terms))
    return s, terms


def gbsa_atom_params(mol):
    """(radius, screen) of every atom (fill_atom_properties, scorer.cpp:735-758)."""
    an = np.ascontiguousarray(mol.atomic_numbers, np.int32)
    bonds = np.ascontiguousarray(np.asarray(mol.all_bonds, np.int32).reshape(-1, 2))
    r, sc = np.zeros(len(an)), np.zeros(len(an))
    lib().or_gbsa_atom_params(len(an), ip(an), len(bonds), ip(bonds if len(bonds) else
                                                             np.zeros((1, 2), np.int32)),
                              dp(r), dp(sc))
    return r, sc


def gbsa_energy(xyz, charge, radius, screen):
    x = f64(xyz).reshape(-1, 3)
    return lib().or_gbsa_energy(len(x), dp(x), dp(f64(charge)), dp(f64(radius)), dp(f64(screen)))


def bscore(bp, receptor, rec_xyz, ligand, lig_xyz, rec_gb=None, lig_gb=None, use_ref=None):
    """Affinity score and its terms (leaves, E_pro, E_lig, E_com).  use_ref: the
    affinity tree's JSON text, to score with the reference's compiled code."""
    rc, lc = receptor.to_c(), ligand.to_c()
    rx, lx = f64(rec_xyz).reshape(-1, 3), f64(lig_xyz).reshape(-1, 3)
    if use_ref is not None:
        gb = np.zeros(3)
        rb = np.ascontiguousarray(np.asarray(receptor.all_bonds, np.int32).reshape(-1, 2))
        lb = np.ascontiguousarray(np.asarray(ligand.all_bonds, np.int32).reshape(-1, 2))
        s = ref().ref_bscore_json(use_ref.encode(), byref(rc), dp(rx),
                                  ip(np.ascontiguousarray(receptor.atomic_numbers, np.int32)),
                                  ip(rb if len(rb) else np.zeros((1, 2), np.int32)), len(rb),
                                  byref(lc), dp(lx),
                                  ip(np.ascontiguousarray(ligand.atomic_numbers, np.int32)),
                                  ip(lb if len(lb) else np.zeros((1, 2), np.int32)), len(lb),
                                  dp(gb))
        return s, gb
    rr, rs = rec_gb if rec_gb is not None else gbsa_atom_params(receptor)
    lr, ls = lig_gb if lig_gb is not None else gbsa_atom_params(ligand)
    terms = np.zeros(ffi.NUM_BSCORE_TERMS)
    s = lib().or_bscore(byref(bp), byref(rc), dp(rx), dp(f64(rr)), dp(f64(rs)), byref(lc),
                        dp(lx), dp(f64(lr)), dp(f64(ls)), dp(terms))
    return s, terms


def ligand_apply(tree, xyz, center, rot, tors):
    tc = tree.to_c()
    x = f64(xyz).reshape(-1, 3)
    out = np.zeros_like(x)
    tt = f64(tors if len(tors) else np.zeros(1))
    lib().or_ligand_apply(byref(tc), len(x), dp(x), dp(f64(center)), dp(f64(rot)), dp(tt),
                          dp(out))
    return out


def receptor_apply(leaves, xyz, tors):
    lc = leaves.to_c()
    x = f64(xyz).reshape(-1, 3)
    out = np.zeros_like(x)
    lib().or_receptor_apply(byref(lc), len(x), dp(x), dp(f64(tors)), dp(out))
    return out


def mc_walker(system, ligand, sp, prm, box, walker, ligand_index, xyz0, own_conformer=False,
              emin=None):
    """One Monte-Carlo walker (searcher.cpp:98-187) with the GPU's Philox draws.
    emin: a 1-element float64 array holding the ligand's lowest
    pre-optimisation energy (updated in place), for the prune."""
    tc = ligand.tree.to_c()
    n = ligand.n_atoms
    bx = np.zeros((n, 3))
    bt = np.zeros(max(system.n_leaves, 1))
    be, bp = np.zeros(1), np.zeros(1)
    ne = np.zeros(1, np.int32)
    npr = np.zeros(1, np.int32)
    steps = lib().or_mc_walker(byref(system.c), byref(tc), byref(sp), byref(prm), byref(box),
                               walker, ligand_index, dp(f64(xyz0).reshape(-1, 3)),
                               1 if own_conformer else 0, dp(emin) if emin is not None else None,
                               dp(bx), dp(bt), dp(be), dp(bp), ip(ne), ip(npr))
    return dict(xyz=bx, tors=bt[:system.n_leaves], energy=float(be[0]), pscore=float(bp[0]),
                nevals=int(ne[0]), steps=int(steps), pruned=int(npr[0]))


def mc_mutate(tree, prm, xyz, walker=1, ligand=0, ctr=0):
    """One walker mutation of xyz; returns (kind, new xyz, next counter)."""
    x = f64(xyz).reshape(-1, 3).copy()
    c = C.c_uint32(ctr)
    tc = tree.to_c()
    k = lib().or_mc_mutate(byref(tc), byref(prm), walker, ligand, C.byref(c), len(x), dp(x))
    return int(k), x, int(c.value)


def mc_draws(kind, count, seed=1, walker=1, ligand=0):
    width = {1: 3, 2: 4}.get(kind, 1)
    out = np.zeros(count * width)
    lib().or_mc_draws(kind, seed, walker, ligand, count, dp(out))
    return out.reshape(count, width) if width > 1 else out
