"""ctypes binding of libpxdock.so, the C ABI of include/pxdock.h.

This is the product path: every call goes to the HIP library.  There is no
CPU fallback: loading fails loudly when the library is missing.  Device
buffers are torch tensors on the current HIP device (torch is plumbing for
memory and streams only).
"""
import ctypes as C
import os

import numpy as np

from . import ffi
from .ffi import f32p, f64p, i32p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpxdock.so")

_lib = None


class PxdError(RuntimeError):
    """A pxd_status other than PXD_OK, with pxd_last_error() text."""

    def __init__(self, status, msg):
        super().__init__(f"pxd status {status}: {msg}")
        self.status = status


def load():
    """Load libpxdock.so (raise if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    # torch ships its own HIP runtime; loading it first makes libpxdock.so bind
    # to that same runtime (one HIP/HSA runtime per process).
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    P, i, d = C.c_void_p, C.c_int32, C.c_double
    PP = C.POINTER(C.c_void_p)
    sig = {
        "pxd_last_error": (C.c_char_p, []),
        "pxd_version": (C.c_char_p, []),
        "pxd_device_count": (i, [i32p]),
        "pxd_lbfgs_default_params": (None, [P, i]),
        "pxd_describe_boundary": (None, [f64p, f64p, P]),
        "pxd_decide_grid_dims": (None, [f64p, f64p, d, P]),
        "pxd_receptor_create": (i, [P, f64p, P, PP]),
        "pxd_receptor_destroy": (None, [P]),
        "pxd_grid_build": (i, [P, P, i32p, i, d, P, PP]),
        "pxd_grid_upload": (i, [P, P, i32p, i, d, f32p, PP]),
        "pxd_grid_download": (i, [P, f32p, C.c_size_t]),
        "pxd_grid_info": (i, [P, P, i32p]),
        "pxd_grid_set_slope": (i, [P, d]),
        "pxd_grid_set_layout": (i, [P, i, P]),
        "pxd_grid_destroy": (None, [P]),
        "pxd_ligand_batch_create": (i, [P, P, i, PP]),
        "pxd_ligand_batch_destroy": (None, [P]),
        "pxd_ligand_batch_natoms": (i, [P, i, i32p]),
        "pxd_score_grad_batch": (i, [P, P, P, i, P, P, P, P, P, P, P, P]),
        "pxd_lbfgs_batch": (i, [P, P, P, P, i, P, P, P, P, P, P, P, P]),
        "pxd_pscore_batch": (i, [P, P, P, i, P, P, P, P, P, P, P]),
        "pxd_lbfgs_plain": (i, [P, P, P, P, i, P, P, P, P, P, P, P, P]),
        "pxd_lbfgs_batch_noalloc": (i, [P, P, P, P, i, P, P, P, P, P, P, P, P]),
        "pxd_receptor_leaf_tables": (i, [P, i]),
        "pxd_receptor_info": (i, [P, i32p, i32p, f64p, f64p]),
        "pxd_mc_default_params": (None, [P]),
        "pxd_dof_grad_batch": (i, [P, i, P, P, P, P, P, P]),
        "pxd_pscore_grad_batch": (i, [P, P, P, i, P, P, P, P, P, P, P, P]),
        "pxd_get_tuning": (None, [P]),
        "pxd_set_tuning": (i, [P]),
        "pxd_lbfgs_last_stats": (i, [P]),
        "pxd_autotune": (i, [P, P, i32p, i, P, i, P, P, P, P, P, P]),
        "pxd_mc_search": (i, [P, P, P, P, P, P, i, P, P, P, P, P, P, i, d, P, P, P]),
        "pxd_mc_search_conformers": (i, [P, P, P, P, P, P, i, i32p, P, P, P, P, P, P, i, d, P,
                                         P, P]),
        "pxd_apply_poses": (i, [P, i, P, P, P, P, P, P, P, P, P]),
        "pxd_vdw_type_tag": (C.c_char_p, [i]),
        "pxd_gbsa_atom_params": (i, [i, i32p, i, i32p, f64p, f64p]),
        "pxd_bscore_batch": (i, [P, f64p, f64p, P, f64p, f64p, P, i, P, P, P, P, P, P, P]),
        "pxd_receptor_key": (i, [P, f64p, P, C.POINTER(C.c_uint64)]),
        "pxd_maps_write": (i, [C.c_char_p, P, i32p, f32p]),
        "pxd_maps_info": (i, [C.c_char_p, P, i32p, i]),
        "pxd_maps_read": (i, [C.c_char_p, P, i32p, i32p, i, f32p, C.c_size_t]),
        "pxd_grid_save": (i, [P, C.c_char_p]),
        "pxd_grid_load": (i, [P, C.c_char_p, i32p, i, d, PP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != 0:
        raise PxdError(status, load().pxd_last_error().decode())


def _ref(s):
    return C.cast(C.pointer(s), C.c_void_p)


def _dev(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def decide_grid_dims(center, size, spacing):
    """decide_grid_dims (engine/src/core/grid.cpp:65-86)."""
    b = ffi.PxdBox()
    load().pxd_decide_grid_dims(ffi.as_f64(center).ctypes.data_as(f64p),
                                ffi.as_f64(size).ctypes.data_as(f64p), spacing, _ref(b))
    return b


def describe_boundary(center, size):
    b = ffi.PxdBox()
    load().pxd_describe_boundary(ffi.as_f64(center).ctypes.data_as(f64p),
                                 ffi.as_f64(size).ctypes.data_as(f64p), _ref(b))
    return b


class Receptor:
    """Device-resident torsional_receptor (engine/include/bytedock/core/data.h:34-71)."""

    def __init__(self, mol):
        self.mol = mol
        self._c = mol.to_c()
        self._lv = mol.leaves.to_c() if mol.leaves is not None else None
        self._xyz = ffi.as_f64(mol.xyz)
        h = C.c_void_p()
        check(load().pxd_receptor_create(_ref(self._c), self._xyz.ctypes.data_as(f64p),
                                         _ref(self._lv) if self._lv is not None else None,
                                         C.byref(h)))
        self.h = h
        self.n_leaves = mol.leaves.n_leaves if mol.leaves is not None else 0

    def table_info(self):
        """(n_leaves, table nodes per leaf or 0, worst midpoint errors e, d)."""
        nl, nn = C.c_int32(), C.c_int32()
        ee, ed = C.c_double(), C.c_double()
        check(load().pxd_receptor_info(self.h, C.byref(nl), C.byref(nn), C.byref(ee),
                                       C.byref(ed)))
        return nl.value, nn.value, ee.value, ed.value

    def set_leaf_tables(self, enable):
        check(load().pxd_receptor_leaf_tables(self.h, 1 if enable else 0))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pxd_receptor_destroy(self.h)
            self.h = None


class Grid:
    """Device-resident receptor_cache maps (grid.h:60-133), FP32 value+gradient."""

    def __init__(self, receptor, box, types, slope=1e6, maps=None):
        self.receptor = receptor
        self.box = box
        self.types = np.ascontiguousarray(sorted(set(int(t) for t in types)), dtype=np.int32)
        h = C.c_void_p()
        if maps is None:
            check(load().pxd_grid_build(receptor.h, _ref(box), self.types.ctypes.data_as(i32p),
                                        len(self.types), slope, _stream(), C.byref(h)))
        else:
            m = np.ascontiguousarray(maps, dtype=np.float32)
            check(load().pxd_grid_upload(receptor.h, _ref(box), self.types.ctypes.data_as(i32p),
                                         len(self.types), slope, m.ctypes.data_as(f32p),
                                         C.byref(h)))
        self.h = h

    @classmethod
    def load(cls, receptor, path, types=None, slope=None):
        """receptor_cache::restore (grid.cpp:346-391): maps saved by save() for
        this receptor; types None loads every saved type."""
        g = cls.__new__(cls)
        g.receptor = receptor
        h = C.c_void_p()
        if types is None:
            tp, n = None, 0
        else:
            g_types = np.ascontiguousarray(sorted(set(int(t) for t in types)), dtype=np.int32)
            tp, n = g_types.ctypes.data_as(i32p), len(g_types)
        check(load().pxd_grid_load(receptor.h, os.fsencode(path), tp, n,
                                   -1.0 if slope is None else float(slope), C.byref(h)))
        g.h = h
        b, nt = ffi.PxdBox(), C.c_int32()
        check(load().pxd_grid_info(h, _ref(b), C.byref(nt)))
        g.box = b
        g.types = (g_types if types is not None
                   else np.ascontiguousarray(maps_info(path)[0].type_list, dtype=np.int32))
        return g

    def save(self, path):
        """receptor_cache::save (grid.cpp:309-344) in the map-directory format."""
        check(load().pxd_grid_save(self.h, os.fsencode(path)))

    def download(self):
        n = (len(self.types) + 1) * self.box.n_points() * 4
        out = np.empty(n, np.float32)
        check(load().pxd_grid_download(self.h, out.ctypes.data_as(f32p), n))
        return out.reshape(len(self.types) + 1, self.box.n_points(), 4)

    def set_layout(self, layout):
        """0: point-major maps; 1: each cell's 8 corners stored together."""
        check(load().pxd_grid_set_layout(self.h, int(layout), _stream()))

    def set_slope(self, slope):
        check(load().pxd_grid_set_slope(self.h, slope))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pxd_grid_destroy(self.h)
            self.h = None


def type_tags():
    """The 99 GAFF2 type tags in index order (kVdwTypeTags, constant.h:131-231)."""
    return [load().pxd_vdw_type_tag(t).decode() for t in range(ffi.NUM_VDW_TYPES)]


def receptor_key(mol):
    """pxd_receptor_key: fingerprint of the fixed receptor atoms the maps sum."""
    c = mol.to_c()
    lv = mol.leaves.to_c() if mol.leaves is not None else None
    xyz = ffi.as_f64(mol.xyz)
    k = C.c_uint64()
    check(load().pxd_receptor_key(_ref(c), xyz.ctypes.data_as(f64p),
                                  _ref(lv) if lv is not None else None, C.byref(k)))
    return k.value


def maps_write(path, box, types, maps, n_atoms, excluded, key, slope=1e6):
    """Host-only write of a map directory; maps is [(len(types)+1), P, 4] FP32."""
    m = ffi.PxdMapsMeta()
    m.box = box
    m.slope = slope
    m.n_atoms = n_atoms
    ex = ffi.as_i32(excluded)
    m.n_excluded = len(ex)
    m.receptor_key = key
    m.n_types = len(types)
    for i, t in enumerate(types):
        m.types[i] = int(t)
    a = np.ascontiguousarray(maps, dtype=np.float32)
    assert a.size == (len(types) + 1) * box.n_points() * 4
    check(load().pxd_maps_write(os.fsencode(path), _ref(m), ex.ctypes.data_as(i32p),
                                a.ctypes.data_as(f32p)))


def maps_info(path):
    """(PxdMapsMeta, excluded atoms) of a saved map directory (host only)."""
    m = ffi.PxdMapsMeta()
    check(load().pxd_maps_info(os.fsencode(path), _ref(m), None, 0))
    ex = np.empty(max(m.n_excluded, 1), np.int32)
    check(load().pxd_maps_info(os.fsencode(path), _ref(m), ex.ctypes.data_as(i32p), len(ex)))
    return m, ex[:m.n_excluded].copy()


def maps_read(path, types=None, expect=None):
    """Host-only read: FP32 [(n+1), P, 4], map 0 Coulomb, then the requested
    types (None: every saved type).  expect = (n_atoms, excluded, key) checks
    that the maps belong to that receptor."""
    meta, _ = maps_info(path)
    sel = meta.type_list if types is None else [int(t) for t in types]
    P = meta.box.n_points()
    out = np.empty((len(sel) + 1, P, 4), np.float32)
    tp = ffi.as_i32(sel if sel else [0])  # never a null list: null asks for every type
    e, ex = None, None
    if expect is not None:
        e = ffi.PxdMapsMeta()
        e.n_atoms, ex, e.receptor_key = expect[0], ffi.as_i32(expect[1]), expect[2]
        e.n_excluded = len(ex)
    check(load().pxd_maps_read(os.fsencode(path), _ref(e) if e is not None else None,
                               ex.ctypes.data_as(i32p) if ex is not None else None,
                               tp.ctypes.data_as(i32p), len(sel),
                               out.ctypes.data_as(f32p), out.size))
    return out


class LigandBatch:
    """Packed uneven ligands (free_ligand, data.h:73-99) on the device."""

    def __init__(self, ligands):
        self.ligands = list(ligands)
        self._c = [m.to_c() for m in self.ligands]
        mols = (ffi.PxdMol * len(self._c))(*self._c)
        trees = None
        if all(m.tree is not None for m in self.ligands):
            self._t = [m.tree.to_c() for m in self.ligands]
            trees = (ffi.PxdFragTree * len(self._t))(*self._t)
        self._arrs = (mols, trees)
        h = C.c_void_p()
        check(load().pxd_ligand_batch_create(C.cast(mols, C.c_void_p),
                                             C.cast(trees, C.c_void_p) if trees is not None else None,
                                             len(self._c), C.byref(h)))
        self.h = h
        self.natoms = np.array([m.n_atoms for m in self.ligands], np.int32)
        self.types = sorted(set(int(t) for m in self.ligands for t in m.vdw_type))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pxd_ligand_batch_destroy(self.h)
            self.h = None


class PoseSet:
    """CSR pose layout: pose p is ligand pose_lig[p] with coordinates
    xyz[pose_atom_off[p] : pose_atom_off[p] + natoms] (FP32, on the device)."""

    def __init__(self, batch, pose_lig, xyz_list, n_leaves, tors=None, device="cuda"):
        import torch
        self.batch = batch
        self.pose_lig = np.ascontiguousarray(pose_lig, dtype=np.int32)
        n = batch.natoms[self.pose_lig]
        self.pose_atom_off = np.zeros(len(n), np.int32)
        if len(n) > 1:
            self.pose_atom_off[1:] = np.cumsum(n)[:-1]
        self.n_atoms_total = int(n.sum())
        flat = np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x in xyz_list]) \
            if len(xyz_list) else np.zeros((0, 3))
        assert len(flat) == self.n_atoms_total, "pose coordinates do not match the ligands"
        self.n_poses = len(self.pose_lig)
        self.n_leaves = n_leaves
        self.d_lig = torch.from_numpy(self.pose_lig).to(device)
        self.d_off = torch.from_numpy(self.pose_atom_off).to(device)
        self.xyz = torch.from_numpy(flat.astype(np.float32).ravel()).to(device)
        t = np.zeros((self.n_poses, max(n_leaves, 1)), np.float32) if tors is None else \
            np.asarray(tors, np.float32).reshape(self.n_poses, -1)
        self.tors = torch.from_numpy(np.ascontiguousarray(t[:, :max(n_leaves, 1)])).to(device)

    def xyz_of(self, p, arr=None):
        a = (self.xyz if arr is None else arr).detach().cpu().numpy().reshape(-1, 3)
        o = self.pose_atom_off[p]
        return a[o:o + self.batch.natoms[self.pose_lig[p]]]


def score_grad(receptor, grid, poses):
    """pxd_score_grad_batch: energies (FP64), ligand gradients, torsion gradients."""
    import torch
    dev = poses.xyz.device
    e = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    g = torch.empty_like(poses.xyz)
    tg = torch.empty((poses.n_poses, max(poses.n_leaves, 1)), dtype=torch.float32, device=dev)
    check(load().pxd_score_grad_batch(receptor.h, grid.h if grid is not None else None,
                                      poses.batch.h, poses.n_poses, _dev(poses.d_lig),
                                      _dev(poses.d_off), _dev(poses.xyz), _dev(poses.tors),
                                      _dev(e), _dev(g), _dev(tg), _stream()))
    return e, g, tg


def lbfgs(receptor, grid, poses, params):
    """pxd_lbfgs_batch: optimises poses.xyz / poses.tors in place."""
    import torch
    dev = poses.xyz.device
    e = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    ne = torch.empty(poses.n_poses, dtype=torch.int32, device=dev)
    cv = torch.empty(poses.n_poses, dtype=torch.int32, device=dev)
    check(load().pxd_lbfgs_batch(receptor.h, grid.h if grid is not None else None, poses.batch.h,
                                 _ref(params), poses.n_poses, _dev(poses.d_lig), _dev(poses.d_off),
                                 _dev(poses.xyz), _dev(poses.tors), _dev(e), _dev(ne), _dev(cv),
                                 _stream()))
    return e, ne, cv


def lbfgs_last_stats():
    """pxd_lbfgs_last_stats: work totals of the last lbfgs() call (dict)."""
    st = ffi.PxdLbfgsStats()
    check(load().pxd_lbfgs_last_stats(_ref(st)))
    return st.as_dict()


def lbfgs_noalloc(receptor, grid, poses, params, out=None):
    """pxd_lbfgs_batch_noalloc: as lbfgs(), with the library allocating
    nothing; out = (energy, nevals, converged) device tensors to fill (made
    here when None).  Returns without waiting for the stream."""
    import torch
    dev = poses.xyz.device
    if out is None:
        out = (torch.empty(poses.n_poses, dtype=torch.float64, device=dev),
               torch.empty(poses.n_poses, dtype=torch.int32, device=dev),
               torch.empty(poses.n_poses, dtype=torch.int32, device=dev))
    e, ne, cv = out
    check(load().pxd_lbfgs_batch_noalloc(receptor.h, grid.h if grid is not None else None,
                                         poses.batch.h, _ref(params), poses.n_poses,
                                         _dev(poses.d_lig), _dev(poses.d_off), _dev(poses.xyz),
                                         _dev(poses.tors), _dev(e), _dev(ne), _dev(cv),
                                         _stream()))
    return e, ne, cv


def lbfgs_plain(receptor, grid, poses, params):
    """pxd_lbfgs_plain: host-driven L-BFGS, one pose per score-kernel launch."""
    import torch
    dev = poses.xyz.device
    e = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    ne = torch.empty(poses.n_poses, dtype=torch.int32, device=dev)
    cv = torch.empty(poses.n_poses, dtype=torch.int32, device=dev)
    check(load().pxd_lbfgs_plain(receptor.h, grid.h if grid is not None else None, poses.batch.h,
                                 _ref(params), poses.n_poses, _dev(poses.d_lig), _dev(poses.d_off),
                                 _dev(poses.xyz), _dev(poses.tors), _dev(e), _dev(ne), _dev(cv),
                                 _stream()))
    return e, ne, cv


def pscore(receptor, poses, params, with_terms=True):
    """pxd_pscore_batch: weighted pose score and the unweighted leaf terms."""
    import torch
    dev = poses.xyz.device
    s = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    t = torch.empty((poses.n_poses, ffi.NUM_TERMS), dtype=torch.float64, device=dev) \
        if with_terms else None
    check(load().pxd_pscore_batch(receptor.h, poses.batch.h, _ref(params), poses.n_poses,
                                  _dev(poses.d_lig), _dev(poses.d_off), _dev(poses.xyz),
                                  _dev(poses.tors), _dev(s), _dev(t), _stream()))
    return s, t


def gbsa_atom_params(mol):
    """pxd_gbsa_atom_params (host): Born radius and screening factor per atom."""
    an = np.ascontiguousarray(mol.atomic_numbers, dtype=np.int32)
    if len(an) != mol.n_atoms:
        raise ValueError("GBSA needs the atomic number of every atom")
    bonds = np.ascontiguousarray(np.asarray(mol.all_bonds, np.int32).reshape(-1, 2))
    r, sc = np.zeros(len(an)), np.zeros(len(an))
    check(load().pxd_gbsa_atom_params(len(an), an.ctypes.data_as(i32p), len(bonds),
                                      bonds.ctypes.data_as(i32p) if len(bonds) else None,
                                      r.ctypes.data_as(f64p), sc.ctypes.data_as(f64p)))
    return r, sc


def bscore(receptor, poses, params, with_terms=True, rec_gb=None, lig_gb=None):
    """pxd_bscore_batch: affinity score per pose and its terms [n, NUM_BSCORE_TERMS]
    (leaves, E_pro, E_lig, E_com).  rec_gb / lig_gb: (radius, screen) of the
    receptor and of the batch's atoms in batch order (default: from the
    molecules' atomic numbers and bonds)."""
    import torch
    if rec_gb is None:
        rec_gb = gbsa_atom_params(receptor.mol)
    if lig_gb is None:
        parts = [gbsa_atom_params(m) for m in poses.batch.ligands]
        lig_gb = (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))
    rr, rs = (np.ascontiguousarray(a, np.float64) for a in rec_gb)
    lr, ls = (np.ascontiguousarray(a, np.float64) for a in lig_gb)
    dev = poses.xyz.device
    s = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    t = torch.empty((poses.n_poses, ffi.NUM_BSCORE_TERMS), dtype=torch.float64, device=dev) \
        if with_terms else None
    check(load().pxd_bscore_batch(receptor.h, rr.ctypes.data_as(f64p), rs.ctypes.data_as(f64p),
                                  poses.batch.h, lr.ctypes.data_as(f64p),
                                  ls.ctypes.data_as(f64p), _ref(params), poses.n_poses,
                                  _dev(poses.d_lig), _dev(poses.d_off), _dev(poses.xyz),
                                  _dev(poses.tors), _dev(s), _dev(t), _stream()))
    return s, t


def mc_default_params():
    """pxd_mc_default_params: the reference's search defaults."""
    p = ffi.PxdMcParams()
    load().pxd_mc_default_params(_ref(p))
    return p


class SearchResult:
    """Per-walker outputs of pxd_mc_search (device tensors) and the clustering."""

    def __init__(self, batch, walkers, best_xyz, best_tors, best_energy, best_pscore, nevals,
                 cluster_idx, cluster_count):
        self.batch = batch
        self.walkers = walkers
        self.best_xyz = best_xyz
        self.best_tors = best_tors
        self.best_energy = best_energy
        self.best_pscore = best_pscore
        self.nevals = nevals
        self.cluster_idx = cluster_idx
        self.cluster_count = cluster_count
        n = batch.natoms
        self.walker_atom_off = np.concatenate([[0], np.cumsum(np.repeat(n, walkers))[:-1]]) \
            .astype(np.int32)

    def walker_xyz(self, w, arr=None):
        a = (self.best_xyz if arr is None else arr).detach().cpu().numpy().reshape(-1, 3)
        o = self.walker_atom_off[w]
        return a[o:o + self.batch.natoms[w // self.walkers]]


def mc_search(receptor, grid, batch, lig_xyz0, score_params, params, box, walkers,
              num_modes=0, min_rmsd=1.0, conformers=None):
    """pxd_mc_search: `walkers` Monte-Carlo walkers per ligand of `batch`.
    lig_xyz0: list of [natoms, 3] input conformations, one per ligand; or,
    with conformers (pxd_mc_search_conformers), conformers[l] is ligand l's
    list of [natoms, 3] conformations (lig_xyz0 is then ignored)."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    if conformers is not None:
        counts = np.array([len(c) for c in conformers], np.int32)
        flat = np.concatenate([np.asarray(x, np.float64).reshape(-1, 3)
                               for c in conformers for x in c])
        assert len(flat) == int((counts * batch.natoms).sum()), "conformers do not match"
    else:
        counts = None
        flat = np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x in lig_xyz0])
        assert len(flat) == int(batch.natoms.sum()), "input conformations do not match the batch"
    x0 = torch.from_numpy(flat.astype(np.float32).ravel()).to(dev)
    W = len(batch.natoms) * walkers
    nl = receptor.n_leaves
    bx = torch.zeros(3 * int(batch.natoms.sum()) * walkers, dtype=torch.float32, device=dev)
    bt = torch.zeros((W, max(nl, 1)), dtype=torch.float32, device=dev)
    be = torch.zeros(W, dtype=torch.float64, device=dev)
    bp = torch.zeros(W, dtype=torch.float64, device=dev)
    ne = torch.zeros(W, dtype=torch.int32, device=dev)
    ci = torch.full((len(batch.natoms), max(num_modes, 1)), -1, dtype=torch.int32, device=dev)
    cc = torch.zeros(len(batch.natoms), dtype=torch.int32, device=dev)
    g = grid.h if grid is not None else None
    if counts is None:
        check(load().pxd_mc_search(receptor.h, g, batch.h, _ref(score_params), _ref(params),
                                   _ref(box), walkers, _dev(x0), _dev(bx), _dev(bt), _dev(be),
                                   _dev(bp), _dev(ne), num_modes, min_rmsd, _dev(ci), _dev(cc),
                                   _stream()))
    else:
        check(load().pxd_mc_search_conformers(
            receptor.h, g, batch.h, _ref(score_params), _ref(params), _ref(box), walkers,
            counts.ctypes.data_as(i32p), _dev(x0), _dev(bx), _dev(bt), _dev(be), _dev(bp),
            _dev(ne), num_modes, min_rmsd, _dev(ci), _dev(cc), _stream()))
    return SearchResult(batch, walkers, bx, bt, be, bp, ne, ci, cc)


def apply_poses(batch, pose_lig, xyz_in, center, rot, tors_list):
    """pxd_apply_poses: rebuild each pose from its input conformation, a new
    centre, a row-major orientation matrix and the ligand's torsion rotations."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    pose_lig = np.ascontiguousarray(pose_lig, dtype=np.int32)
    n = batch.natoms[pose_lig]
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int32)
    toff = np.concatenate([[0], np.cumsum([len(t) for t in tors_list])]).astype(np.int32)
    flat = np.concatenate([np.asarray(x, np.float64).reshape(-1, 3) for x in xyz_in])
    tflat = np.concatenate([np.asarray(t, np.float64).ravel() for t in tors_list] + [np.zeros(1)])
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt).ravel()).to(dev)
    d_lig, d_off, d_toff = T(pose_lig, np.int32), T(off, np.int32), T(toff, np.int32)
    d_x, d_c, d_r, d_t = (T(flat, np.float32), T(center, np.float32), T(rot, np.float32),
                          T(tflat, np.float32))
    out = torch.empty_like(d_x)
    check(load().pxd_apply_poses(batch.h, len(pose_lig), _dev(d_lig), _dev(d_off), _dev(d_x),
                                 _dev(d_c), _dev(d_r), _dev(d_toff), _dev(d_t), _dev(out),
                                 _stream()))
    return out.cpu().numpy().reshape(-1, 3), off


DOF_STRIDE = 6 + 32 - 1


def dof_grad(poses, lig_grad):
    """pxd_dof_grad_batch: [n_poses][DOF_STRIDE] force, torque about the pose
    centre, torsion torques, from the Cartesian gradient of score_grad."""
    import torch
    out = torch.empty((poses.n_poses, DOF_STRIDE), dtype=torch.float32, device=poses.xyz.device)
    check(load().pxd_dof_grad_batch(poses.batch.h, poses.n_poses, _dev(poses.d_lig),
                                    _dev(poses.d_off), _dev(poses.xyz), _dev(lig_grad), _dev(out),
                                    _stream()))
    return out


NUM_SCORE_PARAMS = ffi.NUM_TERMS + 2 * ffi.NUM_VDW_TYPES


def pscore_grad(receptor, poses, params, with_dxyz=True, with_dparams=True):
    """pxd_pscore_grad_batch: pscore per pose, d score / d ligand xyz (CSR like
    poses.xyz) and d score / d parameters [n_poses][NUM_SCORE_PARAMS]."""
    import torch
    dev = poses.xyz.device
    s = torch.empty(poses.n_poses, dtype=torch.float64, device=dev)
    dx = torch.empty_like(poses.xyz) if with_dxyz else None
    dp = torch.empty((poses.n_poses, NUM_SCORE_PARAMS), dtype=torch.float64, device=dev) \
        if with_dparams else None
    check(load().pxd_pscore_grad_batch(receptor.h, poses.batch.h, _ref(params), poses.n_poses,
                                       _dev(poses.d_lig), _dev(poses.d_off), _dev(poses.xyz),
                                       _dev(poses.tors), _dev(s), _dev(dx), _dev(dp), _stream()))
    return s, dx, dp


def get_tuning():
    t = ffi.PxdTuning()
    load().pxd_get_tuning(_ref(t))
    return t


def set_tuning(grid_types_per_pass=8, map_layout=0, lbfgs_slots_per_cu=8,
               score_poses_per_block=1):
    t = ffi.PxdTuning()
    t.grid_types_per_pass = grid_types_per_pass
    t.map_layout = map_layout
    t.lbfgs_slots_per_cu = lbfgs_slots_per_cu
    t.score_poses_per_block = score_poses_per_block
    check(load().pxd_set_tuning(_ref(t)))


def autotune(receptor, box, types, poses):
    """pxd_autotune on these poses; returns the chosen PxdTuning."""
    t = ffi.PxdTuning()
    ty = np.ascontiguousarray(types, dtype=np.int32)
    check(load().pxd_autotune(receptor.h, _ref(box), ty.ctypes.data_as(i32p), len(ty),
                              poses.batch.h, poses.n_poses, _dev(poses.d_lig), _dev(poses.d_off),
                              _dev(poses.xyz), _dev(poses.tors), _stream(), _ref(t)))
    return t
