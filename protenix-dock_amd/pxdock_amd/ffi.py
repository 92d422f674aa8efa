"""ctypes mirror of the descriptor structs declared in include/pxdock.h.

The same structs feed the GPU library (libpxdock.so) and the FP64 oracle, so
both read byte-identical inputs.  Struct layouts must match pxdock.h field for
field.
"""
import ctypes as C

import numpy as np

NUM_VDW_TYPES = 99
VDW_ID_OFFSET = 1000
DIH_TERMS = 6
MAX_LIGAND_ATOMS = 128
MAX_FRAGMENTS = 32
MAX_LEAVES = 64
NUM_TERMS = 9
TERM_NAMES = ("coulomb", "vdw", "ionic", "pication", "f2f", "e2f", "hbond",
              "hydrophobic", "tstrain")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)


class PxdMol(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("vdw_type", i32p), ("sigma", f64p), ("epsilon", f64p), ("charge", f64p),
        ("n_bonds", C.c_int32), ("bond_ij", i32p), ("bond_k", f64p), ("bond_r0", f64p),
        ("n_angles", C.c_int32), ("angle_ijk", i32p), ("angle_k", f64p), ("angle_t0", f64p),
        ("n_dihedrals", C.c_int32), ("dih_ijkl", i32p), ("dih_n", i32p),
        ("dih_phase", f64p), ("dih_k", f64p),
        ("n_pairs14", C.c_int32), ("pairs14", i32p),
        ("n_pairs", C.c_int32), ("pairs", i32p),
        ("n_hydrophobic", C.c_int32), ("hydrophobic", i32p),
        ("n_cation", C.c_int32), ("cation", i32p),
        ("n_anion", C.c_int32), ("anion", i32p),
        ("n_ring5", C.c_int32), ("ring5", i32p),
        ("n_ring6", C.c_int32), ("ring6", i32p),
        ("n_don_charged", C.c_int32), ("don_charged", i32p),
        ("n_don_neutral", C.c_int32), ("don_neutral", i32p),
        ("n_acc_charged", C.c_int32), ("acc_charged", i32p),
        ("n_acc_neutral", C.c_int32), ("acc_neutral", i32p),
        ("n_tstrain", C.c_int32), ("tstrain_coef", f64p), ("tstrain_match_off", i32p),
        ("tstrain_ijkl", i32p),
        ("n_rotatable", C.c_int32),
    ]


class PxdFragTree(C.Structure):
    _fields_ = [
        ("n_frags", C.c_int32), ("perm", i32p), ("frag_start", i32p), ("frag_end", i32p),
        ("frag_parent", i32p), ("frag_bond", i32p), ("n_torsions", C.c_int32),
        ("tors_frag", i32p),
    ]


class PxdLeaves(C.Structure):
    _fields_ = [
        ("n_leaves", C.c_int32), ("origin", f64p), ("axis", f64p), ("atom_off", i32p),
        ("atoms", i32p),
    ]


class PxdBox(C.Structure):
    _fields_ = [
        ("init_xyz", C.c_double * 3), ("npoints", C.c_int32 * 3), ("spacing", C.c_double),
        ("oppo_xyz", C.c_double * 3), ("npoints_minus_one", C.c_int32 * 3),
        ("spacing_inv", C.c_double),
    ]

    @property
    def npts(self):
        return tuple(int(v) for v in self.npoints)

    def n_points(self):
        n = self.npts
        return n[0] * n[1] * n[2]


MAPS_VERSION = 1


class PxdMapsMeta(C.Structure):
    """pxd_maps_meta: metadata of a saved map directory."""
    _fields_ = [
        ("box", PxdBox), ("slope", C.c_double), ("n_atoms", C.c_int32),
        ("n_excluded", C.c_int32), ("receptor_key", C.c_uint64), ("n_types", C.c_int32),
        ("types", C.c_int32 * NUM_VDW_TYPES),
    ]

    @property
    def type_list(self):
        return [int(self.types[i]) for i in range(self.n_types)]


class PxdScoreParams(C.Structure):
    _fields_ = [
        ("coul_coef", C.c_double), ("vdw_coef", C.c_double), ("ionic_coef", C.c_double),
        ("pication_coef", C.c_double), ("f2f_coef", C.c_double), ("e2f_coef", C.c_double),
        ("hbond_coef", C.c_double), ("hydrophobic_coef", C.c_double),
        ("tstrain_coef", C.c_double),
        ("vdw_scale", C.c_double), ("vdw_cutoff", C.c_double),
        ("vdw_sigma_scale", C.c_double * NUM_VDW_TYPES),
        ("vdw_epsilon_scale", C.c_double * NUM_VDW_TYPES),
        ("ionic_lc_coef", C.c_double), ("ionic_pc_coef", C.c_double),
        ("ionic_lc", C.c_double * 2), ("ionic_pc", C.c_double * 2),
        ("pication_sub_coef", C.c_double * 4), ("pication_range", (C.c_double * 2) * 4),
        ("f2f_sub_coef", C.c_double * 4), ("f2f_inter", (C.c_double * 6) * 4),
        ("e2f_sub_coef", C.c_double * 4), ("e2f_inter", (C.c_double * 6) * 4),
        ("hbond_scale", C.c_double), ("hbond_sub_coef", C.c_double * 8),
        ("hbond_inter", (C.c_double * 6) * 8),
        ("hydro_scale", C.c_double), ("hydro_dist_min", C.c_double),
        ("hydro_dist_max", C.c_double), ("hydro_scale_factor", C.c_double),
    ]


NUM_BSCORE_TERMS = NUM_TERMS + 3  # leaves, then GB E_pro, E_lig, E_com


class PxdBscoreParams(C.Structure):
    """pxd_bscore_params: the affinity tree's leaves and the GBSA weight."""
    _fields_ = [("leaves", PxdScoreParams), ("gbsa_coef", C.c_double)]


class PxdLbfgsParams(C.Structure):
    _fields_ = [
        ("history", C.c_int32), ("max_iter", C.c_int32), ("max_eval", C.c_int32),
        ("lr", C.c_double), ("tol_grad", C.c_double), ("tol_change", C.c_double),
        ("c1", C.c_double), ("c2", C.c_double), ("ls_tol_change", C.c_double),
        ("gtd_threshold", C.c_double), ("max_ls", C.c_int32), ("dof_mode", C.c_int32),
    ]


def lbfgs_params(max_iter, history=100, max_eval=None, tol_grad=0.5, tol_change=1e-4,
                 dof_mode=0):
    """Reference defaults of lbfgs_step (optimizer.h:52-66)."""
    p = PxdLbfgsParams()
    p.history = history
    p.max_iter = max_iter
    p.max_eval = max_eval if max_eval else max_iter * 5 // 4
    p.lr = 1.0
    p.tol_grad = tol_grad
    p.tol_change = tol_change
    p.c1 = 1e-4
    p.c2 = 0.9
    p.ls_tol_change = 1e-9
    p.gtd_threshold = 1e5
    p.max_ls = 25
    p.dof_mode = dof_mode
    return p


class PxdMcParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("max_nsteps", C.c_int32), ("mutation", C.c_int32),
        ("amplitude", C.c_double), ("temperature", C.c_double), ("trans_sigma", C.c_double),
        ("rot_sigma", C.c_double), ("max_tries", C.c_int32), ("steps_per_launch", C.c_int32),
        ("relax", PxdLbfgsParams), ("full", PxdLbfgsParams),
        ("mmenergy_threshold", C.c_double),
    ]


def mc_params(seed=0, max_nsteps=200, relax_iters=40, full_iters=240, steps_per_launch=10):
    """monte_carlo_searcher defaults (searcher.h:33, bdock.cpp:60-78)."""
    p = PxdMcParams()
    p.seed = seed
    p.max_nsteps = max_nsteps
    p.mutation = 0
    p.amplitude = 2.0
    p.temperature = 1.2
    p.trans_sigma = 1.0
    p.rot_sigma = 0.3
    p.mmenergy_threshold = -1.0
    p.max_tries = 2000
    p.steps_per_launch = steps_per_launch
    p.relax = lbfgs_params(relax_iters)
    p.full = lbfgs_params(full_iters)
    return p


class PxdLbfgsStats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("evaluations", C.c_uint64),
                ("history_bytes", C.c_uint64), ("map_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class PxdTuning(C.Structure):
    _fields_ = [
        ("grid_types_per_pass", C.c_int32), ("map_layout", C.c_int32),
        ("lbfgs_slots_per_cu", C.c_int32), ("score_poses_per_block", C.c_int32),
        ("grid_ms", C.c_double * 3), ("score_ms", C.c_double * 3),
        ("lbfgs_ms", C.c_double * 4), ("score_ppb_ms", C.c_double * 3),
    ]

    def as_dict(self):
        return {"grid_types_per_pass": self.grid_types_per_pass, "map_layout": self.map_layout,
                "lbfgs_slots_per_cu": self.lbfgs_slots_per_cu,
                "score_poses_per_block": self.score_poses_per_block,
                "grid_ms": {k: v for k, v in zip((4, 8, 16), self.grid_ms)},
                "score_ms": {k: v for k, v in zip((0, 1, 2), self.score_ms)},
                "score_ppb_ms": {k: v for k, v in zip((1, 2, 4), self.score_ppb_ms)},
                "lbfgs_ms": {k: v for k, v in zip((2, 4, 8, 16), self.lbfgs_ms)}}


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (None for None/empty)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


def as_i32(a, width=None):
    a = np.ascontiguousarray(np.asarray(a if a is not None else [], dtype=np.int32))
    if width is not None:
        a = a.reshape(-1, width)
    return a


def as_f64(a, width=None):
    a = np.ascontiguousarray(np.asarray(a if a is not None else [], dtype=np.float64))
    if width is not None:
        a = a.reshape(-1, width)
    return a
