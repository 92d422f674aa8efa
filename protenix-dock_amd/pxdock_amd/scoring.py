"""Scoring-function JSON (pose_selection) -> pxd_score_params.

Follows the reference loader fill(node, ScoringFunctionParams)
(engine/src/core/factory.cpp:144-211): coefficients default to 0, a term's
sub-object is read only when |coef| > 1e-5 (CHECK_PARAMETER_NON_ZERO,
engine/include/bytedock/lib/utility.h:26).  The default file is the
reference's pscore-v7_and_bscore-fake.json, shipped as package data.
"""
import json
import os

from . import ffi

DEFAULT_CONFIG = os.path.join(os.path.dirname(__file__), "data",
                              "pscore-v7_and_bscore-fake.json")

HBOND_KEYS = ("hldcc", "hldcn", "hldnc", "hldnn", "hpdcc", "hpdcn", "hpdnc", "hpdnn")
STACK_KEYS = ("l5r5", "l5r6", "l6r5", "l6r6")
PICAT_KEYS = ("lc5", "lc6", "pc5", "pc6")


def active(c):
    return c > 1e-5 or c < -1e-5


def params_from_dict(node, pad_vdw_types=False):
    """Fill a PxdScoreParams from one scoring tree (e.g. config['pose_selection']).

    The reference reads 99 vdW type scales whatever the list holds
    (scorer.cpp:210-217); a shorter list is an error here unless
    pad_vdw_types, which gives the missing types scale 1 (the affinity tree of
    the default configuration lists 38 types, all of scale 1, and the
    reference's read past them is undefined)."""
    p = ffi.PxdScoreParams()
    for t in range(ffi.NUM_VDW_TYPES):
        p.vdw_sigma_scale[t] = 1.0
        p.vdw_epsilon_scale[t] = 1.0
    g = lambda k: float(node.get(k, 0.0))
    p.coul_coef = g("interaction_coulomb_energy_coef")
    p.vdw_coef = g("interaction_vdw_energy_coef")
    p.ionic_coef = g("ionic_energy_coef")
    p.pication_coef = g("pi_cationic_energy_coef")
    p.f2f_coef = g("f2f_pi_stacking_energy_coef")
    p.e2f_coef = g("e2f_pi_stacking_energy_coef")
    p.hbond_coef = g("hbond_energy_coef")
    p.hydrophobic_coef = g("hydrophobic_energy_coef")
    p.tstrain_coef = g("torsion_strain_penalty_coef")
    if active(p.vdw_coef):
        v = node["interaction_vdw_energy"]
        p.vdw_scale = float(v["scale"])
        p.vdw_cutoff = float(v["cutoff"])
        types = v["types"]
        if len(types) < ffi.NUM_VDW_TYPES and not pad_vdw_types:
            # the reference reads 99 entries regardless (scorer.cpp:213-217)
            raise ValueError("interaction_vdw_energy.types needs 99 entries")
        for t in range(min(len(types), ffi.NUM_VDW_TYPES)):
            p.vdw_sigma_scale[t] = float(types[t]["sigma_scale"])
            p.vdw_epsilon_scale[t] = float(types[t]["epsilon_scale"])
    if active(p.ionic_coef):
        v = node["ionic_energy"]
        p.ionic_lc_coef = float(v["lc_coef"])
        p.ionic_pc_coef = float(v["pc_coef"])
        p.ionic_lc[0], p.ionic_lc[1] = v["lc_inter"]["dist_min"], v["lc_inter"]["dist_max"]
        p.ionic_pc[0], p.ionic_pc[1] = v["pc_inter"]["dist_min"], v["pc_inter"]["dist_max"]
    if active(p.pication_coef):
        v = node["pi_cationic_energy"]
        for i, k in enumerate(PICAT_KEYS):
            p.pication_sub_coef[i] = float(v[k + "_coef"])
            p.pication_range[i][0] = float(v[k + "_inter"]["dist_min"])
            p.pication_range[i][1] = float(v[k + "_inter"]["dist_max"])
    for name, sub, inter in (("f2f", p.f2f_sub_coef, p.f2f_inter),
                             ("e2f", p.e2f_sub_coef, p.e2f_inter)):
        if active(getattr(p, name + "_coef")):
            v = node[name + "_pi_stacking_energy"]
            for i, k in enumerate(STACK_KEYS):
                sub[i] = float(v[k + "_coef"])
                it = v[k + "_inter"]
                for j, f in enumerate(("dist_min", "dist_max", "angle_min", "angle_max",
                                       "sigma_min", "sigma_max")):
                    inter[i][j] = float(it[f])
    if active(p.hbond_coef):
        v = node["hbond_energy"]
        p.hbond_scale = float(v["scale"])
        for i, k in enumerate(HBOND_KEYS):
            p.hbond_sub_coef[i] = float(v[k + "_coef"])
            it = v[k + "_inter"]
            for j, f in enumerate(("dist_min", "dist_max", "angle_min", "angle_max",
                                   "dist_center", "angle_center")):
                p.hbond_inter[i][j] = float(it[f])
    if active(p.hydrophobic_coef):
        v = node["hydrophobic_energy"]
        p.hydro_scale = float(v["scale"])
        p.hydro_dist_min = float(v["dist_min"])
        p.hydro_dist_max = float(v["dist_max"])
        p.hydro_scale_factor = float(v["scale_factor"])
    return p


def bscore_params_from_dict(node):
    """pxd_bscore_params from an affinity_ranking tree, with the weights the
    reference's root_scorer ends up using (scorer.cpp:53-115): the
    rotatable-energy coefficient is stored under the torsion-strain key, so
    when both are active it replaces the torsion-strain weight, and the
    rotatable-energy leaf (named 'rotatable_energy') has no weight at all;
    molecular_weight_reward has no scorer.  The GBSA delta is weighted by
    gbsa_energy_delta_coef."""
    b = ffi.PxdBscoreParams()
    b.leaves = params_from_dict(node, pad_vdw_types=True)
    rot = float(node.get("rotatable_energy_coef", 0.0))
    if active(b.leaves.tstrain_coef) and active(rot):
        b.leaves.tstrain_coef = rot
    b.gbsa_coef = float(node.get("gbsa_energy_delta_coef", 0.0))
    return b


def load_affinity_ranking(path=None):
    """scoring_function_factory (factory.cpp:213-238), affinity_ranking tree."""
    with open(path or DEFAULT_CONFIG) as fh:
        cfg = json.load(fh)
    return bscore_params_from_dict(cfg["affinity_ranking"])


def load_pose_selection(path=None):
    """scoring_function_factory (factory.cpp:213-238), pose_selection tree only."""
    with open(path or DEFAULT_CONFIG) as fh:
        cfg = json.load(fh)
    return params_from_dict(cfg["pose_selection"])


def params_to_vector(p):
    """Flat view of the trainable pscore parameters: the 9 term weights then
    (sigma_scale, epsilon_scale) for each of the 99 types."""
    import numpy as np
    w = [p.coul_coef, p.vdw_coef, p.ionic_coef, p.pication_coef, p.f2f_coef, p.e2f_coef,
         p.hbond_coef, p.hydrophobic_coef, p.tstrain_coef]
    return np.array(w + list(p.vdw_sigma_scale) + list(p.vdw_epsilon_scale))
