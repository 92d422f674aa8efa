"""The FP64 oracle against the reference's own gtest vectors.

Each test mirrors one gtest case (engine/utest/*.cpp) with its inputs,
expected values and EXPECT_NEAR tolerances transcribed into
tests/golden/gtest_vectors.json and the JSON fixtures beside it.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from conftest import golden_path
from pxdock_amd import data
from pxdock_amd.data import Molecule


def _mol(n, **kw):
    base = dict(xyz=np.zeros((n, 3)), charge=np.zeros(n), sigma=np.zeros(n),
                epsilon=np.zeros(n))
    base.update(kw)
    return Molecule(**base)


def _check(ref, got, tol):
    np.testing.assert_allclose(np.asarray(got), np.asarray(ref), rtol=0, atol=tol)


# ---- test_math.cpp ----

def test_algebra_multiply_matrix(golden):
    g = golden["math_matmul"]
    out = np.zeros(9)
    O.lib().or_matmul3(O.dp(O.f64(g["lhs"])), O.dp(O.f64(g["rhs"])), O.dp(out))
    _check(g["out"], out, g["tol"])


def test_geometry_dihedral_and_gradient(golden):
    g = golden["math_dihedral"]
    p = [O.f64(x) for x in g["pos"]]
    th = O.lib().or_dihedral(*[O.dp(x) for x in p])
    assert abs(th - g["theta"]) < g["theta_tol"]
    gr = [np.zeros(3) for _ in range(4)]
    th2 = O.lib().or_dihedral_grad(*[O.dp(x) for x in p], *[O.dp(x) for x in gr])
    assert abs(th2 - g["theta"]) < g["theta_tol"]
    for ref, got in zip(g["grad"], gr):
        _check(ref, got, g["grad_tol"])


def test_geometry_rotation_matrix_and_gradient(golden):
    g = golden["math_rotation"]
    ax = O.f64(g["axis_raw"])
    ax = ax / np.linalg.norm(ax)
    R = np.zeros(9)
    O.lib().or_rotmat_axis(O.dp(ax), math.radians(g["theta_deg"]), O.dp(R))
    _check(g["out"], R.reshape(3, 3) @ np.array(g["vin"], float), g["tol"])
    J = np.zeros(9)
    R2 = np.zeros(9)
    O.lib().or_rotmat_axis_grad(O.dp(O.f64(g["axis_unit"])), math.pi / 4, O.dp(R2), O.dp(J))
    _check(g["out"], R2.reshape(3, 3) @ np.array(g["vin"], float), g["tol"])
    _check(g["jacobian"], J.reshape(3, 3) @ np.array(g["vin"], float), g["tol"])


# ---- test_interaction.cpp ----

def test_intra_harmonic_bond(golden):
    g = golden["bond"]
    x = O.f64(g["xyz"])
    b = np.array(g["bonds"])
    m = _mol(len(x), bond_ij=b[:, :2].astype(np.int32), bond_k=b[:, 2], bond_r0=b[:, 3])
    mc = m.to_c()
    grad = np.zeros_like(x)
    e = O.lib().or_bond_energy(O.byref(mc), O.dp(x), O.dp(grad))
    assert abs(e - g["energy"]) < g["energy_tol"]
    _check(g["grad"], grad, g["grad_tol"])


def test_intra_harmonic_angle(golden):
    g = golden["angle"]
    x = O.f64(g["xyz"])
    a = np.array(g["angles_deg"])
    m = _mol(len(x), angle_ijk=a[:, :3].astype(np.int32), angle_k=a[:, 3],
             angle_t0=a[:, 4] / 180.0 * math.pi)
    mc = m.to_c()
    grad = np.zeros_like(x)
    e = O.lib().or_angle_energy(O.byref(mc), O.dp(x), O.dp(grad))
    assert abs(e - g["energy"]) < g["energy_tol"]
    _check(g["grad"], grad, g["grad_tol"])


def test_intra_fourier_dihedral(golden):
    g = golden["dihedral"]
    x = O.f64(g["xyz"])
    rows = g["impropers"] + g["propers"]
    m = _mol(len(x), n_impropers=len(g["impropers"]),
             dih_ijkl=np.array([r[0] for r in rows], np.int32),
             dih_n=np.array([r[1] for r in rows], np.int32),
             dih_phase=np.array([r[2] for r in rows], float) * math.pi,
             dih_k=np.array([r[3] for r in rows], float))
    mc = m.to_c()
    grad = np.zeros_like(x)
    e = O.lib().or_dihedral_energy(O.byref(mc), O.dp(x), O.dp(grad))
    assert abs(e - g["energy"]) < g["energy_tol"]
    _check(g["grad"], grad, g["grad_tol"])


def test_intra_self_nonbonded(golden):
    g = golden["nonbonded"]
    x = O.f64(g["xyz"])
    m = _mol(len(x), charge=O.f64(g["charges"]), sigma=O.f64(g["sigma"]),
             epsilon=O.f64(g["epsilon"]), pairs14=np.array(g["pairs14"], np.int32),
             pairs=np.array(g["pairs"], np.int32))
    mc = m.to_c()
    grad = np.zeros_like(x)
    e = O.lib().or_nonbonded_energy(O.byref(mc), O.dp(x), O.dp(grad))
    assert abs(e - g["energy"]) < g["energy_tol"]
    _check(g["grad"], grad, g["grad_tol"])


def test_inter_molecular_brute_force(golden):
    g = golden["inter_brute"]
    rx, lx = O.f64(g["receptor_xyz"]), O.f64(g["ligand_xyz"])
    rec = _mol(len(rx), xyz=rx, charge=O.f64(g["receptor_charges"]),
               sigma=O.f64(g["receptor_sigma"]), epsilon=O.f64(g["receptor_epsilon"]))
    lig = Molecule(xyz=lx[None], charge=O.f64(g["ligand_charges"]),
                   vdw_type=np.array(g["ligand_types"], np.int32))
    rc, lc = rec.to_c(), lig.to_c()
    rg, lg = np.zeros_like(rx), np.zeros_like(lx)
    e = O.lib().or_inter_energy(O.byref(rc), O.dp(rx), None, O.byref(lc), O.dp(lx), O.dp(rg),
                                O.dp(lg))
    assert abs(e - g["energy"]) < g["energy_tol"]
    _check(g["receptor_grad"], rg, g["grad_tol"])
    _check(g["ligand_grad"], lg, g["grad_tol"])


def _system_fixture(g):
    rec = data.parse_receptor(golden_path(g["receptor_json"]))
    lig = data.parse_ligand(golden_path(g["ligand_json"]))
    return rec, lig


def test_inter_receptor_xyz_to_torsion(golden):
    g = golden["system_torsions"]
    rec, lig = _system_fixture(g)
    assert rec.leaves.n_leaves == 3
    s = O.System(rec, lig)
    for case in g["cases"]:
        t = np.radians(case["torsions_deg"])
        e, _, tg = s.energy(lig.xyz[0], t)
        assert abs(e - case["energy"]) < case["energy_tol"]
        _check(case["tors_grad"], tg, case["tors_grad_tol"])


# ---- test_optimizer.cpp ----

@pytest.mark.parametrize("idx", [0, 1])
def test_optimizer_lbfgs(golden, idx):
    g = golden["lbfgs"]
    case = g["cases"][idx]
    rec, lig = _system_fixture(g)
    s = O.System(rec, lig)
    from pxdock_amd import ffi
    p = ffi.lbfgs_params(case["max_iter"], max_eval=case["max_eval"])
    out = s.lbfgs(lig.xyz[0], np.array(case["torsions"], float), p)
    assert abs(out["energy"] - case["energy"]) < case["energy_tol"]
    _check(case["out_torsions"], out["tors"], case["tors_tol"])
    _check(case["out_xyz"], out["xyz"], case["xyz_tol"])


# ---- test_conformer.cpp ----

def _eval_expr(s):
    return float(eval(s, {"pi": math.pi, "sqrt": math.sqrt}))


def test_conformer_apply_receptor_parameters(golden):
    g = golden["receptor_apply"]
    rec = data.parse_receptor(golden_path(g["receptor_json"]))
    t = np.array([_eval_expr(s) for s in g["torsions"]])
    out = O.receptor_apply(rec.leaves, rec.xyz, t)
    _check(g["out_xyz"], out, g["tol"])


def test_conformer_translate_rotate_mutate_ligand(golden):
    g = golden["ligand_apply"]
    lig = data.parse_ligand(golden_path(g["ligand_json"]))
    x = lig.xyz[0]
    c_old = x.mean(axis=0)
    zeros = np.zeros(lig.tree.n_torsions)
    out = O.ligand_apply(lig.tree, x, g["translate"]["center"], np.eye(3).ravel(), zeros)
    _check(x - c_old, out - np.array(g["translate"]["center"]), g["translate"]["tol"])
    rot = np.array(g["rotate"]["rot"], float)
    out = O.ligand_apply(lig.tree, x, [0, 0, 0], rot, zeros)
    exp = (x - c_old) * np.array([-1, -1, 1])
    _check(exp, out, g["rotate"]["tol"])
    mu = g["mutate"]
    rot = np.array([_eval_expr(s) for s in mu["rot_expr"]])
    tors = np.array([_eval_expr(s) for s in mu["torsions"]])
    out = O.ligand_apply(lig.tree, x, mu["center"], rot, tors)
    _check(mu["out_xyz"], out, mu["tol"])


# ---- test_scorer.cpp ----

def test_scorer_torsion_penalty(golden):
    g = golden["tstrain"]
    x = O.f64(g["xyz"])
    n = len(x)
    off = [0]
    flat = []
    for m in g["matched"]:
        flat.extend(m)
        off.append(len(flat))
    lig = Molecule(xyz=x[None], charge=np.zeros(n), vdw_type=np.zeros(n, np.int32),
                   tstrain_coef=np.array(g["coefs"], float), tstrain_off=np.array(off, np.int32),
                   tstrain_ijkl=np.array(flat, np.int32))
    rec = _mol(1)
    from pxdock_amd import ffi
    sp = ffi.PxdScoreParams()
    sp.tstrain_coef = 1.0
    s, terms = O.pscore(sp, rec, rec.xyz, lig, x)
    assert abs(terms[ffi.TERM_NAMES.index("tstrain")] - g["score"]) < g["tol"]
    assert abs(s - g["score"]) < g["tol"]
