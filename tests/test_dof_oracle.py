"""The rigid-body + torsion variables of the oracle (translation, rotation
vector, ligand torsions, receptor torsions; or_dof_energy / or_dof_grad),
pinned by FP64 central finite differences of the energy through
free_ligand::apply_parameters (data.cpp:592-641) on the brute-force path
(the grid path's stored gradients are not the derivative of its interpolated
energy, so finite differences there only agree to the grid's own error), and
by the identities at zero: no motion, and the translation gradient equal to
the summed Cartesian gradient.  The L-BFGS over these variables keeps every
bond length (only rigid motions and torsions), unlike the Cartesian one.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import ffi, synth


@pytest.fixture(scope="module")
def sysd():
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    lig = synth.make_ligand(61, n_heavy=18, n_torsions=5, n_rings=1)
    x0 = np.asarray(synth.ligand_apply(lig.tree, lig.xyz[0], rec.pocket_center + 1.0, np.eye(3),
                                       np.zeros(lig.tree.n_torsions)))
    return dict(rec=rec, lig=lig, s=O.System(rec, lig, None), x0=x0)


def _nvar(d):
    return 6 + d["lig"].tree.n_torsions + d["rec"].leaves.n_leaves


def test_zero_variables_are_the_input_pose(sysd):
    x = np.zeros(_nvar(sysd))
    e, g, pose = sysd["s"].dof_energy(sysd["x0"], x)
    np.testing.assert_allclose(pose, sysd["x0"], atol=1e-12)
    e0, gc, tg = sysd["s"].energy(sysd["x0"], np.zeros(sysd["rec"].leaves.n_leaves))
    assert e == pytest.approx(e0, rel=1e-13)
    np.testing.assert_allclose(g[:3], gc.sum(axis=0), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(g[6 + sysd["lig"].tree.n_torsions:], tg, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_central_differences(sysd, seed):
    rng = np.random.default_rng(seed)
    nt = sysd["lig"].tree.n_torsions
    x = np.concatenate([rng.normal(size=3) * 0.4, rng.normal(size=3) * 0.5,
                        rng.uniform(-math.pi, math.pi, nt),
                        rng.normal(size=sysd["rec"].leaves.n_leaves) * 0.5])
    e, g, _ = sysd["s"].dof_energy(sysd["x0"], x)
    fd = np.zeros_like(x)
    h = 1e-6
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (sysd["s"].dof_energy(sysd["x0"], xp)[0] -
                 sysd["s"].dof_energy(sysd["x0"], xm)[0]) / (2 * h)
    scale = max(np.linalg.norm(fd), 1.0)
    assert np.linalg.norm(g - fd) <= 1e-6 * scale, (np.abs(g - fd).max(), scale)


def test_rigid_torsion_lbfgs_keeps_bond_lengths(sysd):
    lig = sysd["lig"]
    ij = np.asarray(lig.bond_ij)
    b0 = np.linalg.norm(sysd["x0"][ij[:, 0]] - sysd["x0"][ij[:, 1]], axis=1)
    e0, _, _ = sysd["s"].energy(sysd["x0"], np.zeros(sysd["rec"].leaves.n_leaves))
    out = sysd["s"].lbfgs(sysd["x0"], None, ffi.lbfgs_params(60, dof_mode=1))
    b1 = np.linalg.norm(out["xyz"][ij[:, 0]] - out["xyz"][ij[:, 1]], axis=1)
    np.testing.assert_allclose(b1, b0, atol=1e-9)
    assert out["energy"] < e0
    # the reported energy is the energy of the returned pose
    e1, _, _ = sysd["s"].energy(out["xyz"], out["tors"])
    assert out["energy"] == pytest.approx(e1, rel=1e-12)
    # while the Cartesian optimiser moves bond lengths
    outr = sysd["s"].lbfgs(sysd["x0"], None, ffi.lbfgs_params(60))
    br = np.linalg.norm(outr["xyz"][ij[:, 0]] - outr["xyz"][ij[:, 1]], axis=1)
    assert np.abs(br - b0).max() > 1e-6
