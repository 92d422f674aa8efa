"""The oracle's Monte-Carlo draws and mutations pinned by their mathematics.

The walker stream cannot equal the reference's mt19937 + libstdc++ draws,
so these tests check each distribution the reference specifies
(random.cpp:22-49, searcher.cpp:189-223) with goodness-of-fit tests on
thousands of draws.  A mutation is measured from its result alone: the root
fragment gives the rigid motion (Kabsch fit about the input's geometric
centre, the point free_ligand::apply_parameters rotates about,
data.cpp:592-641), a moved fragment the torsion change.  A slip in a range
or an amplitude (for instance uniform_int(0, n_tors) instead of
(0, n_tors + 1), an amplitude of 1/R_g instead of 2/R_g, a torsion drawn
from [0, pi)) moves one of these statistics far outside its bound.

Fixed seeds make every statistic a fixed number; the bounds (p > 1e-3 for
the fits, several standard errors for moments) are far from the values a
correct sampler gives with these seeds.
"""
import math

import numpy as np
import pytest
from scipy import stats

import pyoracle as O
from pxdock_amd import ffi, synth

N = 6000


def test_uniform_draws():
    u = O.mc_draws(0, N, seed=3)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def test_point_in_unit_ball():
    v = O.mc_draws(1, N, seed=4)
    r = np.linalg.norm(v, axis=1)
    assert r.max() < 1.0
    # P(|v| <= r) = r^3 for a uniform ball (random_in_sphere, random.cpp:42-49)
    assert stats.kstest(r ** 3, "uniform").pvalue > 1e-3
    # isotropic: each coordinate has mean 0 and variance 1/5
    assert np.all(np.abs(v.mean(axis=0)) < 4 * math.sqrt(0.2 / N))
    assert np.all(np.abs(v.var(axis=0) - 0.2) < 0.02)


def test_unit_quaternion_is_uniform_on_rotations():
    q = O.mc_draws(2, N, seed=5)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    # uniform on S^3 (random_orientation, random.cpp:29-40): every component
    # squared has mean 1/4, and the rotation angle theta has the Haar density
    # (1 - cos theta) / pi, CDF (theta - sin theta) / pi
    assert np.all(np.abs((q ** 2).mean(axis=0) - 0.25) < 0.02)
    th = 2.0 * np.arccos(np.clip(np.abs(q[:, 0]), 0, 1))
    assert stats.kstest((th - np.sin(th)) / math.pi, "uniform").pvalue > 1e-3


def test_standard_normal():
    z = O.mc_draws(3, N, seed=6)
    assert stats.kstest(z, "norm").pvalue > 1e-3


@pytest.mark.parametrize("k", [0, 1, 5])
def test_uniform_integers_are_inclusive(k):
    z = O.mc_draws(4 + k, N, seed=7).astype(int)
    assert z.min() == 0 and z.max() == k
    if k:
        counts = np.bincount(z, minlength=k + 1)
        assert stats.chisquare(counts).pvalue > 1e-3


# ---------------------------------------------------------------- mutations

@pytest.fixture(scope="module")
def lig():
    m = synth.make_ligand(61, n_heavy=22, n_torsions=4, n_rings=1)
    assert m.tree.n_torsions == 4
    x = np.asarray(synth.ligand_apply(m.tree, m.xyz[0], np.array([1.0, -2.0, 3.0]), np.eye(3),
                                      np.zeros(m.tree.n_torsions)))
    return m, x


def _root_atoms(tree):
    f = [i for i in range(tree.n_frags) if tree.frag_parent[i] < 0][0]
    return [int(tree.perm[p]) for p in range(tree.frag_start[f], tree.frag_end[f])]


def _rigid(x0, x1, atoms):
    """Rotation R and translation delta with x1 = R (x0 - c) + c + delta on
    the given atoms, c the geometric centre of x0 (Kabsch)."""
    c = x0.mean(axis=0)
    a = x0[atoms] - c
    b = x1[atoms] - c
    ca, cb = a.mean(axis=0), b.mean(axis=0)
    H = (a - ca).T @ (b - cb)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    R = Vt.T @ D @ U.T
    delta = cb - R @ ca
    return R, delta


def _angle(R):
    return math.acos(max(-1.0, min(1.0, (np.trace(R) - 1.0) / 2.0)))


def _torsion_change(tree, x0, x1, t):
    """Signed rotation of torsion t's moved atoms about its bond axis."""
    f = int(tree.tors_frag[t])
    pa, mi = (int(tree.perm[s]) for s in tree.frag_bond[f])
    axis = x0[mi] - x0[pa]
    axis /= np.linalg.norm(axis)
    best, ang = 0.0, None
    for a in tree.moved_atoms(t):
        v0 = x0[a] - x0[mi]
        v0 -= axis * (v0 @ axis)
        v1 = x1[a] - x1[mi]
        v1 -= axis * (v1 @ axis)
        if np.linalg.norm(v0) > best:
            best = np.linalg.norm(v0)
            ang = math.atan2(np.cross(v0, v1) @ axis, v0 @ v1)
    return ang


def _mutations(m, x, prm, n):
    out = []
    for w in range(n):
        k, y, _ = O.mc_mutate(m.tree, prm, x, walker=w + 1, ligand=0, ctr=0)
        out.append((k, y))
    return out


def test_reference_mutation(lig):
    m, x = lig
    prm = ffi.mc_params(seed=12345)
    assert prm.mutation == 0 and prm.amplitude == 2.0
    nt = m.tree.n_torsions
    muts = _mutations(m, x, prm, N)
    kinds = np.array([k for k, _ in muts])
    # mutation kind uniform_int(0, n_tors + 1) inclusive (searcher.cpp:192)
    counts = np.bincount(kinds, minlength=nt + 2)
    assert len(counts) == nt + 2 and counts.min() > 0
    assert stats.chisquare(counts).pvalue > 1e-3
    root = _root_atoms(m.tree)
    c = x.mean(axis=0)
    gr = math.sqrt(((x - c) ** 2).sum(axis=1).mean())
    trans, rot, tors = [], [], {t: [] for t in range(nt)}
    for k, y in muts:
        R, d = _rigid(x, y, root)
        if k == 0:
            # pure translation: every atom moves by the same delta
            np.testing.assert_allclose(y - x, np.broadcast_to(d, x.shape), atol=1e-9)
            assert _angle(R) < 1e-6
            trans.append(np.linalg.norm(d))
        elif k == 1:
            # rotation about the geometric centre, no translation, rigid
            assert np.linalg.norm(d) < 1e-9
            np.testing.assert_allclose((y - c), (x - c) @ R.T, atol=1e-9)
            rot.append(_angle(R))
        else:
            # only torsion k - 2 moves, and only its subtree
            assert _angle(R) < 1e-9 and np.linalg.norm(d) < 1e-9
            moved = set(m.tree.moved_atoms(k - 2))
            still = [a for a in range(m.n_atoms) if a not in moved]
            np.testing.assert_allclose(y[still], x[still], atol=1e-9)
            tors[k - 2].append(_torsion_change(m.tree, x, y, k - 2))
    # translation: amplitude x a uniform point in the unit ball
    # (searcher.cpp:197-201): (|delta| / 2)^3 uniform, E|delta| = 1.5 A
    trans = np.array(trans)
    assert stats.kstest((trans / 2.0) ** 3, "uniform").pvalue > 1e-3
    assert abs(trans.mean() - 1.5) < 4 * 2.0 * math.sqrt(3 / 80) / math.sqrt(len(trans))
    # rotation: angle |2 u / R_g|, u uniform in the unit ball (searcher.cpp:204-210)
    rot = np.array(rot)
    assert rot.max() < 2.0 / gr + 1e-9
    assert stats.kstest((rot * gr / 2.0) ** 3, "uniform").pvalue > 1e-3
    # torsion: the change is uniform in [-pi, pi) (uniform_torsion, random.h)
    for t, a in tors.items():
        a = np.array(a)
        assert len(a) > 300
        assert stats.kstest((a + math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3


def test_gaussian_mutation(lig):
    m, x = lig
    prm = ffi.mc_params(seed=777)
    prm.mutation = 1
    prm.trans_sigma, prm.rot_sigma = 1.0, 0.3
    nt = m.tree.n_torsions
    muts = _mutations(m, x, prm, N)
    kinds = np.array([k for k, _ in muts])
    # exactly one torsion per step, uniform over the torsions
    assert kinds.min() >= 2 and kinds.max() <= nt + 1
    assert stats.chisquare(np.bincount(kinds - 2, minlength=nt)).pvalue > 1e-3
    root = _root_atoms(m.tree)
    deltas, angles, tors = [], [], []
    for k, y in muts:
        R, d = _rigid(x, y, root)
        deltas.append(d)
        angles.append(_angle(R))
        # the torsion's own change: undo the rigid motion first
        yr = (y - x.mean(axis=0) - d) @ R + x.mean(axis=0)
        tors.append(_torsion_change(m.tree, x, yr, k - 2))
    deltas = np.array(deltas)
    # translation N(0, sigma_t^2) per axis
    for a in range(3):
        assert stats.kstest(deltas[:, a] / prm.trans_sigma, "norm").pvalue > 1e-3
    # rotation vector N(0, sigma_r^2 I): its angle / sigma_r is chi with 3
    # degrees of freedom (Maxwell)
    assert stats.kstest(np.array(angles) / prm.rot_sigma, stats.chi(3).cdf).pvalue > 1e-3
    assert stats.kstest((np.array(tors) + math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3


def test_mutation_draws_advance_the_counter(lig):
    """Consecutive mutations of one walker use fresh draws: the counter moves
    on and the next mutation differs."""
    m, x = lig
    prm = ffi.mc_params(seed=5)
    k1, y1, c1 = O.mc_mutate(m.tree, prm, x, walker=3, ctr=0)
    k2, y2, c2 = O.mc_mutate(m.tree, prm, x, walker=3, ctr=c1)
    assert c1 > 0 and c2 > c1
    k3, y3, _ = O.mc_mutate(m.tree, prm, x, walker=3, ctr=0)
    assert k3 == k1 and np.array_equal(y3, y1)
    assert not np.array_equal(y1, y2)
