"""Pins of oracle functions that no reference gtest vector covers, by what
the mathematics fixes:

  * Philox4x32-10 against the Random123 known-answer vectors (Salmon et al.,
    SC'11; kat_vectors of the Random123 distribution);
  * receptor_cache::get (grid.cpp:197-303): a map holding a linear field
    with its constant gradient is interpolated exactly; a node returns its
    stored value; outside the box the energy gains |outside| * slope, the
    gradient +-slope per dimension and no map gradient;
  * receptor_cache::populate (grid.cpp:164-194): a node of the Coulomb map is
    the brute-force inter energy of a unit probe charge (difference of two
    one-atom ligands that differ only in charge) and a node of a type map
    is the brute-force inter energy of an uncharged probe of that type, with
    the gradients of the same brute force (inter brute force is pinned by the
    reference's gtest vector in test_oracle_golden.py);
  * pose_cluster::preprocess (writer.cpp:124-177): with min_rmsd 0 it is a
    top-k by pscore, with min_rmsd beyond every distance it keeps the single
    best pose, and failed candidates (nevals 0) never appear.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import data, ffi, synth


def _philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    O.lib().or_philox4x32(c, k, o)
    return list(o)


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert _philox(ctr, key) == expect


def _box(init, npts, spacing):
    b = ffi.PxdBox()
    for d in range(3):
        b.init_xyz[d] = init[d]
        b.npoints[d] = npts[d]
        b.npoints_minus_one[d] = npts[d] - 1
        b.oppo_xyz[d] = init[d] + (npts[d] - 1) * spacing
    b.spacing = spacing
    b.spacing_inv = 1.0 / spacing
    return b


def _grid_get(g, t, q, x):
    e = np.zeros(1)
    gr = np.zeros(3)
    hit = O.lib().or_grid_get(O.byref(g.c), t, q, O.dp(O.f64(x)), O.dp(e), O.dp(gr))
    return hit, e[0], gr


@pytest.fixture(scope="module")
def linear_grid():
    box = _box([-3.0, 1.0, 2.5], [9, 7, 11], 0.5)
    nx, ny, nz = 9, 7, 11
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pts = np.stack([ii, jj, kk], -1).reshape(-1, 3) * 0.5 + np.array(box.init_xyz)
    a_c, b_c = np.array([0.7, -1.3, 0.4]), 2.0     # Coulomb map field
    a_t, b_t = np.array([-0.2, 0.9, 1.7]), -5.0    # type map field
    maps = np.zeros((2, len(pts), 4))
    maps[0, :, 0] = pts @ a_c + b_c
    maps[0, :, 1:] = a_c
    maps[1, :, 0] = pts @ a_t + b_t
    maps[1, :, 1:] = a_t
    g = O.Grid(box, [7], maps, 1e6, [])
    return dict(g=g, box=box, pts=pts, a_c=a_c, b_c=b_c, a_t=a_t, b_t=b_t, maps=maps)


def test_grid_interpolates_a_linear_field_exactly(linear_grid):
    L = linear_grid
    rng = np.random.default_rng(0)
    lo = np.array(L["box"].init_xyz)
    hi = np.array(L["box"].oppo_xyz)
    for _ in range(200):
        x = rng.uniform(lo, hi)
        q = rng.normal()
        hit, e, g = _grid_get(L["g"], 7, q, x)
        assert hit == 1
        assert e == pytest.approx(q * (x @ L["a_c"] + L["b_c"]) + x @ L["a_t"] + L["b_t"],
                                  rel=1e-12, abs=1e-12)
        np.testing.assert_allclose(g, q * L["a_c"] + L["a_t"], rtol=1e-12, atol=1e-12)


def test_grid_returns_node_values(linear_grid):
    L = linear_grid
    rng = np.random.default_rng(1)
    for idx in rng.integers(0, len(L["pts"]), 50):
        x = L["pts"][idx]
        hit, e, _ = _grid_get(L["g"], 7, 1.0, x)
        assert e == pytest.approx(L["maps"][0, idx, 0] + L["maps"][1, idx, 0], rel=1e-13, abs=1e-13)


def test_grid_wall_penalty(linear_grid):
    L = linear_grid
    hi = np.array(L["box"].oppo_xyz)
    lo = np.array(L["box"].init_xyz)
    inside = (lo + hi) / 2
    for d in range(3):
        for side, sign in ((hi, +1.0), (lo, -1.0)):
            delta = 0.37
            x = inside.copy()
            x[d] = side[d] + sign * delta
            xc = x.copy()
            xc[d] = side[d]  # clamped point: the map energy is read there
            hit, e, g = _grid_get(L["g"], 7, 0.5, x)
            assert hit == 0
            e_map = 0.5 * (xc @ L["a_c"] + L["b_c"]) + xc @ L["a_t"] + L["b_t"]
            assert e == pytest.approx(e_map + delta * 1e6, rel=1e-12)
            expect = np.zeros(3)
            expect[d] = sign * 1e6  # no map gradient on a miss
            np.testing.assert_allclose(g, expect, rtol=0, atol=1e-9)


def _one_atom_ligand(t, q):
    return data.Molecule(xyz=np.zeros((1, 1, 3)), charge=np.array([q]),
                         vdw_type=np.array([t], np.int32))


def test_grid_node_is_the_brute_force_probe_energy():
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    rc = rec.to_c()
    rxyz = O.f64(rec.xyz)
    excl = np.zeros(rec.n_atoms, np.uint8)  # brute-force inter sees every receptor atom
    types = np.array([3, 17, 40], np.int32)
    rng = np.random.default_rng(2)
    for _ in range(12):
        p = rec.pocket_center + rng.normal(size=3) * 3.0
        out = np.zeros((len(types) + 1) * 4)
        O.lib().or_grid_point(O.byref(rc), O.dp(rxyz), excl.ctypes.data_as(C.c_void_p), O.dp(p),
                              O.ip(types), len(types), O.dp(out))
        out = out.reshape(-1, 4)

        def inter(t, q):
            lig = _one_atom_ligand(int(t), q)
            lc = lig.to_c()
            rg = np.zeros_like(rxyz)
            lg = np.zeros(3)
            e = O.lib().or_inter_energy(O.byref(rc), O.dp(rxyz), None, O.byref(lc), O.dp(O.f64(p)),
                                        O.dp(rg), O.dp(lg))
            return e, lg

        for k, t in enumerate(types):
            e1, g1 = inter(t, 1.0)
            e0, g0 = inter(t, 0.0)
            assert out[k + 1, 0] == pytest.approx(e0, rel=1e-10, abs=1e-10)
            np.testing.assert_allclose(out[k + 1, 1:], g0, rtol=1e-9, atol=1e-9)
            assert out[0, 0] == pytest.approx(e1 - e0, rel=1e-9, abs=1e-9)
            np.testing.assert_allclose(out[0, 1:], g1 - g0, rtol=1e-8, atol=1e-8)


def _cluster(xyz, ps, ne, modes, rmsd):
    n_cand, n = xyz.shape[0], xyz.shape[1]
    out = np.zeros(modes, np.int32)
    cnt = O.lib().or_pose_cluster(n_cand, n, O.dp(np.ascontiguousarray(xyz, np.float64)),
                                  O.dp(O.f64(ps)), O.ip(np.ascontiguousarray(ne, np.int32)),
                                  modes, rmsd, O.ip(out))
    return list(out[:cnt])


def test_cluster_properties():
    rng = np.random.default_rng(5)
    n_cand, n = 40, 12
    xyz = rng.normal(size=(n_cand, n, 3)) * 4.0
    ps = rng.normal(size=n_cand)
    ne = np.ones(n_cand, np.int32)
    ne[[3, 17, 29]] = 0
    ok = np.flatnonzero(ne)
    # min_rmsd 0: the top-k by pscore, ascending
    top = _cluster(xyz, ps, ne, 9, 0.0)
    assert top == list(ok[np.argsort(ps[ok], kind="stable")][:9])
    # min_rmsd beyond every distance: the single best candidate
    assert _cluster(xyz, ps, ne, 9, 1e9) == [int(ok[np.argmin(ps[ok])])]
    # duplicates of one pose collapse onto its best-scored copy
    dup = np.repeat(xyz[:1], 5, axis=0)
    dps = np.array([3.0, 1.0, 2.0, 0.5, 4.0])
    assert _cluster(dup, dps, np.ones(5, np.int32), 9, 0.5) == [3]
    # failed candidates never appear
    for k in (1, 5, 9):
        assert not set(_cluster(xyz, ps, ne, k, 2.0)) & {3, 17, 29}
