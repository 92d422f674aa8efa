"""pxd_lbfgs_batch_noalloc (the L-BFGS entry point that allocates nothing)
against pxd_lbfgs_batch through the C ABI: the same poses give bit-identical
poses, energies, evaluation counts and convergence flags, on two calls in a
row (the second smaller, so stale scratch contents from the first would
show), with a short history (every tuned work slot fits the scratch) and
the reference's 100 (fewer slots than tuned)."""
import numpy as np
import pytest

from pxdock_amd import ffi, lib, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sysm():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    ligs = [synth.make_ligand(900 + i, n_heavy=12 + 7 * i, n_torsions=2 + i, n_rings=i % 3)
            for i in range(8)]
    ligs.append(synth.make_ligand(999, n_heavy=75, n_torsions=10, n_rings=2))  # widest class
    box = lib.decide_grid_dims(rec.pocket_center, [16.0] * 3, 0.5)
    R = lib.Receptor(rec)
    G = lib.Grid(R, box, sorted(set(int(t) for m in ligs for t in m.vdw_type)))
    return dict(rec=rec, ligs=ligs, R=R, G=G, B=lib.LigandBatch(ligs))


def _poses(s, n, seed):
    rng = np.random.default_rng(seed)
    pl = rng.integers(0, len(s["ligs"]), n).astype(np.int32)
    xyz = []
    for l in pl:
        lig = s["ligs"][l]
        c = s["rec"].pocket_center + rng.normal(0.0, 2.0, 3)
        R = synth.quat_to_rot(synth.random_quaternion(rng))
        t = rng.uniform(-np.pi, np.pi, lig.tree.n_torsions)
        xyz.append(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t))
    tors = rng.uniform(-1.0, 1.0, (n, s["rec"].leaves.n_leaves))
    return pl, xyz, tors


@pytest.mark.parametrize("history,iters", [(8, 60), (100, 40)])
def test_gpu_lbfgs_noalloc_equals_lbfgs_batch(sysm, history, iters):
    import torch
    prm = ffi.lbfgs_params(iters, history=history)
    nl = sysm["rec"].leaves.n_leaves
    for call, n in enumerate((600, 150)):
        print(f"\n[noalloc] history={history} call={call} poses={n}", flush=True)
        pl, xyz, tors = _poses(sysm, n, 100 * history + call)
        Pa = lib.PoseSet(sysm["B"], pl, xyz, nl, tors=tors)
        Pb = lib.PoseSet(sysm["B"], pl, xyz, nl, tors=tors)
        ea, nea, cva = lib.lbfgs(sysm["R"], sysm["G"], Pa, prm)
        eb, neb, cvb = lib.lbfgs_noalloc(sysm["R"], sysm["G"], Pb, prm)
        torch.cuda.synchronize()
        assert torch.equal(ea, eb)
        assert torch.equal(nea, neb)
        assert torch.equal(cva, cvb)
        assert torch.equal(Pa.xyz, Pb.xyz)
        assert torch.equal(Pa.tors, Pb.tors)
        assert bool(torch.isfinite(eb).all()) and int(neb.min()) >= 1


def test_gpu_lbfgs_noalloc_checks_arguments(sysm):
    prm = ffi.lbfgs_params(10, history=129)  # above the 128-entry history cap
    pl, xyz, tors = _poses(sysm, 4, 7)
    P = lib.PoseSet(sysm["B"], pl, xyz, sysm["rec"].leaves.n_leaves, tors=tors)
    with pytest.raises(lib.PxdError) as e:
        lib.lbfgs_noalloc(sysm["R"], sysm["G"], P, prm)
    assert e.value.status == 5  # PXD_ERR_RANGE

