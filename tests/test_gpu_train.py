"""Backward pass of the pose score (pxd_pscore_grad_batch) and its
torch.autograd.Function, against FP64 central finite differences of the
oracle's pscore (the reference has no pscore gradient to compare with).

Finite differences use h = 1e-6 on FP64, so a step crosses a ramp or clamp
breakpoint only rarely; the comparison is norm-wise per pose, 1e-3 relative
(FP32 device arithmetic against FP64 differences).
"""
import math

import numpy as np
import pytest

import pyoracle as O
from pxdock_amd import ffi, lib, scoring, synth, train

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def tw():
    import torch
    assert torch.cuda.is_available(), "a GPU is required"
    rec = synth.make_receptor(3, n_heavy=500, radius=16.0, pocket_offset=7.0, n_ring6=3,
                              n_ring5=1, n_leaves=3)
    ligs = [synth.make_ligand(70 + i, n_heavy=h, n_torsions=t, n_rings=r)
            for i, (h, t, r) in enumerate([(20, 4, 2), (26, 5, 3), (14, 2, 1)])]
    rng = np.random.default_rng(5)
    pose_lig, xyz = [], []
    for l, lig in enumerate(ligs):
        for _ in range(3):
            c = rec.pocket_center + rng.normal(size=3) * 1.5
            R = synth.quat_to_rot(synth.random_quaternion(rng))
            t = rng.uniform(-math.pi, math.pi, lig.tree.n_torsions)
            xyz.append(_f32(synth.ligand_apply(lig.tree, lig.xyz[0], c, R, t)))
            pose_lig.append(l)
    tors = _f32(rng.uniform(-1, 1, (len(xyz), rec.leaves.n_leaves)))
    Rh = lib.Receptor(rec)
    B = lib.LigandBatch(ligs)
    P = lib.PoseSet(B, pose_lig, xyz, rec.leaves.n_leaves, tors=tors)
    return dict(rec=rec, ligs=ligs, pose_lig=pose_lig, xyz=xyz, tors=tors, R=Rh, B=B, P=P,
                sp=scoring.load_pose_selection())


def _oracle_score(tw, p, x, sp):
    rx = O.receptor_apply(tw["rec"].leaves, tw["rec"].xyz, tw["tors"][p])
    s, _ = O.pscore(sp, tw["rec"], rx, tw["ligs"][tw["pose_lig"][p]], x)
    return s


def test_gpu_pscore_grad_value_equals_pscore(tw):
    s0, _ = lib.pscore(tw["R"], tw["P"], tw["sp"], with_terms=False)
    s1, _, _ = lib.pscore_grad(tw["R"], tw["P"], tw["sp"])
    s0, s1 = s0.cpu().numpy(), s1.cpu().numpy()
    np.testing.assert_allclose(s1, s0, rtol=1e-5, atol=1e-5)
    for p in range(len(s0)):
        ref = _oracle_score(tw, p, tw["xyz"][p], tw["sp"])
        assert abs(s1[p] - ref) <= 1e-4 * max(abs(ref), 1.0)


def test_gpu_pscore_coordinate_gradient_matches_fd(tw):
    _, dx, _ = lib.pscore_grad(tw["R"], tw["P"], tw["sp"])
    dx = dx.cpu().numpy().reshape(-1, 3)
    h = 1e-6
    for p in range(len(tw["xyz"])):
        x = tw["xyz"][p]
        fd = np.zeros_like(x)
        for a in range(len(x)):
            for d in range(3):
                xp, xm = x.copy(), x.copy()
                xp[a, d] += h
                xm[a, d] -= h
                fd[a, d] = (_oracle_score(tw, p, xp, tw["sp"]) -
                            _oracle_score(tw, p, xm, tw["sp"])) / (2 * h)
        o = tw["P"].pose_atom_off[p]
        got = dx[o:o + len(x)]
        err = np.linalg.norm(got - fd)
        assert err <= 1e-3 * max(np.linalg.norm(fd), 1.0), (p, err, np.linalg.norm(fd),
                                                             np.abs(got - fd).max())


def test_gpu_pscore_parameter_gradient_matches_fd(tw):
    _, _, dp = lib.pscore_grad(tw["R"], tw["P"], tw["sp"])
    dp = dp.cpu().numpy()
    theta = train.params_to_vector(tw["sp"])
    types = sorted(set(int(t) for m in tw["ligs"] for t in m.vdw_type))
    idx = list(range(ffi.NUM_TERMS)) + [ffi.NUM_TERMS + t for t in types] + \
        [ffi.NUM_TERMS + ffi.NUM_VDW_TYPES + t for t in types]
    h = 1e-6
    for p in range(len(tw["xyz"])):
        fd = np.zeros(len(idx))
        for k, i in enumerate(idx):
            vp, vm = theta.copy(), theta.copy()
            vp[i] += h
            vm[i] -= h
            fd[k] = (_oracle_score(tw, p, tw["xyz"][p], train.vector_to_params(vp, tw["sp"])) -
                     _oracle_score(tw, p, tw["xyz"][p], train.vector_to_params(vm, tw["sp"]))) / (2 * h)
        got = dp[p, idx]
        err = np.linalg.norm(got - fd)
        assert err <= 1e-3 * max(np.linalg.norm(fd), 1.0), (p, err, np.abs(got - fd).max())
        # types absent from the pose and inactive terms have no derivative
        others = np.setdiff1d(np.arange(lib.NUM_SCORE_PARAMS), idx)
        assert (dp[p, others] == 0).all()


def test_gpu_autograd_function_and_training_step(tw):
    import torch
    dev = tw["P"].xyz.device
    xyz = tw["P"].xyz.clone().requires_grad_(True)
    theta = torch.tensor(train.params_to_vector(tw["sp"]), device=dev, requires_grad=True)
    rng = np.random.default_rng(9)
    target = torch.tensor(rng.normal(-8.0, 2.0, tw["P"].n_poses), device=dev)
    s = train.pose_score(xyz, theta, tw["R"], tw["P"], tw["sp"])
    loss = ((s - target) ** 2).sum()
    loss.backward()
    # the chain rule by hand from the kernel's own derivatives
    s2, dx, dp = lib.pscore_grad(tw["R"], tw["P"], tw["sp"])
    gs = 2.0 * (s2 - target)
    atom_pose = np.repeat(np.arange(tw["P"].n_poses), tw["B"].natoms[tw["P"].pose_lig])
    want_x = (dx.view(-1, 3).cpu().numpy() * gs.cpu().numpy()[atom_pose][:, None]).ravel()
    np.testing.assert_allclose(xyz.grad.cpu().numpy(), want_x, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(theta.grad.cpu().numpy(), (gs @ dp).cpu().numpy(), rtol=1e-10,
                               atol=1e-10)
    # a few steps on the parameters lower the loss
    theta2 = theta.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([theta2], lr=1e-3)
    losses = []
    for _ in range(20):
        opt.zero_grad()
        s = train.pose_score(tw["P"].xyz, theta2, tw["R"], tw["P"], tw["sp"])
        lo = ((s - target) ** 2).sum()
        lo.backward()
        opt.step()
        losses.append(float(lo))
    assert losses[-1] < losses[0], losses
