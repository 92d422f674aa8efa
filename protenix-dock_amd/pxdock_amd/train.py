"""Training binding: the pose score as a differentiable torch function.

The reference has no training code (its scoring function is fixed); this
module fits the pose_selection scoring function on the GPU.  The forward
pass is pxd_pscore_grad_batch (one wavefront per pose), which returns the
scores together with their analytic derivatives with respect to the ligand
coordinates and to the parameter vector

    theta = [9 term coefficients (PXD_TERM_* order),
             vdw_sigma_scale[99], vdw_epsilon_scale[99]]

so the backward pass is two small products: d loss / d xyz = g_p * dS_p/dxyz
per pose and d loss / d theta = g^T (dS/dtheta).  torch only carries the
tensors; every score and derivative comes from the HIP kernel.
"""
import ctypes as C

import numpy as np

from . import ffi, lib

N_PARAMS = lib.NUM_SCORE_PARAMS
COEF_FIELDS = ("coul_coef", "vdw_coef", "ionic_coef", "pication_coef", "f2f_coef", "e2f_coef",
               "hbond_coef", "hydrophobic_coef", "tstrain_coef")


def params_to_vector(sp):
    """pxd_score_params -> theta (numpy FP64)."""
    v = np.zeros(N_PARAMS)
    for k, f in enumerate(COEF_FIELDS):
        v[k] = getattr(sp, f)
    v[ffi.NUM_TERMS:ffi.NUM_TERMS + ffi.NUM_VDW_TYPES] = np.ctypeslib.as_array(sp.vdw_sigma_scale)
    v[ffi.NUM_TERMS + ffi.NUM_VDW_TYPES:] = np.ctypeslib.as_array(sp.vdw_epsilon_scale)
    return v


def vector_to_params(v, base):
    """theta -> a copy of base (pxd_score_params) with those fields replaced."""
    sp = ffi.PxdScoreParams()
    C.memmove(C.addressof(sp), C.addressof(base), C.sizeof(sp))
    v = np.asarray(v, dtype=np.float64)
    for k, f in enumerate(COEF_FIELDS):
        setattr(sp, f, float(v[k]))
    for t in range(ffi.NUM_VDW_TYPES):
        sp.vdw_sigma_scale[t] = float(v[ffi.NUM_TERMS + t])
        sp.vdw_epsilon_scale[t] = float(v[ffi.NUM_TERMS + ffi.NUM_VDW_TYPES + t])
    return sp


def _fn():
    import torch

    class PoseScore(torch.autograd.Function):
        """S = pscore(xyz; theta) for every pose of a PoseSet."""

        @staticmethod
        def forward(ctx, xyz, theta, receptor, poses, base):
            sp = vector_to_params(theta.detach().cpu().numpy(), base)
            saved = poses.xyz
            poses.xyz = xyz.detach().contiguous()
            try:
                s, dx, dp = lib.pscore_grad(receptor, poses, sp)
            finally:
                poses.xyz = saved
            atom_pose = getattr(poses, "_atom_pose", None)
            if atom_pose is None or atom_pose.device != xyz.device:
                # pose index of every atom of the CSR layout, built once per PoseSet
                atom_pose = torch.from_numpy(np.repeat(np.arange(poses.n_poses),
                                                       poses.batch.natoms[poses.pose_lig])).to(
                    xyz.device)
                poses._atom_pose = atom_pose
            ctx.save_for_backward(dx, dp, atom_pose)
            return s

        @staticmethod
        def backward(ctx, gs):
            dx, dp, atom_pose = ctx.saved_tensors
            g_xyz = None
            g_theta = None
            if ctx.needs_input_grad[0]:
                per_atom = gs.to(dx.dtype)[atom_pose]
                g_xyz = (dx.view(-1, 3) * per_atom[:, None]).view(-1)
            if ctx.needs_input_grad[1]:
                # g^T dS/dtheta as a column sum over the poses: a 1 x n_poses
                # by n_poses x N_PARAMS product handed to a GEMM runs its long
                # inner dimension in two tiles, ~100x slower than this
                g_theta = (dp * gs.to(torch.float64)[:, None]).sum(0)
            return g_xyz, g_theta, None, None, None

    return PoseScore


_POSE_SCORE = None


def pose_score(xyz, theta, receptor, poses, base):
    """Differentiable pscore of every pose: xyz (FP32, the PoseSet CSR layout)
    and theta (FP64, N_PARAMS) may require grad."""
    global _POSE_SCORE
    if _POSE_SCORE is None:
        _POSE_SCORE = _fn()
    return _POSE_SCORE.apply(xyz, theta, receptor, poses, base)
