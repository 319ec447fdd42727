"""Test-only autograd wrapper of the raycast-depth loss adjoint (C ABI), for gradcheck.

DepthLoss(xi, vals) evaluates L = sum (d_ray - d_obs)^2 for the pose pose0 * exp(xi)
(xi = [t; w], the ICP increment convention) over a volume whose voxels `idx` hold
`vals` (every other voxel fixed). Forward and backward both run on the GPU; tensors
are CPU float64 (autograd plumbing only) and are converted to the kernel's float32.
"""
import ctypes

import numpy as np
import torch

from azu_amd import native as N
from azu_amd.fusion import pose_cm


def se3_right(T, xi):
    """T * [R(w) | t] with R(w) by Rodrigues, in float64."""
    xi = np.asarray(xi, np.float64)
    t, w = xi[:3], xi[3:]
    th = np.linalg.norm(w)
    if th < 1e-12:
        R = np.eye(3) + np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    else:
        k = w / th
        Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx
    D = np.eye(4)
    D[:3, :3] = R
    D[:3, 3] = t
    return np.asarray(T, np.float64) @ D


class LossProblem:
    """A volume + observation set-up shared by the forward and backward calls."""

    def __init__(self, vol, pose0, K, W, H, d_obs, step_voxels=0.5):
        self.vol = vol
        self.pose0 = np.asarray(pose0, np.float64)
        self.K = N.f32(K)
        self.W, self.H = W, H
        self.obs = N.device_array(N.f32(d_obs))
        self.step = step_voxels
        self.res = vol.res
        self.tsdf, self.weight = vol.voxelData()
        self._edited = set()

    def set_voxels(self, idx, vals):
        """The volume with voxels idx set to vals and every other voxel at its original
        value (voxels edited by an earlier call and not in idx are restored), written
        voxel by voxel (azu_vol_write_voxels) rather than by a whole-volume upload."""
        flat_t, flat_w = self.tsdf.reshape(-1), self.weight.reshape(-1)
        edits = dict(zip((int(i) for i in idx), (np.float32(v) for v in vals)))
        for i in sorted(self._edited - set(edits)):
            edits_i = np.array([flat_t[i]], np.float32), np.array([flat_w[i]], np.float32)
            N.check(N.lib.azu_vol_write_voxels(self.vol.h, i, 1, edits_i[0].ctypes.data, edits_i[1].ctypes.data))
        for i, v in edits.items():
            t1, w1 = np.array([v], np.float32), np.array([flat_w[i]], np.float32)
            N.check(N.lib.azu_vol_write_voxels(self.vol.h, i, 1, t1.ctypes.data, w1.ctypes.data))
        self._edited = set(edits)

    def set_obs(self, d_obs):
        self.obs = N.device_array(N.f32(d_obs))

    def depth(self, xi):
        """Ray depths (H, W) at pose0 * exp(xi), 0 on a miss."""
        dbuf = N.DeviceBuffer(self.W * self.H * 4)
        T = se3_right(self.pose0, xi).astype(np.float32)
        out = (ctypes.c_double * 8)()
        N.check(N.lib.azu_raycast_depth_loss(self.vol.h, N.fptr(pose_cm(T)), N.fptr(self.K), self.W, self.H,
                                             self.obs.ptr, self.step, 0, None, dbuf.ptr, out, None))
        return dbuf.download(np.float32, (self.H, self.W))

    def run(self, xi, backward=False, vals=None, idx=None):
        if vals is not None:
            self.set_voxels(np.atleast_1d(idx), np.atleast_1d(vals))
        T = se3_right(self.pose0, xi).astype(np.float32)
        out = (ctypes.c_double * 8)()
        grad = None
        gbuf = None
        if backward:
            gbuf = N.DeviceBuffer(self.res ** 3 * 4)
            zeros = np.zeros(self.res ** 3, np.float32)  # referenced until the copy has run
            N.check(N.lib.azu_memcpy_h2d(gbuf.ptr, zeros.ctypes.data, zeros.nbytes, None))
            N.synchronize()
        N.check(N.lib.azu_raycast_depth_loss(self.vol.h, N.fptr(pose_cm(T)), N.fptr(self.K), self.W, self.H,
                                             self.obs.ptr, self.step, 1 if backward else 0,
                                             gbuf.ptr if gbuf else None, None, out, None))
        if backward:
            grad = gbuf.download(np.float32, (self.res ** 3,))
        return np.array(out[:]), grad

    def mask_discontinuities(self, d_obs, h):
        """Drops the observations of pixels whose ray depth is not smooth under a +-h pose
        change along any of the six axes (hit/miss switches at silhouettes, bracket
        switches on a curved field): the one-sided changes d(+h) - d(0) and d(0) - d(-h)
        of a differentiable pixel agree to first order. The loss is differentiable on the
        rest."""
        keep = np.asarray(d_obs) > 0
        d0 = self.depth(np.zeros(6)).astype(np.float64)
        keep &= d0 > 0
        for i in range(6):
            e = np.zeros(6)
            e[i] = h
            dp, dm = self.depth(e).astype(np.float64), self.depth(-e).astype(np.float64)
            hit = (dp > 0) & (dm > 0)
            curv = np.abs((dp - d0) - (d0 - dm))
            keep &= hit & (curv <= 0.1 * np.abs(dp - dm) + 2e-7)
        masked = np.where(keep, d_obs, np.float32(0)).astype(np.float32)
        self.set_obs(masked)
        return masked, keep


class DepthLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, vals, prob, idx):
        out, _ = prob.run(xi.detach().numpy(), False, vals.detach().numpy().astype(np.float32), idx)
        ctx.save_for_backward(xi, vals)
        ctx.prob, ctx.idx = prob, idx
        return torch.tensor(out[0], dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        xi, vals = ctx.saved_tensors
        out, grad = ctx.prob.run(xi.detach().numpy(), True, vals.detach().numpy().astype(np.float32), ctx.idx)
        gxi = torch.tensor(out[1:7], dtype=torch.float64) * g
        gvals = torch.tensor(grad[ctx.idx].astype(np.float64)) * g
        return gxi, gvals, None, None
