"""The raycast-depth loss adjoint at its full workload: a 512^3 FP32 TSDF at 2 mm,
640 x 480 rays, step 0.5 voxel, march recomputed in the backward.

* Plane closed form: on a planar TSDF (linear in every cell a bracket uses) the secant is
  exact, so L is the sum of the squared offsets and dL/dxi follows from the reference's
  point-to-plane ICP Jacobian J = [n_c ; q x n_c] (src/tracking/ICPTracker.cpp:307-312)
  at each hit q: d(depth)/dxi = -J / ((n_c . r_hat) |r|).
* Central finite differences (h = 1e-4 on the six pose components, 1e-3 on the TSDF
  voxels the loss depends on most) on a volume fused from ten frames of the 512^3
  sequence, observed with the tenth frame's depth, over the pixels whose depth is smooth
  under the +-h pose changes.
"""
import ctypes

import numpy as np
import pytest

from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import TSDFParams, TSDFVolume
from tests.adjoint_autograd import LossProblem
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

RES, VS, TAU = 512, 0.002, 0.008
ORIGIN = sc.ORIGIN


def test_plane_closed_form_512():
    n_w = np.array([0.15, -0.1, -1.0])
    n_w /= np.linalg.norm(n_w)
    # z = 0.7 + 0.15 x - 0.1 y in [0.6, 0.8] over the image: every ray meets it inside
    # the volume (|x| <= 0.61 z < 0.512, |y| <= 0.46 z)
    p0 = np.array([0.0, 0.0, 0.7])
    ax = np.arange(RES, dtype=np.float64) * VS
    # s(x, y, z) = n . (X - p0), a sum of one term per axis
    sx = n_w[0] * (ORIGIN[0] + ax - p0[0])
    sy = n_w[1] * (ORIGIN[1] + ax - p0[1])
    sz = n_w[2] * (ORIGIN[2] + ax - p0[2])
    t = np.clip((sz[:, None, None] + sy[None, :, None] + sx[None, None, :]) / TAU, -1, 1).astype(np.float32)
    vol = TSDFVolume(TSDFParams(RES, VS, TAU, 128.0, ORIGIN, 0.3, 5.0), fmt=N.FMT_F32)
    vol.upload(t, np.ones_like(t))
    del t
    K = sc.K640
    u, v = np.meshgrid(np.arange(640, dtype=np.float64), np.arange(480, dtype=np.float64))
    rc = np.stack([(u - K[2]) / K[0], (v - K[3]) / K[1], np.ones_like(u)], -1)
    rn = np.linalg.norm(rc, axis=-1)
    d_true = (n_w @ p0) / (rc @ n_w)
    offset = 0.004 + 0.002 * np.sin(rc[..., 0] * 7)
    prob = LossProblem(vol, np.eye(4), K, 640, 480, (d_true + offset).astype(np.float32))
    out, _ = prob.run(np.zeros(6), backward=True)
    r = -offset
    L_ref = float(np.sum(r * r))
    rhat = rc / rn[..., None]
    q = rhat * (d_true * rn)[..., None]
    J = np.concatenate([np.broadcast_to(n_w, q.shape), np.cross(q, n_w)], -1)
    g_ref = np.sum((2 * r)[..., None] * (-J / ((rhat @ n_w) * rn)[..., None]), axis=(0, 1))
    say("512^3 plane: loss %.9g (closed form %.9g), pixels %d" % (out[0], L_ref, out[7]))
    say("grad", out[1:7], "ICP-Jacobian form", g_ref)
    assert out[7] == 640 * 480
    assert abs(out[0] - L_ref) <= 1e-4 * L_ref
    np.testing.assert_allclose(out[1:7], g_ref, rtol=2e-3, atol=1e-6 * np.abs(g_ref).max())
    vol.close()


@pytest.fixture(scope="module")
def fused():
    frames, truth, K = sc.sequence("c2_512", frames=10)
    p = sc.volume_params("c2_512")
    vol = TSDFVolume(TSDFParams(**p), fmt=N.FMT_F32)
    for k in range(10):
        d = np.where((frames[k] > 0), frames[k].astype(np.float32) * np.float32(0.001), np.float32(0))
        vol.integrate(d, K, truth[k])
    obs = np.where(frames[9] > 0, frames[9].astype(np.float32) * np.float32(0.001), np.float32(0)).astype(np.float32)
    # pixels away from the border and from depth edges
    dd = obs.astype(np.float64)
    smooth = np.ones_like(dd, bool)
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0), (2, 0), (0, 2), (-2, 0), (0, -2)):
        smooth &= np.abs(np.roll(np.roll(dd, dy, 0), dx, 1) - dd) < 0.01
    smooth[:8] = smooth[-8:] = False
    smooth[:, :8] = smooth[:, -8:] = False
    obs = np.where(smooth, obs, np.float32(0))
    prob = LossProblem(vol, truth[9], K, 640, 480, obs)
    # Residuals of more than 1 cm (the model's holes and edges against this frame) are
    # left out: a single 20 cm residual would dominate the voxels it reads, and its
    # difference quotient would resolve its float32 depth change only to ~10 %.
    d0 = prob.depth(np.zeros(6))
    obs = np.where(np.abs(d0 - obs) <= 0.01, obs, np.float32(0)).astype(np.float32)
    prob.set_obs(obs)
    _, keep = prob.mask_discontinuities(obs, 1e-4)
    say("512^3 fused: differentiable observed pixels", int(keep.sum()), "of", int((obs > 0).sum()))
    assert keep.sum() > 0.5 * (obs > 0).sum()
    yield prob
    vol.close()


def test_pose_gradient_finite_differences_512(fused):
    prob = fused
    out, _ = prob.run(np.zeros(6), backward=True)
    h = 1e-4
    fd = np.zeros(6)
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        fd[i] = (prob.run(e)[0][0] - prob.run(-e)[0][0]) / (2 * h)
    for i in range(6):
        say("d/dxi[%d] adjoint %+.6e  central difference %+.6e" % (i, out[1 + i], fd[i]))
    np.testing.assert_allclose(out[1:7], fd, rtol=5e-2, atol=2e-2 * np.abs(fd).max())


def test_tsdf_gradient_finite_differences_512(fused):
    """The voxels with the largest |dL/dtsdf|: the loss is piecewise smooth in a voxel too
    (an edit can flip a sample's sign and move a ray's bracket to the next sample), so a
    voxel counts only where the one-sided differences L(+h) - L(0) and L(0) - L(-h) agree
    within 10 %, as the pose test keeps only the pixels smooth under its steps; at least
    eight of the 16 largest must."""
    prob = fused
    L0, g = prob.run(np.zeros(6), backward=True)
    L0 = L0[0]
    cand = np.argsort(-np.abs(g))[:16]
    gmax = float(np.abs(g[cand[0]]))
    t0 = prob.tsdf.reshape(-1)
    # A voxel edit of h moves the hits of the few rays whose brackets read it by about
    # h * truncation (6 um here), some 100 float32 ulps of a 0.8 m depth: the difference
    # quotient carries about 1 % of quantisation noise at the largest gradient (more,
    # relatively, at smaller ones), hence the absolute term of 1 % of the largest.
    h = 1e-3
    checked = 0
    for i in cand:
        vp, vm = np.float32(t0[i] + h), np.float32(t0[i] - h)
        Lp = prob.run(np.zeros(6), vals=[vp], idx=[i])[0][0]
        Lm = prob.run(np.zeros(6), vals=[vm], idx=[i])[0][0]
        fd = (Lp - Lm) / (float(vp) - float(vm))
        smooth = abs((Lp - L0) - (L0 - Lm)) <= 0.1 * abs(Lp - Lm)
        say("voxel %d: adjoint %+.6e central difference %+.6e%s" % (i, g[i], fd, "" if smooth else " (not smooth)"))
        if smooth:
            checked += 1
            assert abs(fd - g[i]) <= 0.05 * abs(g[i]) + 1e-2 * gmax, (i, fd, g[i])
    prob.set_voxels([], [])
    assert checked >= 8


def test_skip_bounds_leave_the_loss_unchanged_512(fused):
    """The march passes over the lattice samples the brick map's per-pixel bounds prove
    far; with skipping off every sample is read. Ray depths, the loss, the contributing
    pixels and the (fixed-order) pose gradient are bit-identical either way."""
    prob = fused
    vol = prob.vol
    runs = {}
    for on in (True, False):
        vol.set_skip(on)
        out, _ = prob.run(np.zeros(6), backward=True)
        runs[on] = (out, prob.depth(np.zeros(6)))
    vol.set_skip(True)
    (a, da), (b, db) = runs[True], runs[False]
    say("loss with / without skipping: %.12g / %.12g, pixels %d / %d" % (a[0], b[0], a[7], b[7]))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
