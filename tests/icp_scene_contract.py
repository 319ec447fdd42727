"""The reference's ICP contracts on its synthetic room (tests/icp_refpose_contract.cpp:
33-90, tests/icp_lowres_model_contract.cpp:40-136), shared by the oracle and the GPU.

`ops` supplies a fused volume (`fuse(depth, pose, n)`), a model raycast at a size
(`model(pose, w, h, K)` -> vertices, normals) and a track
(`track(depth, model, est, ref, iters, dist, angle)` -> result dict with a 4x4 pose,
tracking_ok, inliers)."""
from tests import room_scene as rs

K = (525.0, 525.0, 319.5, 239.5)


def refpose(ops):
    """A: with the model's raycast pose as reference, ICP converges within 1 mm /
    0.05 deg and succeeds; B: with the previous pose as reference it is detectably worse."""
    ops.fuse(rs.render_depth(rs.P_MODEL), rs.P_MODEL, 5)
    model = ops.model(rs.P_MODEL, 640, 480, K)
    live = rs.render_depth(rs.P_LIVE_REF)
    good = ops.track(live, model, rs.P_PREV, rs.P_MODEL)
    eg = rs.pose_error(good["pose"], rs.P_LIVE_REF)
    assert good["tracking_ok"] and eg[0] < 1e-3 and eg[1] < 0.05, eg
    bad = ops.track(live, model, rs.P_PREV, rs.P_PREV)
    eb = rs.pose_error(bad["pose"], rs.P_LIVE_REF)
    assert not bad["tracking_ok"] or eb[0] > 5e-3 or eb[1] > 0.25, eb
    return good, bad


def lowres(ops):
    """B: a 640 x 480 model within 1 mm / 0.05 deg; C: a 320 x 240 model (intrinsics
    scaled) within 2 mm / 0.1 deg; D: level 2 only, 30 iterations, from 8 deg / 3 cm
    against a 160 x 120 model lands within 1 cm / 0.5 deg; E: the 160 x 120 model
    rendered with unscaled intrinsics is measurably wrong."""
    ops.fuse(rs.render_depth(rs.P_MODEL), rs.P_MODEL, 5)
    live = rs.render_depth(rs.P_LIVE_LOWRES)
    out = {}
    for name, (w, h), tol in (("B", (640, 480), (1e-3, 0.05)), ("C", (320, 240), (2e-3, 0.1))):
        model = ops.model(rs.P_MODEL, w, h, rs.scale_intrinsics(K, 640, 480, w, h))
        r = ops.track(live, model, rs.P_MODEL, rs.P_MODEL)
        e = rs.pose_error(r["pose"], rs.P_LIVE_LOWRES)
        assert e[0] < tol[0] and e[1] < tol[1], (name, e)
        out[name] = r
    far = rs.render_depth(rs.P_FAR)
    coarse = dict(iters=(0, 0, 30), dist=0.25, angle=45.0)
    small = ops.model(rs.P_MODEL, 160, 120, rs.scale_intrinsics(K, 640, 480, 160, 120))
    rd = ops.track(far, small, rs.P_MODEL, rs.P_MODEL, **coarse)
    ed = rs.pose_error(rd["pose"], rs.P_FAR)
    assert ed[0] < 0.01 and ed[1] < 0.5, ed
    wrong = ops.model(rs.P_MODEL, 160, 120, K)
    re = ops.track(far, wrong, rs.P_MODEL, rs.P_MODEL, **coarse)
    ee = rs.pose_error(re["pose"], rs.P_FAR)
    assert re["inliers"] < 100 or ee[0] > 0.02 or ee[1] > 1.0, ee
    out["D"], out["E"] = rd, re
    return out
