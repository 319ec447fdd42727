"""The reference tests' synthetic room scene (tests/support/SyntheticScene.h:21-127 of
iphands/azu), restated in numpy: a room box seen from inside, a box and a sphere,
rendered as Z-depth by sphere tracing; makePose and poseError as that header defines
them. Used by the ICP reference-pose and low-resolution-model contracts."""
import numpy as np

F = np.float32


def _box(p, c, h):
    q = np.abs(p - np.asarray(c, F)) - np.asarray(h, F)
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)


def sdf(p):
    room = -_box(p, (0.0, -0.05, 1.15), (1.2, 0.95, 1.25))
    box = _box(p, (0.35, 0.55, 1.5), (0.3, 0.35, 0.25))
    sphere = np.linalg.norm(p - np.asarray((-0.5, 0.2, 1.7), F), axis=-1) - F(0.3)
    return np.minimum(np.minimum(room, box), sphere)


def render_depth(pose, K=(525.0, 525.0, 319.5, 239.5), W=640, H=480):
    """renderDepth (SyntheticScene.h:35-66) without noise: float32 Z-depth, 0 = miss."""
    fx, fy, cx, cy = (F(k) for k in K)
    v, u = np.mgrid[0:H, 0:W].astype(F)
    rc = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], -1).reshape(-1, 3)
    norm = np.linalg.norm(rc, axis=-1).astype(F)
    T = np.asarray(pose, F)
    d = (rc @ T[:3, :3].T) / norm[:, None]
    o = T[:3, 3]
    t = np.full(len(rc), F(0.05))
    depth = np.zeros(len(rc), F)
    active = np.ones(len(rc), bool)
    for _ in range(256):
        idx = np.nonzero(active & (t < 8.0))[0]
        if idx.size == 0:
            break
        s = sdf(o + d[idx] * t[idx, None]).astype(F)
        hit = s < F(1e-5)
        depth[idx[hit]] = t[idx[hit]] / norm[idx[hit]]
        active[idx[hit]] = False
        t[idx[~hit]] += s[~hit]
    return depth.reshape(H, W)


def make_pose(axis_angle, t):
    """makePose (SyntheticScene.h:113-119): rotation by the axis-angle vector, then t."""
    a = np.asarray(axis_angle, np.float64)
    T = np.eye(4)
    n = np.linalg.norm(a)
    if n > 0:
        k = a / n
        Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        T[:3, :3] = np.eye(3) + np.sin(n) * Kx + (1 - np.cos(n)) * Kx @ Kx
    T[:3, 3] = t
    return T.astype(F)


def pose_error(a, b):
    """poseError (SyntheticScene.h:121-127): |t| of a^-1 b in m, its rotation angle in deg."""
    d = np.linalg.inv(np.asarray(a, np.float64)) @ np.asarray(b, np.float64)
    c = np.clip((np.trace(d[:3, :3]) - 1) * 0.5, -1, 1)
    return float(np.linalg.norm(d[:3, 3])), float(np.degrees(np.arccos(c)))


def scale_intrinsics(K, rw, rh, w, h):
    """scaleIntrinsics (include/sensor/CameraIntrinsics.h:33-40) in float32."""
    if (w, h) == (rw, rh):
        return tuple(float(k) for k in K)
    sx, sy = F(w) / F(rw), F(h) / F(rh)
    fx, fy, cx, cy = (F(k) for k in K)
    return (float(fx * sx), float(fy * sy), float((cx + F(0.5)) * sx - F(0.5)), float((cy + F(0.5)) * sy - F(0.5)))


# tests/icp_refpose_contract.cpp:45-66 and tests/icp_lowres_model_contract.cpp:72-121
P_MODEL = make_pose((0.0, 0.05, 0.0), (0.02, 0.0, 0.1))
P_PREV = (P_MODEL.astype(np.float64) @ make_pose((0.02, 0.1, 0.0), (0.03, 0.0, 0.01))).astype(F)
P_LIVE_REF = (P_PREV.astype(np.float64) @ make_pose((0.0, 0.015, 0.003), (0.004, 0.0, 0.002))).astype(F)
P_LIVE_LOWRES = (P_MODEL.astype(np.float64) @ make_pose((0.0, 0.02, 0.004), (0.006, 0.0, 0.003))).astype(F)
P_FAR = (P_MODEL.astype(np.float64) @ make_pose((0.0, 0.1396, 0.0), (0.03, 0.0, 0.0))).astype(F)
VOLUME = dict(resolution=256, voxel_size=0.01, truncation=0.03, max_weight=128.0, origin=(-1.28, -1.28, 0.0),
              min_depth=0.3, max_depth=3.0)
