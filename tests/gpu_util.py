"""Shared helpers for the GPU parity tests (fixtures built from the synthetic scene)."""
import numpy as np

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import ICPTracker, TSDFParams, TSDFVolume


def small_params(res=128, vs=0.008, trunc=None):
    """A res^3 volume over the 1.024 m cube centred on the sphere."""
    trunc = 3 * vs if trunc is None else trunc
    return dict(resolution=res, voxel_size=vs, truncation=trunc, max_weight=128.0, origin=sc.ORIGIN,
                min_depth=0.3, max_depth=5.0)


def gpu_volume(p, fmt=N.FMT_F32, mode=N.MODE_AUTO):
    return TSDFVolume(TSDFParams(**p), fmt=fmt, mode=mode)


def oracle_volume(p, dtype=np.float32):
    return oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                         p["min_depth"], p["max_depth"], dtype=dtype)


def frames(n, boxes=10, jitter=1, W=640, H=480):
    out = []
    for f in range(n):
        T = sc.pose(f, jitter)
        out.append((sc.render(T, sc.K640, W, H, boxes, sc.SEED, True, f), T))
    return out


def say(*a):
    print(*a, flush=True)


def pose_diff(Ta, Tb):
    """(max abs translation difference in m, rotation angle between them in rad), in float64."""
    Ta = np.asarray(Ta, np.float64)
    Tb = np.asarray(Tb, np.float64)
    dt = float(np.abs(Ta[:3, 3] - Tb[:3, 3]).max())
    dR = Ta[:3, :3] @ Tb[:3, :3].T
    s = 0.5 * np.array([dR[2, 1] - dR[1, 2], dR[0, 2] - dR[2, 0], dR[1, 0] - dR[0, 1]])
    ang = float(np.arctan2(np.linalg.norm(s), 0.5 * (np.trace(dR) - 1.0)))
    return dt, ang


def steps_per_level(packed):
    return tuple((int(packed) >> (8 * l)) & 255 for l in range(3))


def icp_pose_parity(gpu_pose, gpu_steps, ref_pose, ref_steps):
    """The ICP parity rule. When both runs take the same accepted steps per level the
    poses agree within 1e-5 m / 1e-5 rad. The stop test |x| < 5e-5 is discrete, and the
    reference's own result moves by one step when its input moves by one float ulp, so
    when the step counts differ the poses must agree within one step: 1e-4 m / 1e-4 rad.
    Returns (ok, dt, ang, same_steps)."""
    dt, ang = pose_diff(gpu_pose, ref_pose)
    same = steps_per_level(gpu_steps) == steps_per_level(ref_steps)
    tol = 1e-5 if same else 1e-4
    return (dt < tol and ang < tol), dt, ang, same
