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
