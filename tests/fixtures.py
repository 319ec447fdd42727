"""Fixtures re-expressed from the reference's own CPU contracts, with the same numbers.

Each builder returns plain numpy inputs plus the closed-form expectations the reference
test derives by hand, so the same fixture checks the CPU oracle and the GPU path.
"""
import numpy as np

FX, FY, CX, CY = 525.0, 525.0, 319.5, 239.5
K640 = (FX, FY, CX, CY)


# ---------------------------------------------------------------------------
# tests/icp_weighting_contract.cpp:62-153 - 11 correspondences at z = 1 m, model
# normals +z, residual = depth gap: 10 inliers (e = 0.01, w = 1) as five pixel pairs
# point-symmetric about the optical centre, one outlier (e = 0.08, w = 0.25). With one
# level-0 iteration and the damping escalated to 1.0 (zero eigenvalues):
#   t_z = -0.12 / 11.25, error = 0.0038 / 11.
# ---------------------------------------------------------------------------
ICP_TZ = -0.12 / 11.25
ICP_TZ_UNWEIGHTED = -0.12 / 12.0
ICP_TZ_W2 = -0.105 / 11.0625
ICP_ERR = 0.0038 / 11.0
ICP_ERR_WE2 = 0.0014 / 11.0
ICP_TOL = 5e-6
ICP_GAP = 2e-4


def icp_weighting(corruption="clean"):
    W, H = 640, 480
    lv = np.zeros((H, W, 3), np.float32)
    ln = np.zeros((H, W, 3), np.float32)
    mv = np.zeros((H, W, 3), np.float32)
    mn = np.zeros((H, W, 3), np.float32)

    def put(px, py, err, inf_normal=False, nan_depth=False):
        z = np.float32(1.0)
        vx = (np.float32(px) - np.float32(CX)) / np.float32(FX) * z
        vy = (np.float32(py) - np.float32(CY)) / np.float32(FY) * z
        lv[py, px] = (vx, vy, z)
        ln[py, px] = (0, 0, 1)
        mv[py, px] = (vx, vy, z - np.float32(err))
        if nan_depth:
            mv[py, px, 2] = np.nan
        mn[py, px] = (0, 0, np.inf) if inf_normal else (0, 0, 1)

    for k in range(5):
        put(319 - k, 239 - k, 0.01)
        put(320 + k, 240 + k, 0.01)
    put(320, 239, 0.08)
    if corruption == "inf_normal":
        put(340, 260, 0.01, inf_normal=True)
    elif corruption == "nan_depth":
        put(300, 220, 0.01, nan_depth=True)
    levels_v = [lv, np.zeros((H // 2, W // 2, 3), np.float32), np.zeros((H // 4, W // 4, 3), np.float32)]
    levels_n = [ln, np.zeros((H // 2, W // 2, 3), np.float32), np.zeros((H // 4, W // 4, 3), np.float32)]
    return levels_v, levels_n, mv, mn


# ---------------------------------------------------------------------------
# tests/icp_weighting_contract.cpp:277-391 - pyramid depth-jump guard on a 16x16 field.
# ---------------------------------------------------------------------------
def pyramid_field():
    W = H = 16
    d = np.ones((H, W), np.float32)
    v = np.zeros((H, W, 3), np.float32)
    n = np.zeros((H, W, 3), np.float32)
    for y in range(H):
        for x in range(W):
            v[y, x] = (np.float32(0.001) * np.float32(x), np.float32(0.002) * np.float32(y), 1.0)
            n[y, x] = (0, 0, 1)
    d[0, 1] = 1.1
    d[0, 2], d[0, 3], d[1, 2], d[1, 3] = 1.3, 1.2, 1.1, 1.0
    d[2, 1] = 1.025
    d[3, 3] = 0.0
    d[4, 4] = d[4, 5] = d[5, 4] = d[5, 5] = 0.0
    return d, v, n


def check_pyramid_field(src, l1):
    """Asserts of icp_weighting_contract.cpp:338-375 on level 1 = (depth, vertices, normals)."""
    d0, v0, _ = src
    d1, v1, n1 = l1
    assert d1.shape == (8, 8)
    assert v1[0, 0, 2] == 1.0 and d1[0, 0] == 1.0  # nearest, not the 1.025 average
    assert np.array_equal(v1[0, 0], v0[0, 0])
    assert d1[0, 1] == 1.0  # (x=1, y=0): nearest 1.0, not the 1.15 average
    assert np.array_equal(v1[0, 1], v0[1, 3])  # the minimum sat at the far corner
    assert np.array_equal(n1[0, 0], [0, 0, 1]) and np.array_equal(n1[0, 1], [0, 0, 1])
    assert abs(float(d1[1, 0]) - (1.0 + 1.025 + 1.0 + 1.0) / 4.0) <= 1e-6
    smooth_v = (v0[2, 0] + v0[2, 1] + v0[3, 0] + v0[3, 1]) * np.float32(0.25)
    assert np.linalg.norm(v1[1, 0] - smooth_v) <= 1e-6
    assert abs(float(d1[1, 1]) - 1.0) <= 1e-6  # three valid samples
    assert d1[2, 2] == 0.0 and not v1[2, 2].any() and not n1[2, 2].any()
    assert abs(float(d1[3, 3]) - 1.0) <= 1e-6 and v1[3, 3, 2] == 1.0


# ---------------------------------------------------------------------------
# tests/tsdf_integration_race_contract.cpp:232-324 - the hand-derived voxel fold: a 4x4
# frame with one valid pixel on the optical axis, two frames folded into a 64^3 volume.
# ---------------------------------------------------------------------------
FOLD_PARAMS = dict(resolution=64, voxel_size=0.01, truncation=0.03, max_weight=128.0, origin=(-0.32, -0.32, 0.0),
                   min_depth=0.3, max_depth=5.0)
FOLD_K = (500.0, 500.0, 2.0, 2.0)
FOLD_DEPTHS = (0.5015, 0.5065)
FOLD_BAND = (0.1, 3.0)


def fold_frames():
    out = []
    for D in FOLD_DEPTHS:
        d = np.zeros((4, 4), np.float32)
        d[2, 2] = D
        out.append(d)
    return out


def fold_expect(k):
    """(tsdf, weight) of axis voxel k after both frames, folded in double."""
    vs, trunc = 0.01, 0.03
    t, w = 1.0, 0.0
    for D in FOLD_DEPTHS:
        sdf = float(np.float32(D)) - float(np.float32(k) * np.float32(vs))
        if sdf < -trunc:
            continue
        tn = min(1.0, sdf / trunc)
        t = (t * w + tn) / (w + 1.0)
        w = min(w + 1.0, 128.0)
    return t, w


def check_fold(tsdf, weight):
    """Asserts of tsdf_integration_race_contract.cpp:296-318 on (z, y, x) arrays."""
    ox = oy = 32
    surface = 0
    for k in range(1, 64):
        t, w = fold_expect(k)
        assert weight[k, oy, ox] == np.float32(w), k
        assert abs(float(tsdf[k, oy, ox]) - t) < 2e-5, (k, float(tsdf[k, oy, ox]), t)
        if w == 2.0 and abs(t) < 0.5:
            surface += 1
    assert surface >= 2
    assert weight[50, oy, ox] == 2.0
    assert weight[54, oy, ox] == 0.0
    assert weight[50, oy, ox + 1] == 0.0
    # nothing off the axis was touched
    touched = np.argwhere(weight > 0)
    assert np.all(touched[:, 1] == oy) and np.all(touched[:, 2] == ox)
