"""The ten scenarios of the reference's ICP numeric-policy contract, re-expressed.

Source: tests/icp_numeric_policy_contract.cpp:84-849 (iphands/azu). Each builder places
correspondences the way that contract's makeCorr does (a live pixel back-projected at a
depth, its model vertex offset along one axis, the model pixel found by the tracker's
floor(f x / z + c + 1/2) projection) and returns the live level-0 arrays, the model
image, the ICP parameters and the closed-form expectations the contract derives by hand
in double. The same scenario then runs through the CPU oracle and the GPU tracker.
"""
import math

import numpy as np

FX, FY, CX, CY = 525.0, 525.0, 319.5, 239.5
K640 = (FX, FY, CX, CY)
W, H = 640, 480
F = np.float32

# The contract's literals (icp_numeric_policy_contract.cpp:84-99).
DAMP_L0, DAMP_OTHER, DAMP_ESC = 0.01, 0.1, 1.0
TRANS_CAP, ROT_CAP = 0.2, 0.5
CONV_STEP, ACCEPT_STEP = 5e-5, 1e-3
HUBER_K = 0.02
TOL_MATCH = 1e-7
GAP_FLOOR = 2e-4
GAP_FLOOR_S = 1e-5


def back_project(px, py, z):
    """backProject (:123-126): ((px - CX) / FX * z, (py - CY) / FY * z, z) in float."""
    z = F(z)
    return np.array([(F(px) - F(CX)) / F(FX) * z, (F(py) - F(CY)) / F(FY) * z, z], F)


def projected_pixel(v):
    """projectedPixel (:116-121): floor(FX * x * (1/z) + CX + 0.5) in float."""
    inv_z = F(1.0) / F(v[2])
    sx = F(F(F(FX) * F(v[0])) * inv_z)
    sy = F(F(F(FY) * F(v[1])) * inv_z)
    mx = math.floor(float(F(F(sx + F(CX)) + F(0.5))))
    my = math.floor(float(F(F(sy + F(CY)) + F(0.5))))
    return mx, my


class Corr:
    """One correspondence (:106-113)."""

    def __init__(self, px, py, z, n, axis, offset):
        # makeCorr (:130-142)
        self.lpx, self.lpy = px, py
        self.lv = back_project(px, py, z)
        self.ln = np.asarray(n, F).copy()
        self.mv = self.lv.copy()
        self.mv[axis] = F(self.lv[axis] - F(offset))
        self.mn = np.asarray(n, F).copy()
        self.mpx, self.mpy = projected_pixel(self.lv)


def corr_arrays(cs):
    """runFixture (:144-166): level 0 of the live pyramid and the model image; every
    other pixel is zero (rejected by the live depth and model validity gates)."""
    lv = np.zeros((H, W, 3), F)
    ln = np.zeros((H, W, 3), F)
    mv = np.zeros((H, W, 3), F)
    mn = np.zeros((H, W, 3), F)
    for c in cs:
        lv[c.lpy, c.lpx] = c.lv
        ln[c.lpy, c.lpx] = c.ln
        mv[c.mpy, c.mpx] = c.mv
        mn[c.mpy, c.mpx] = c.mn
    levels_v = [lv, np.zeros((H // 2, W // 2, 3), F), np.zeros((H // 4, W // 4, 3), F)]
    levels_n = [ln, np.zeros((H // 2, W // 2, 3), F), np.zeros((H // 4, W // 4, 3), F)]
    return levels_v, levels_n, mv, mn


def grid(x0, x1, y0, y1, z, offset, n=(0.0, 0.0, 1.0)):
    """gridFixture (:229-238)."""
    return [Corr(x, y, z, n, 2, offset) for y in range(y0, y1 + 1) for x in range(x0, x1 + 1)]


def rotation_leak(per_row):
    """rotationLeakFixture (:319-329): rows y = 450 (+1 m) and y = 29 (-1 m) at z = 2 m."""
    x0 = 320 - per_row // 2
    cs = [Corr(x0 + i, 450, 2.0, (0, 0, 1), 2, 1.0) for i in range(per_row)]
    cs += [Corr(x0 + i, 29, 2.0, (0, 0, 1), 2, -1.0) for i in range(per_row)]
    return cs


def escalation():
    """testAdaptiveDamping (a) (:471-477): 12 symmetric pairs at 0.01 + one at 0.08."""
    cs = []
    for k in range(6):
        cs.append(Corr(319 - k, 239 - k, 1.0, (0, 0, 1), 2, 0.01))
        cs.append(Corr(320 + k, 240 + k, 1.0, (0, 0, 1), 2, 0.01))
    cs.append(Corr(320, 239, 1.0, (0, 0, 1), 2, 0.08))
    return cs


def well_conditioned():
    """testAdaptiveDamping (b) (:494-512): a 24-pixel +z wall (0.01) and 1600 +x and
    1600 +y pixels (0) at depths {1, 1.5, 2, 2.5}."""
    cs = []
    for x in range(314, 326):
        cs.append(Corr(x, 150, 1.0, (0, 0, 1), 2, 0.01))
        cs.append(Corr(x, 329, 1.0, (0, 0, 1), 2, 0.01))
    depths = (1.0, 1.5, 2.0, 2.5)
    for x in range(300, 340):
        for y in range(200, 240):
            cs.append(Corr(x, y, depths[(x + y) & 3], (1, 0, 0), 0, 0.0))
        for y in range(240, 280):
            cs.append(Corr(x, y, depths[(2 * x + y) & 3], (0, 1, 0), 0, 0.0))
    return cs


def model_normal(extra_normal=None):
    """modelNormalFixture (:541-554): 10 clean correspondences + one at (340, 260) whose
    model normal is extra_normal."""
    cs = []
    for k in range(5):
        cs.append(Corr(319 - k, 239 - k, 1.0, (0, 0, 1), 2, 0.01))
        cs.append(Corr(320 + k, 240 + k, 1.0, (0, 0, 1), 2, 0.01))
    if extra_normal is not None:
        c = Corr(340, 260, 1.0, (0, 0, 1), 2, 0.01)
        c.mn = np.asarray(extra_normal, F).copy()
        cs.append(c)
    return cs


MODEL_NORMAL_CASES = [  # (:567-577) name, normal, accepted
    ("unit +z", (0, 0, 1), True),
    ("boundary in (sq 0.9025)", (0, 0, 0.95), True),
    ("boundary out (sq 0.8836)", (0, 0, 0.94), False),
    ("half length", (0, 0, 0.5), False),
    ("zero", (0, 0, 0), False),
    ("NaN", (0, 0, np.nan), False),
    ("+Inf", (0, 0, np.inf), False),
    ("NaN in x", (np.nan, 0, 1), False),
]


def live_normal(live_n, offset=0.01, dist_offset=0.0):
    """liveNormalFixture (:626-640): a tangent wall, model normals (1, 0, 0), live normal
    live_n, model vertices optionally pushed dist_offset along z."""
    cs = []
    for k in range(5):
        for i in range(2):
            px = 320 + k if i else 319 - k
            py = 240 + k if i else 239 - k
            c = Corr(px, py, 1.0, (1, 0, 0), 0, offset)
            c.ln = np.asarray(live_n, F).copy()
            if dist_offset != 0.0:
                c.mv[2] = F(c.mv[2] + F(dist_offset))
            cs.append(c)
    return cs


def orthogonal_normals():
    """testAngleThresholdClampAndGate (:731-741): model (0, 0, 1), live (1, 0, 0)."""
    cs = []
    for k in range(5):
        for i in range(2):
            px = 320 + k if i else 319 - k
            py = 240 + k if i else 239 - k
            c = Corr(px, py, 1.0, (0, 0, 1), 2, 0.01)
            c.ln = np.array([1, 0, 0], F)
            cs.append(c)
    return cs


ANGLE_CLAMP_CASES = [  # (:711-718) input, sanitised
    (120.0, 85.0), (float("nan"), 30.0), (float("inf"), 30.0), (-5.0, 0.0), (45.0, 45.0), (30.0, 30.0)]


def reflection():
    """testSvdReturnsProperRotation (:779-806): a reflected, non-orthonormal estimate
    diag(-2, 1, 2) and 12 live pixels at (0, 0, 1) whose R v = (0, 0, 2) is the model
    vertex, so the residual is zero and the update only re-orthonormalises."""
    est = np.eye(4, dtype=F)
    est[0, 0], est[1, 1], est[2, 2] = -2.0, 1.0, 2.0
    v_ref = np.array([0, 0, 2], F)
    mx, my = projected_pixel(v_ref)
    cs = []
    for i in range(12):
        c = Corr(300 + i, 200, 1.0, (0, 0, 1), 2, 0.0)
        c.lv = np.array([0, 0, 1], F)
        c.ln = np.array([0, 0, 1], F)
        c.mv = v_ref.copy()
        c.mn = np.array([0, 0, 1], F)
        c.mpx, c.mpy = mx, my
        cs.append(c)
    return cs, est


def row_vy():
    return (450.0 - CY) * 2.0 / FY
