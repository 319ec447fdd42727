"""Closed forms for the live frame geometry and the constant-velocity estimate, shared
by the CPU oracle tests and the GPU tests.

* Back-projection (src/sensor/FrameData.cpp:113-133): a depth map rendered from a
  tilted plane Z = z0 + a X + b Y gives vertices (u d, v d, d) with u = (x - cx) / fx,
  v = (y - cy) / fy, evaluated here in double from the float depth.
* Normals (FrameData.cpp:196-251): (v[r] - v[l]) x (v[d] - v[u]), normalised, is the
  plane's unit normal (-a, -b, 1) / |(-a, -b, 1)| with that sign (pointing away from the
  camera: for a fronto-parallel plane the right and down differences are +x and +y);
  border pixels and pixels next to a hole or a depth jump larger than
  max(0.03, 0.05 d) are zero.
* Constant-velocity estimate (src/app/PipelineController.cpp:1162-1187): for poses
  P_k = L^k P0 (a constant world-frame motion L) or P_k = P0 M^k (a constant body-frame
  motion M), cur (last^-1 cur) is exactly the next pose; with the composition order
  reversed (last cur^-1, or a left product) it is not.
"""
import numpy as np

FX, FY, CX, CY = 525.0, 525.0, 319.5, 239.5
K640 = (FX, FY, CX, CY)
W, H = 640, 480

PLANE = (1.5, 0.3, -0.2)  # z0, a, b


def plane_depth(z0=PLANE[0], a=PLANE[1], b=PLANE[2], W=W, H=H):
    """Float32 depth of the plane Z = z0 + a X + b Y along each pixel's ray."""
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    u = (x - CX) / FX
    v = (y - CY) / FY
    return (z0 / (1.0 - a * u - b * v)).astype(np.float32)


def exact_vertices(depth):
    """(u d, v d, d) in double for the float depth d (0 where d <= 0)."""
    H_, W_ = depth.shape
    y, x = np.mgrid[0:H_, 0:W_].astype(np.float64)
    d = depth.astype(np.float64)
    v = np.stack([(x - CX) / FX * d, (y - CY) / FY * d, d], axis=-1)
    v[depth <= 0] = 0.0
    return v


def plane_normal(a=PLANE[1], b=PLANE[2]):
    n = np.array([-a, -b, 1.0])
    return n / np.linalg.norm(n)


def check_plane(vertices, normals, depth, a=PLANE[1], b=PLANE[2]):
    """Vertices within 4 float ulps of the exact back-projection, interior normals
    within 2e-4 of the plane normal (rounding of the 2-pixel chords), borders zero."""
    ve = exact_vertices(depth)
    err = np.abs(vertices.astype(np.float64) - ve)
    assert err.max() <= 1e-6 * max(1.0, float(depth.max())), err.max()
    n = normals.astype(np.float64)
    inner = n[1:-1, 1:-1]
    dn = np.abs(inner - plane_normal(a, b)).max()
    assert dn <= 2e-4, dn
    for edge in (n[0], n[-1], n[:, 0], n[:, -1]):
        assert not edge.any()


def step_depth(d0, step, x0=320, W=W, H=H):
    """A fronto-parallel plane at d0 with d0 + step for x >= x0."""
    d = np.full((H, W), d0, np.float32)
    d[:, x0:] = np.float32(d0 + step)
    return d


# (d0, step, zeroed): the jump threshold max(0.03, 0.05 d) on both sides of each term.
# At d = 1 the 5 % term (0.05) decides: 0.04 is kept, 0.06 cut; at d = 0.4 the 0.03
# floor decides: 0.025 is kept (a 5 %-only threshold, 0.02, would cut it), 0.035 cut.
STEP_CASES = [(1.0, 0.04, False), (1.0, 0.06, True), (0.4, 0.025, False), (0.4, 0.035, True)]


def check_step(normals, zeroed, x0=320):
    n = normals
    inner = np.abs(n[1:-1, 1:-1]).sum(axis=-1) > 0
    cols = np.arange(1, n.shape[1] - 1)
    near = (cols == x0 - 1) | (cols == x0)
    if zeroed:
        assert not inner[:, near].any(), "normals next to the jump must be zero"
        assert inner[:, ~near].all()
    else:
        assert inner.all(), "a step below the threshold keeps every interior normal"


def hole_depth():
    d = plane_depth()
    d[100, 100] = 0.0
    return d


def check_hole(vertices, normals):
    assert not vertices[100, 100].any()
    for (y, x) in ((100, 100), (100, 99), (100, 101), (99, 100), (101, 100)):
        assert not normals[y, x].any()
    assert normals[100, 98].any() and normals[98, 100].any()


def rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    Kx = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * Kx + (1 - np.cos(ang)) * Kx @ Kx


def rigid(R, t):
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def uniform_motion(k, kind):
    """P_k = L^k P0 (kind "world": an orbit step L about a centre) or P0 M^k (kind "body":
    a body-frame step M). Either way last^-1 cur is the same relative motion every
    frame and cur (last^-1 cur) = P_{k+1}."""
    P0 = rigid(rot((0.2, 1.0, -0.3), 0.7), (0.1, -0.2, 0.9))
    if kind == "world":
        c = np.array([0.0, 0.0, 0.8])
        Lr = rot((0, 1, 0), np.deg2rad(2.0))
        return np.linalg.matrix_power(rigid(Lr, c - Lr @ c), k) @ P0
    M = rigid(rot((1.0, 0.2, 0.1), 0.01), (0.003, -0.001, 0.002))
    return P0 @ np.linalg.matrix_power(M, k)
