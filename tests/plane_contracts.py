"""The reference's tilted-plane contracts, restated with their fixtures and bars:

* tests/fusion_bias_contract.cpp:53-160: noise-free depth of the plane z = 1 + 0.2 x
  (640x480, fx = fy = 525, cx = 319.5, cy = 239.5, identity pose), integrated N times
  into a 256^3 / 1 cm volume (trunc 3 cm, max weight 64, origin (-1.28, -1.28, 0), band
  [0.3, 3.0]). A: marching cubes of 10 frames has > 10000 vertices, the mean signed
  vertex distance to the plane is below 0.5 mm and the 95th percentile of |distance|
  below 2 mm. B: after N in {1, 3, 70} frames every observed voxel has weight
  min(N, 64), and the observed count is the full-scan count. C: the volume is
  bit-identical across runs (the reference: across OpenMP thread counts).
* tests/raycast_accuracy_contract.cpp:42-130: the same plane fused 5 times; A: from the
  fusing pose >= 90 % of the pixels hit, every hit within 1 mm of the plane, every
  normal within n . n_plane > 0.99; B: from behind the plane (camera at z = 2 looking
  back) no pixel hits; C: from 45 degrees off axis (eye (-1, 0, 0) looking at (0, 0, 1))
  more than 1000 hits, at most 0.1 % of them more than 2 mm off the plane.
"""
import numpy as np

W, H = 640, 480
K = (525.0, 525.0, 319.5, 239.5)
SLOPE, Z0 = 0.2, 1.0
NORM = np.sqrt(1.0 + SLOPE * SLOPE)
PARAMS = dict(resolution=256, voxel_size=0.01, truncation=0.03, max_weight=64.0, origin=(-1.28, -1.28, 0.0),
              min_depth=0.3, max_depth=3.0)
BAND = (0.3, 3.0)
PLANE_N = np.array([SLOPE, 0.0, -1.0]) / NORM  # toward the fusing camera


def plane_depth():
    """planeDepth (fusion_bias_contract.cpp:61-70): z = 1 / (1 - 0.2 (u - cx) / fx),
    evaluated in double and stored as float."""
    u = np.arange(W, dtype=np.float64)
    a = (u - K[2]) / K[0]
    row = (Z0 / (1.0 - SLOPE * a)).astype(np.float32)
    return np.ascontiguousarray(np.broadcast_to(row, (H, W)))


def plane_distance(p):
    p = np.asarray(p, np.float64)
    return (SLOPE * p[..., 0] - p[..., 2] + Z0) / NORM


def look_from(eye, forward):
    """lookFrom (raycast_accuracy_contract.cpp:70-81), y-down image convention."""
    z = np.asarray(forward, np.float64)
    z = z / np.linalg.norm(z)
    x = np.cross([0.0, 1.0, 0.0], z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, z, eye
    return T.astype(np.float32)


POSE_FRONT = np.eye(4, dtype=np.float32)
POSE_BEHIND = look_from(np.array([0.0, 0.0, 2.0]), np.array([0.0, 0.0, -1.0]))
POSE_OBLIQUE = look_from(np.array([-1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]) - np.array([-1.0, 0.0, 0.0]))


def check_bias(positions):
    assert len(positions) > 10000, len(positions)
    d = plane_distance(positions)
    mean, p95 = float(d.mean()), float(np.sort(np.abs(d))[len(d) * 95 // 100])
    assert abs(mean) < 0.5e-3, mean
    assert p95 < 2.0e-3, p95
    return mean, p95


def check_weights(weight, n, observed_count=None):
    want = min(float(n), PARAMS["max_weight"])
    obs = weight > 0
    assert obs.any()
    assert (weight[obs] == want).all(), (n, np.unique(weight[obs])[:5])
    if observed_count is not None:
        assert observed_count == int(obs.sum())


def cast_stats(v, n):
    hit = np.abs(v).sum(-1) > 0
    d = np.abs(plane_distance(v[hit]))
    dots = n[hit].astype(np.float64) @ PLANE_N
    return dict(hits=int(hit.sum()), off_1mm=int((d > 1e-3).sum()), off_2mm=int((d > 2e-3).sum()),
                bad_normal=int((~(dots > 0.99)).sum()))


def check_front(s):
    assert s["hits"] >= W * H * 90 // 100, s
    assert s["off_1mm"] == 0 and s["bad_normal"] == 0, s


def check_behind(s):
    assert s["hits"] == 0, s


def check_oblique(s):
    assert s["hits"] > 1000 and s["off_2mm"] <= s["hits"] // 1000, s
