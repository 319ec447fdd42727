"""Raw Kinect v1 frames for the depth conditioning tests: the synthetic room rendered
at a pose, encoded with the raw curve, with per-pixel code noise, raw-0 holes (single
pixels and blobs), raw-2047 sentinel streaks, a far patch whose codes fall outside the
[0.3, 2.5] m band, and a guidance luma image in [0, 1] (a shaded gradient with a
texture and a step edge, as a sharpened camera image would give).

Everything is seeded, so CPU and GPU tests see the same frames."""
import numpy as np

from tests import room_scene as rs

A, B = -0.0030711016, 3.3309495161


def encode(m):
    """Metres -> nearest raw code (double arithmetic; 0 for no depth)."""
    m = np.asarray(m, np.float64)
    with np.errstate(divide="ignore"):
        code = (1.0 / np.where(m > 0, m, np.inf) - B) / A
    return np.where(m > 0, np.clip(np.rint(code), 1, 2046), 0).astype(np.uint16)


def luma_image(W=640, H=480, seed=0):
    rng = np.random.default_rng(seed + 1000)
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    g = 0.35 + 0.3 * (x / W) + 0.1 * np.sin(y / 23.0) * np.cos(x / 31.0)
    g += np.where(x > 0.55 * W, 0.2, 0.0)  # a step edge
    g += 0.03 * rng.standard_normal((H, W))
    return np.clip(g, 0.0, 1.0).astype(np.float32)


def raw_frame(k=0, W=640, H=480, noise=1.5, seed=0):
    """Frame k of a slow orbit; code noise ~N(0, noise) codes."""
    rng = np.random.default_rng(seed + 17 * k)
    pose = rs.make_pose((0.0, 0.02 * k, 0.0), (0.01 * k, 0.0, 0.0))
    K = rs.scale_intrinsics((525.0, 525.0, 319.5, 239.5), 640, 480, W, H)
    d = rs.render_depth(pose, K, W, H)
    raw = encode(d).astype(np.int32)
    raw = np.where(raw > 0, raw + np.rint(noise * rng.standard_normal((H, W))).astype(np.int32), 0)
    raw = np.clip(raw, 0, 2046)
    # single-pixel dropouts and blob holes
    drop = rng.random((H, W)) < 0.02
    raw[drop] = 0
    for _ in range(12):
        cy, cx, r = rng.integers(0, H), rng.integers(0, W), rng.integers(1, 5)
        yy, xx = np.ogrid[0:H, 0:W]
        raw[(yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] = 0
    # sentinel streaks
    for _ in range(4):
        y0, x0 = rng.integers(0, H - 1), rng.integers(0, W - 40)
        raw[y0, x0:x0 + 40] = 2047
    # a patch beyond the 2.5 m band edge (codes > 954 decode past 2.5 m, and codes
    # past the pole at ~1084.6 decode to negative metres)
    raw[H // 8:H // 8 + 20, W // 8:W // 8 + 30] = rng.integers(960, 1200, (20, 30))
    return raw.astype(np.uint16)
