"""The far-sample bound behind the raycast's empty-space skipping (CPU, float32 numpy).

A brick stays clear while each of its voxels is unobserved or has tsdf in
[kFarTsdf, 1.5]. A trilinear sample whose eight corners all lie in [kFarTsdf, 1.5] must
then interpolate to a finite value >= 0.999f under both of the raycast's evaluation
orders (the exact one, c0 (1 - t) + c1 t, and the fast one, fma(c1 - c0, t, c0)), so the
march takes its far step there. The threshold is read from the kernel header.
"""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def far_threshold():
    src = open(os.path.join(ROOT, "azu_amd", "csrc", "common.hpp")).read()
    m = re.search(r"constexpr float kFarTsdf = ([0-9.eE+-]+)f;", src)
    assert m, "kFarTsdf not found"
    return np.float32(float(m.group(1)))


def lerp_exact(a, b, t):
    one = np.float32(1.0)
    return a * (one - t) + b * t


def lerp_fast(a, b, t):
    d = (b - a).astype(np.float64)
    return (d * t.astype(np.float64) + a.astype(np.float64)).astype(np.float32)


def trilinear(c, tx, ty, tz, lerp):
    v00 = lerp(c[0], c[1], tx)
    v10 = lerp(c[2], c[3], tx)
    v01 = lerp(c[4], c[5], tx)
    v11 = lerp(c[6], c[7], tx)
    v0 = lerp(v00, v10, ty)
    v1 = lerp(v01, v11, ty)
    return lerp(v0, v1, tz)


def test_threshold_constants():
    k = far_threshold()
    assert k > np.float32(0.999)
    # the FP16 values just below 1 that count as far: 1 - 2^-11 and 1 - 2^-10
    assert np.float32(np.float16(1.0 - 2.0 ** -10)) >= k
    assert np.float32(np.nextafter(np.float16(1.0 - 2.0 ** -10), np.float16(0))) < k


def test_far_corners_interpolate_far():
    k = far_threshold()
    rng = np.random.default_rng(7)
    n = 1 << 21
    # corners: the threshold itself (the worst case), 1.0, 1.5 and uniform draws
    pick = rng.integers(0, 4, size=(8, n))
    uni = (k + (np.float32(1.5) - k) * rng.random((8, n), dtype=np.float32)).astype(np.float32)
    c = np.where(pick == 0, k, np.where(pick == 1, np.float32(1.0), np.where(pick == 2, np.float32(1.5), uni)))
    c = np.clip(c, k, np.float32(1.5)).astype(np.float32)
    c[:, : n // 8] = k  # all corners at the threshold

    def frac(size):
        t = rng.random(size, dtype=np.float32)
        special = np.array([0.0, 2.0 ** -24, 0.25, 0.5, 1.0 - 2.0 ** -24], np.float32)
        m = rng.random(size) < 0.2
        t[m] = special[rng.integers(0, special.size, size=int(m.sum()))]
        return t

    tx, ty, tz = frac(n), frac(n), frac(n)
    lim = np.float32(0.999)
    for lerp in (lerp_exact, lerp_fast):
        f = trilinear(c, tx, ty, tz, lerp)
        assert np.isfinite(f).all()
        assert f.min() >= lim, (lerp.__name__, float(f.min()))
        assert f.max() <= np.float32(1.5)


def test_fp16_packed_far_test_matches_rounding():
    """integrate decides far-ness of a packed FP16 word from its float tsdf before the
    binary16 rounding: the thresholds must give exactly the words whose half value lies
    in [kFarTsdf, 1.5]."""
    k = far_threshold()
    src = open(os.path.join(ROOT, "azu_amd", "csrc", "common.hpp")).read()
    lo_hi = re.findall(r"farPacked\(float t\) \{ return t >= ([0-9.]+)f && t <= ([0-9.]+)f; \}", src)
    assert len(lo_hi) == 1, lo_hi
    lo, hi = np.float32(float(lo_hi[0][0])), np.float32(float(lo_hi[0][1]))
    # every float32 in a band around both thresholds, plus a random spread
    def around(x, n):
        b = np.array([x], np.float32).view(np.int32)[0]
        return np.arange(b - n, b + n, dtype=np.int32).view(np.float32)
    rng = np.random.default_rng(3)
    t = np.concatenate([around(lo, 1 << 20), around(hi, 1 << 20),
                        rng.uniform(0.9, 1.6, 1 << 20).astype(np.float32)])
    h = t.astype(np.float16).astype(np.float32)  # round to nearest even
    want = (h >= k) & (h <= np.float32(1.5))
    got = (t >= lo) & (t <= hi)
    assert np.array_equal(want, got)
