"""Pins for the depth conditioning stages (src/sensor/SignalConditioner_omp.cpp:442-716),
shared by the CPU oracle tests and the GPU tests.

* The reference's own fixtures with their hand-derived numbers:
  tests/depth_ema_determinism_contract.cpp (two-frame EMA: wall raw 534, centre raw 575
  and not the contaminated 586; reset lifecycle) and the hole-fill section of
  tests/depth_domain_contract.cpp:425-560 (uniform wall -> 700, raw 2047 and an
  out-of-band raw untouched, a sentinel-only window stays a hole, a mixed window -> 500,
  a 500/548 checkerboard -> 528, only raw 0 ever changes and only to in-band codes).
* Independent restatements in double precision, vectorised with numpy (borders by
  np.pad(mode="symmetric"), which is cpuReflectCoord: -1 -> 0, n -> n - 1). The float
  stages agree with them except where a double value lies within float rounding of a
  half-code boundary; `near_half` finds those pixels.
"""
import numpy as np

A64, B64 = -0.0030711016, 3.3309495161
A32, B32 = np.float32(-0.0030711016), np.float32(3.3309495161)
W, H = 640, 480
LO, HI = 0.3, 2.5


# ---- the raw curve, in double -------------------------------------------------------
def meters64(raw, lo=LO, hi=HI):
    """cpuDepthMeters in double (0 = unusable)."""
    r = np.asarray(raw).astype(np.float64)
    valid = (r != 0) & (r < 2047)
    with np.errstate(divide="ignore"):
        m = 1.0 / (r * A64 + B64)
    return np.where(valid & np.isfinite(m) & (m >= lo) & (m <= hi), m, 0.0)


def meters32(raw, lo=LO, hi=HI):
    """cpuDepthMeters in float32, as the stages use it (0 = unusable)."""
    r = np.asarray(raw).astype(np.float32)
    valid = (r != 0) & (r < 2047)
    with np.errstate(divide="ignore"):
        m = np.float32(1.0) / (r * A32 + B32)
    return np.where(valid & np.isfinite(m) & (m >= np.float32(lo)) & (m <= np.float32(hi)), m, np.float32(0))


def raw_value64(m):
    """(1/m - B) / A in double: the pre-rounding code of cpuDepthMetersToRaw."""
    with np.errstate(divide="ignore"):
        return (1.0 / m - B64) / A64


def encode64(m, lo=LO, hi=HI):
    m = np.asarray(m, np.float64)
    ok = np.isfinite(m) & (m >= lo) & (m <= hi)
    raw = np.where(ok, raw_value64(np.where(ok, m, 1.0)), 0.0)
    c = np.floor(raw + 0.5)
    return np.where(ok & (c >= 1) & (c <= 2046), c, 0).astype(np.uint16)


def near_half(m, tol=2e-3):
    """Pixels whose double pre-rounding code is within tol of a half integer."""
    raw = raw_value64(np.where(m > 0, m, 1.0))
    return (m > 0) & (np.abs(raw - np.floor(raw) - 0.5) < tol)


def _windows(a, r):
    """(2r+1)^2 shifted views of a, reflected at the borders, in row-major tap order."""
    p = np.pad(a, r, mode="symmetric")
    Hh, Ww = a.shape
    return [(dy, dx, p[r + dy:r + dy + Hh, r + dx:r + dx + Ww]) for dy in range(-r, r + 1) for dx in range(-r, r + 1)]


# ---- independent double restatements ------------------------------------------------
def median64(raw, lo=LO, hi=HI):
    """denoiseDepthSpatial: rank count // 2 of the usable 3x3 samples (centre included),
    for a usable centre with count >= 3."""
    m = meters32(raw, lo, hi)  # the window holds the float metres; the median is one of them
    stack = np.stack([w for _, _, w in _windows(m, 1)], axis=0)
    cnt = (stack > 0).sum(axis=0)
    srt = np.sort(np.where(stack > 0, stack, np.inf), axis=0)
    med = np.take_along_axis(srt, (cnt // 2)[None], axis=0)[0]
    out = np.asarray(raw, np.uint16).copy()
    sel = (m > 0) & (cnt >= 3)
    out[sel] = encode64(med[sel].astype(np.float64), lo, hi)
    return out, np.where(sel, med, 0.0)


def fill_holes64(raw, lo=LO, hi=HI):
    m = meters64(raw, lo, hi)
    acc = np.zeros_like(m)
    wsum = np.zeros_like(m)
    for dy, dx, w in _windows(m, 2):
        if dy == 0 and dx == 0:
            continue
        wt = 1.0 / (dy * dy + dx * dx + 1)
        acc += np.where(w > 0, wt * w, 0.0)
        wsum += np.where(w > 0, wt, 0.0)
    raw = np.asarray(raw, np.uint16)
    hole = (raw == 0) & (wsum > 1e-6)
    mean = np.where(hole, acc / np.where(wsum > 0, wsum, 1.0), 0.0)
    out = raw.copy()
    out[hole] = encode64(mean[hole], lo, hi)
    return out, mean


def guided64(raw, luma=None, lo=LO, hi=HI):
    """guidedDepthFilter in double. The edge factor multiplies every tap of a pixel's
    window alike, so it cancels from the weighted mean; it is carried anyway."""
    raw = np.asarray(raw, np.uint16)
    m = meters64(raw, lo, hi)
    lu = np.zeros(raw.shape) if luma is None else np.asarray(luma, np.float64)
    # computeDepthGradient on the raw codes (1 at the border or next to an unusable code)
    r = raw.astype(np.float64)
    edge = np.ones(raw.shape)
    inner = np.zeros(raw.shape, bool)
    inner[1:-1, 1:-1] = True
    u = m > 0
    ok = inner.copy()
    ok[1:-1, 1:-1] &= u[1:-1, 1:-1] & u[1:-1, 2:] & u[1:-1, :-2] & u[2:, 1:-1] & u[:-2, 1:-1]
    g = np.zeros(raw.shape)
    g[1:-1, 1:-1] = np.hypot(r[1:-1, 2:] - r[1:-1, :-2], r[2:, 1:-1] - r[:-2, 1:-1])
    edge = np.where(ok, np.minimum(g / 100.0, 1.0), 1.0)
    ef = 1.0 - edge * 0.5
    lwin = dict(((dy, dx), w) for dy, dx, w in _windows(lu, 4))
    sw = np.zeros(raw.shape)
    sd = np.zeros(raw.shape)
    for dy, dx, nd in _windows(m, 4):
        ws = np.exp(-(dy * dy + dx * dx) / 32.0)
        wl = np.exp(-((lwin[(dy, dx)] - lu) ** 2) / (2 * 0.1 * 0.1 + 0.01))
        wd = np.exp(-((nd - m) ** 2) / (2 * 0.04 * 0.04 + 0.01))
        w = np.where(nd > 0, ws * wl * wd * ef, 0.0)
        sw += w
        sd += w * nd
    out = raw.copy()
    sel = u & (sw > 1e-6)
    mean = np.where(sel, sd / np.where(sw > 0, sw, 1.0), 0.0)
    out[sel] = encode64(mean[sel], lo, hi)
    return out, mean


# ---- the reference's fixtures ---------------------------------------------------------
EMA_WALL1, EMA_WALL2, EMA_CENTRE = 500, 548, 586
EMA_CX, EMA_CY = 320, 240
EMA_EXPECT_WALL, EMA_EXPECT_CENTRE, EMA_CONTAMINATED = 534, 575, 586


def ema_frame(wall, centre):
    d = np.full((H, W), wall, np.uint16)
    d[EMA_CY, EMA_CX] = centre
    return d


def ema64(raw, state, lo=LO, hi=HI):
    """applyDepthEma in double (the contract's own rule, oracleEma,
    depth_ema_determinism_contract.cpp:113-150), neighbours from the source frame."""
    raw = np.asarray(raw, np.uint16)
    m = meters64(raw, lo, hi)
    prev = np.asarray(state, np.float64)
    delta = np.abs(m - prev)
    reset = ~(prev > 0) | (delta > 0.05)
    s = np.zeros_like(m)
    n = np.zeros_like(m)
    for dy, dx, w in _windows(m, 1):
        if dy == 0 and dx == 0:
            continue
        s += w
        n += w > 0
    check = ~reset & (delta > 0.025) & (n > 0)
    reset |= check & (np.abs(m - s / np.where(n > 0, n, 1)) > 0.05)
    filt = np.where(reset, m, 0.7 * m + 0.3 * prev)
    code = encode64(filt, lo, hi)
    usable = m > 0
    out = np.where(usable, code, raw).astype(np.uint16)
    new_state = np.where(usable & (code != 0), filt, 0.0)
    return out, new_state


def hole_fixtures():
    """(name, frame, (y, x), expected code at (y, x) or None) of depth_domain_contract's
    hole-fill section (tests/depth_domain_contract.cpp:425-525)."""
    out = []
    d = np.full((H, W), 700, np.uint16)
    d[100, 100] = 0
    out.append(("uniform wall", d, (100, 100), 700))
    d = np.full((H, W), 700, np.uint16)
    d[100, 200] = 2047
    out.append(("sentinel untouched", d, (100, 200), 2047))
    d = np.full((H, W), 700, np.uint16)
    d[100, 300] = 1500
    out.append(("out-of-band raw untouched", d, (100, 300), 1500))
    d = np.full((H, W), 700, np.uint16)
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            d[100 + dy, 400 + dx] = 2047 if (abs(dx) + abs(dy)) % 2 else 1500
    d[100, 400] = 0
    out.append(("sentinel-only window", d, (100, 400), 0))
    d = np.full((H, W), 700, np.uint16)
    pattern = (700, 500, 2047, 0, 1500, 548)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            d[200 + dy, 500 + dx] = pattern[(abs(dx) + abs(dy)) % 6]
    d[200, 500] = 0
    out.append(("mixed window", d, (200, 500), 500))
    d = np.full((H, W), 700, np.uint16)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            d[300 + dy, 100 + dx] = 548 if (abs(dx) + abs(dy)) % 2 else 500
    d[300, 100] = 0
    out.append(("checkerboard", d, (300, 100), 528))
    return out


def hole_sentinel_block():
    """depth_domain_contract.cpp:527-555: a 10 x 10 block of raw 0 / 2047 in a 700 wall."""
    d = np.full((H, W), 700, np.uint16)
    for y in range(50, 60):
        for x in range(50, 60):
            d[y, x] = 0 if (x * 7 + y * 3) % 5 == 0 else 2047
    return d


def check_hole_block(before, after, lo=LO, hi=HI):
    moved = after != before
    assert not (moved & (before != 0)).any(), "only raw 0 may change"
    assert (meters32(after[moved], lo, hi) > 0).all(), "a fill writes an in-band code"
    planted = (before != 0) & ((before >= 2047) | (before == 0))
    assert (after[planted] == before[planted]).all(), "planted sentinels stay"
    assert moved.any(), "holes next to the wall are filled"
