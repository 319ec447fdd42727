"""The oracle's depth conditioning stages (oracle/oracle_condition.hpp, restating
src/sensor/SignalConditioner_omp.cpp:442-716) pinned against the reference's own
fixtures and against independent double-precision restatements (tests/cond_pins.py)."""
import numpy as np
import pytest

import oracle
from tests import cond_pins as cp
from tests import cond_scene as cs


# ---- applyDepthEma: tests/depth_ema_determinism_contract.cpp ------------------------
def test_ema_two_frame_contract():
    st = np.zeros((cp.H, cp.W), np.float32)
    f1 = cp.ema_frame(cp.EMA_WALL1, cp.EMA_WALL2)
    f2 = cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE)
    oracle.cond_ema(f1, st)
    out = oracle.cond_ema(f2, st)
    assert out[cp.EMA_CY, cp.EMA_CX + 40] == cp.EMA_EXPECT_WALL
    assert out[cp.EMA_CY, cp.EMA_CX] == cp.EMA_EXPECT_CENTRE != cp.EMA_CONTAMINATED
    # every pixel and the retained history against the contract's double rule
    s64 = np.zeros((cp.H, cp.W))
    _, s64 = cp.ema64(f1, s64)
    want, s64 = cp.ema64(f2, s64)
    assert np.array_equal(out, want)
    assert np.abs(st.astype(np.float64) - s64).max() <= 1e-6
    # no emitted code is a sentinel or out of band
    nz = out[out != 0]
    assert (cp.meters32(nz) > 0).all()


def test_ema_reads_the_source_snapshot():
    """A destination-reading EMA (neighbours already blended) resets the centre and emits
    586: planted in a copy of the rule, it fails the contract's literal."""
    f1 = cp.ema_frame(cp.EMA_WALL1, cp.EMA_WALL2)
    f2 = cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE)
    s = np.zeros((cp.H, cp.W))
    _, s = cp.ema64(f1, s)
    out, _ = cp.ema64(f2, s)
    # the four lower-index neighbours already hold the blended wall value
    m = cp.meters64(f2)
    nb = [cp.meters64(np.array([cp.EMA_WALL2]))[0]] * 4 + [0.7 * m[0, 0] + 0.3 * s[0, 0]] * 4
    centre_m = m[cp.EMA_CY, cp.EMA_CX]
    contaminated_reset = abs(centre_m - np.mean(nb)) > 0.05
    assert contaminated_reset and out[cp.EMA_CY, cp.EMA_CX] == cp.EMA_EXPECT_CENTRE


def test_ema_reset_lifecycle():
    st = np.zeros((cp.H, cp.W), np.float32)
    oracle.cond_ema(cp.ema_frame(cp.EMA_WALL1, cp.EMA_WALL2), st)
    assert (st > 0).any()
    st[:] = 0.0  # resetEMA
    f = cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE)
    assert np.array_equal(oracle.cond_ema(f, st), f), "the frame after a reset is the identity on in-band raws"
    p = f.copy()
    p[cp.EMA_CY, cp.EMA_CX] = 0
    p[cp.EMA_CY, cp.EMA_CX + 5] = 2047
    p[cp.EMA_CY, cp.EMA_CX + 9] = 1500
    out = oracle.cond_ema(p, st)
    for dx, code in ((0, 0), (5, 2047), (9, 1500)):
        assert st[cp.EMA_CY, cp.EMA_CX + dx] == 0.0
        assert out[cp.EMA_CY, cp.EMA_CX + dx] == code, "invalid codes pass through"


def test_ema_against_double_rule_on_scene_sequence():
    st = np.zeros((cp.H, cp.W), np.float32)
    s64 = np.zeros((cp.H, cp.W))
    for k in range(3):
        f = cs.raw_frame(k)
        got = oracle.cond_ema(f, st)
        want, s64 = cp.ema64(f, s64)
        bad = got != want
        # float vs double only near a half-code boundary or a reset threshold
        assert bad.sum() <= 5, bad.sum()


# ---- fillDepthHoles: tests/depth_domain_contract.cpp:425-560 -----------------------
@pytest.mark.parametrize("case", range(6))
def test_hole_fill_contract(case):
    name, d, (y, x), expect = cp.hole_fixtures()[case]
    out = oracle.cond_fill_holes(d)
    assert out[y, x] == expect, name
    moved = out != d
    assert not (moved & (d != 0)).any(), name + ": only raw 0 may change"
    if case == 0:
        moved[y, x] = False
        assert not moved.any(), "a lone hole is the only pixel that changes"


def test_hole_fill_sentinel_block():
    d = cp.hole_sentinel_block()
    cp.check_hole_block(d, oracle.cond_fill_holes(d))


def test_hole_fill_against_double_on_scene():
    for k in range(2):
        d = cs.raw_frame(k)
        got = oracle.cond_fill_holes(d)
        want, mean = cp.fill_holes64(d)
        bad = got != want
        assert (bad <= cp.near_half(mean)).all()
        assert bad.sum() <= 1e-4 * bad.size, bad.sum()
        assert ((d == 0) & (got != 0)).sum() > 1000


# ---- denoiseDepthSpatial ------------------------------------------------------------
def test_median_pins():
    d = np.full((48, 64), 700, np.uint16)
    assert np.array_equal(oracle.cond_median(d), d), "a flat wall is a fixed point"
    d[10, 10] = 760  # a spike: 8 of 9 samples say 700
    assert oracle.cond_median(d)[10, 10] == 700
    # rank count // 2 of the sorted usable samples (centre included): an 8-sample
    # window (one hole) takes the upper median
    e = np.full((48, 64), 700, np.uint16)
    e[20, 19:22] = (600, 800, 650)
    e[19, 19:22] = (610, 0, 660)
    e[21, 19:22] = (620, 630, 640)
    win = np.sort(cp.meters32(e[19:22, 19:22]).reshape(-1))
    win = win[win > 0]
    assert len(win) == 8
    want = cp.encode64(np.float64(win[4]))
    assert oracle.cond_median(e)[20, 20] == want == 640
    # fewer than 3 usable samples: unchanged; a hole stays a hole
    f = np.zeros((48, 64), np.uint16)
    f[30, 30], f[30, 31] = 700, 720
    out = oracle.cond_median(f)
    assert out[30, 30] == 700 and out[30, 31] == 720 and out[29, 30] == 0
    f[31, 30] = 710  # now 3 usable: rank 1 of (700, 710, 720)
    assert oracle.cond_median(f)[30, 30] == 710


def test_median_reflect_border():
    """cpuReflectCoord mirrors including the edge sample: at the corner the 3x3 window
    holds the corner sample four times."""
    d = np.full((8, 8), 700, np.uint16)
    d[0, 0] = 600
    d[0, 1] = 600
    # window at (0,0): rows (0,0,1) x cols (0,0,1): 600 x4 (corner), 600 x2 ((0,1) twice), 700 x3
    assert oracle.cond_median(d)[0, 0] == 600
    d[0, 1] = 700  # now 600 x4, 700 x5 -> rank 4 is 700
    assert oracle.cond_median(d)[0, 0] == 700


def test_median_against_sort_on_scene():
    for k in range(2):
        d = cs.raw_frame(k)
        want, _ = cp.median64(d)
        assert np.array_equal(oracle.cond_median(d), want)


# ---- guidedDepthFilter -------------------------------------------------------------
def test_guided_flat_is_identity():
    d = np.full((40, 48), 700, np.uint16)
    lu = cs.luma_image(48, 40)
    assert np.array_equal(oracle.cond_guided(d, lu), d)


def test_guided_against_double_on_scene():
    lu = cs.luma_image()
    for k, luma in ((0, None), (1, lu)):
        d = oracle.cond_fill_holes(oracle.cond_median(cs.raw_frame(k)))
        got = oracle.cond_guided(d, luma)
        want, mean = cp.guided64(d, luma)
        bad = got != want
        # float sums of 81 taps vs double: only codes whose double value is within
        # 2e-3 of a half-code boundary may round the other way
        assert (bad <= cp.near_half(mean)).all(), bad.sum()
        assert bad.sum() <= 1e-4 * bad.size, bad.sum()
        assert (got != d).sum() > 50000, "the filter moves most codes of a noisy frame"


def test_guided_luma_edge_stops_smoothing():
    """Two depth planes meeting at a vertical luma edge: with the guidance, pixels next
    to the edge keep their side's depth; without it the depth term alone decides."""
    d = np.full((32, 64), 600, np.uint16)
    d[:, 32:] = 606  # ~1.3 cm apart at ~0.66 m
    lu = np.zeros((32, 64), np.float32)
    lu[:, 32:] = 1.0
    with_l = oracle.cond_guided(d, lu)
    without = oracle.cond_guided(d)
    assert with_l[16, 31] == 600 and with_l[16, 32] == 606
    assert without[16, 31] != 600 and without[16, 32] != 606
    want, _ = cp.guided64(d, lu)
    assert np.array_equal(with_l, want)


def test_guided_expf_choice_is_invisible_on_scene_frames():
    """The reference's weights come from the C library's expf (not correctly rounded for
    ~6e-4 of arguments); the oracle's mode 1 rounds the double exponential. On the scene
    frames the codes agree at every pixel."""
    lu = cs.luma_image()
    d = oracle.cond_fill_holes(oracle.cond_median(cs.raw_frame(2)))
    a = oracle.cond_guided(d, lu, exp_mode=0)
    b = oracle.cond_guided(d, lu, exp_mode=1)
    assert (a != b).sum() <= 3


# ---- processCpu order ----------------------------------------------------------------
def test_process_order_and_minimal():
    lu = cs.luma_image()
    d = cs.raw_frame(0)
    st = np.zeros(d.shape, np.float32)
    st2 = np.zeros(d.shape, np.float32)
    full = oracle.cond_process(d, st, lu)
    step = oracle.cond_ema(oracle.cond_guided(oracle.cond_fill_holes(oracle.cond_median(d)), lu), st2)
    assert np.array_equal(full, step) and np.array_equal(st, st2)
    st3 = np.zeros(d.shape, np.float32)
    assert np.array_equal(oracle.cond_process(d, st3, lu, stages=oracle.COND_MEDIAN), oracle.cond_median(d))
    assert not st3.any()
