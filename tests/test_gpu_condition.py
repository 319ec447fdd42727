"""Depth conditioning on the GPU (azu_cond_*, the depth half of SignalConditioner,
src/sensor/SignalConditioner_omp.cpp:337-716) against the oracle, bit for bit:

  * each stage on the scene frames (raw codes with noise, holes, sentinel streaks, an
    out-of-band patch), the guided filter with and without guidance luma;
  * the full chain over a 5-frame sequence, codes and the EMA history's float bits;
  * the reference's fixtures (EMA 534 / 575, hole fill 700 / 500 / 528, lifecycle);
  * odd frame sizes (tiles cut by the border), minimal mode, aliased in/out.

The guided filter's Gaussians are the correctly rounded expf on the GPU; it is compared
with the oracle's mode 1 (the same rounding), and with the reference's C-library expf
(mode 0) to the measured bound."""
import numpy as np
import pytest

import oracle
from azu_amd import conditioner as C
from azu_amd.native import AzuError
from tests import cond_pins as cp
from tests import cond_scene as cs
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

STAGES = [(C.MEDIAN, "median"), (C.HOLES, "holes"), (C.GUIDED, "guided"), (C.EMA, "ema")]


def _oracle_stage(stage, d, luma, state, exp_mode=1):
    if stage == C.MEDIAN:
        return oracle.cond_median(d)
    if stage == C.HOLES:
        return oracle.cond_fill_holes(d)
    if stage == C.GUIDED:
        return oracle.cond_guided(d, luma, exp_mode=exp_mode)
    return oracle.cond_ema(d, state)


@pytest.mark.parametrize("stage,name", STAGES)
def test_stage_bitwise(stage, name):
    sc = C.SignalConditioner()
    lu = cs.luma_image()
    state = np.zeros((480, 640), np.float32)
    for k in range(2):
        d = cs.raw_frame(k)
        for luma in ((None, lu) if stage == C.GUIDED else (None,)):
            say("cond stage %s frame %d luma %s" % (name, k, luma is not None))
            got = sc.run_stages(d, stage, luma)
            want = _oracle_stage(stage, d, luma, state)
            assert np.array_equal(got, want), (name, k, int((got != want).sum()))
    if stage == C.EMA:
        assert np.array_equal(sc.emaState().view(np.uint32), state.view(np.uint32))
    sc.close()


def test_chain_sequence_bitwise():
    sc = C.SignalConditioner()
    lu = cs.luma_image()
    state = np.zeros((480, 640), np.float32)
    for k in range(5):
        d = cs.raw_frame(k)
        say("cond process frame %d" % k)
        got = sc.process(d, lu)
        want = oracle.cond_process(d, state, lu, exp_mode=1)
        assert np.array_equal(got, want), (k, int((got != want).sum()))
        assert np.array_equal(sc.emaState().view(np.uint32), state.view(np.uint32)), k
        assert (got != d).sum() > 100000
    sc.close()


def test_guided_vs_library_expf():
    """Against the reference's own weights (the C library's expf): the codes differ only
    where that expf's rounding moves a weighted mean across a half-code boundary."""
    sc = C.SignalConditioner()
    lu = cs.luma_image()
    bad = 0
    for k in range(3):
        d = oracle.cond_fill_holes(oracle.cond_median(cs.raw_frame(k)))
        say("guided vs library expf frame %d" % k)
        got = sc.run_stages(d, C.GUIDED, lu)
        want = oracle.cond_guided(d, lu, exp_mode=0)
        diff = got != want
        assert (np.abs(got.astype(int) - want.astype(int))[diff] <= 1).all()
        bad += int(diff.sum())
    assert bad <= 1e-5 * 3 * d.size, bad
    sc.close()


def test_reference_fixtures():
    sc = C.SignalConditioner()
    say("EMA two-frame fixture")
    sc.applyDepthEmaForTests(cp.ema_frame(cp.EMA_WALL1, cp.EMA_WALL2))
    out = sc.applyDepthEmaForTests(cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE))
    assert out[cp.EMA_CY, cp.EMA_CX + 40] == cp.EMA_EXPECT_WALL
    assert out[cp.EMA_CY, cp.EMA_CX] == cp.EMA_EXPECT_CENTRE
    s64 = np.zeros((480, 640))
    _, s64 = cp.ema64(cp.ema_frame(cp.EMA_WALL1, cp.EMA_WALL2), s64)
    want, s64 = cp.ema64(cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE), s64)
    assert np.array_equal(out, want)
    assert np.abs(sc.emaState().astype(np.float64) - s64).max() <= 1e-6
    # lifecycle: reset -> identity; invalid codes clear their own history only
    sc.resetEMA()
    assert not sc.emaState().view(np.uint32).any()
    f = cp.ema_frame(cp.EMA_WALL2, cp.EMA_CENTRE)
    assert np.array_equal(sc.applyDepthEmaForTests(f), f)
    p = f.copy()
    p[cp.EMA_CY, cp.EMA_CX], p[cp.EMA_CY, cp.EMA_CX + 5], p[cp.EMA_CY, cp.EMA_CX + 9] = 0, 2047, 1500
    out = sc.applyDepthEmaForTests(p)
    st = sc.emaState()
    for dx, code in ((0, 0), (5, 2047), (9, 1500)):
        assert st[cp.EMA_CY, cp.EMA_CX + dx] == 0.0 and out[cp.EMA_CY, cp.EMA_CX + dx] == code
    assert st[cp.EMA_CY, cp.EMA_CX + 20] > 0
    say("hole-fill fixtures")
    for name, d, (y, x), expect in cp.hole_fixtures():
        out = sc.fillDepthHolesForTests(d)
        assert out[y, x] == expect, name
        assert not ((out != d) & (d != 0)).any(), name
    d = cp.hole_sentinel_block()
    cp.check_hole_block(d, sc.fillDepthHolesForTests(d))
    sc.close()


@pytest.mark.parametrize("W,H", [(37, 23), (641, 9), (4, 4), (160, 120)])
def test_odd_sizes(W, H):
    rng = np.random.default_rng(W * 1000 + H)
    d = rng.integers(500, 620, (H, W)).astype(np.uint16)
    d[rng.random((H, W)) < 0.1] = 0
    d[rng.random((H, W)) < 0.03] = 2047
    lu = rng.random((H, W)).astype(np.float32)
    sc = C.SignalConditioner(W, H)
    state = np.zeros((H, W), np.float32)
    for k in range(2):
        say("cond %dx%d frame %d" % (W, H, k))
        got = sc.process(d, lu)
        want = oracle.cond_process(d, state, lu, exp_mode=1)
        assert np.array_equal(got, want), (W, H, k)
        d = np.clip(d.astype(int) + rng.integers(-3, 4, (H, W)), 0, 2047).astype(np.uint16)
    sc.close()


def test_minimal_alias_and_errors():
    sc = C.SignalConditioner()
    d = cs.raw_frame(3)
    sc.setMinimalDepth(True)
    say("minimal mode")
    assert np.array_equal(sc.process(d), oracle.cond_median(d))
    assert not sc.emaState().any(), "minimal mode leaves the history alone"
    sc.setMinimalDepth(False)
    # in place (d_out == d_in)
    buf = sc._in
    buf.upload(d)
    sc.process_device(buf.ptr, buf.ptr)
    state = np.zeros((480, 640), np.float32)
    assert np.array_equal(buf.download(np.uint16, d.shape), oracle.cond_process(d, state, exp_mode=1))
    with pytest.raises(AzuError):
        sc.run_stages_device(buf.ptr, sc._out.ptr, 16)
    sc.close()
    with pytest.raises(AzuError):
        C.SignalConditioner(3, 480)
