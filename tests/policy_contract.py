"""Checks of the reference's ICP numeric-policy contract, shared by the CPU oracle and
the GPU tracker (tests/icp_numeric_policy_contract.cpp:241-849 of iphands/azu).

Every check takes `run(cs, dist=0.1, angle=30.0, est=None)`, which tracks the
correspondence set `cs` with one level-0 iteration ({1, 0, 0}) and returns the result
as a dict (pose 4x4 row-major, error, inliers, converged, tracking_ok, final_step,
valid_live_points, valid_model_points, projected_points, dist_filtered,
angle_filtered). The expected numbers are the contract's closed forms in double.
"""
import math

import numpy as np

from tests import policy_fixtures as pf

I4 = np.eye(4, dtype=np.float32)


def counters(r, live, model, projected, dist_f, angle_f):
    got = (r["valid_live_points"], r["valid_model_points"], r["projected_points"], r["dist_filtered"],
           r["angle_filtered"])
    assert got == (live, model, projected, dist_f, angle_f), got


def same_run(a, b):
    """checkDeterminism (:202-204): a repeated run is bit-identical."""
    assert np.array_equal(a["pose"].view(np.uint32), b["pose"].view(np.uint32))
    for k in ("error", "inliers", "converged", "tracking_ok", "final_step", "valid_live_points",
              "valid_model_points", "projected_points", "dist_filtered", "angle_filtered"):
        assert a[k] == b[k] or (k in ("error", "final_step") and math.isnan(a[k]) and math.isnan(b[k])), k


def twice(run, cs, **kw):
    a = run(cs, **kw)
    same_run(a, run(cs, **kw))
    return a


def pose_is(r, T):
    return np.array_equal(np.asarray(r["pose"], np.float32).view(np.uint32), np.asarray(T, np.float32).view(np.uint32))


def acceptable_final_step(run):
    """testAcceptableFinalStepIsSuccess (:240-267)."""
    cs = pf.grid(300, 339, 237, 242, 1.0, 5e-4)
    assert len(cs) == 240
    r = twice(run, cs)
    counters(r, 240, 240, 240, 0, 0)
    assert r["inliers"] == 240 and not r["converged"]
    assert pf.CONV_STEP < r["final_step"] <= pf.ACCEPT_STEP
    assert r["tracking_ok"]
    oracle = -240.0 * 5e-4 / (240.0 + pf.DAMP_ESC)
    assert abs(float(r["pose"][2, 3]) - oracle) <= pf.TOL_MATCH
    assert abs(float(r["pose"][0, 3])) <= pf.TOL_MATCH and abs(float(r["pose"][1, 3])) <= pf.TOL_MATCH
    assert abs(float(r["error"]) - 5e-4 * 5e-4) <= pf.TOL_MATCH


def inlier_floor(run):
    """testInlierFloorStillTrackingLost (:269-301): 100 inliers fail, 101 succeed."""
    cs = pf.grid(310, 329, 237, 241, 1.0, 5e-4)
    assert len(cs) == 100
    r = twice(run, cs)
    counters(r, 100, 100, 100, 0, 0)
    assert r["inliers"] == 100 and not r["converged"] and r["final_step"] <= pf.ACCEPT_STEP
    assert not r["tracking_ok"]
    assert abs(float(r["pose"][2, 3]) - (-100.0 * 5e-4 / (100.0 + pf.DAMP_ESC))) <= pf.TOL_MATCH
    plus = cs + [pf.Corr(200, 200, 1.0, (0, 0, 1), 2, 5e-4)]
    r101 = twice(run, plus)
    assert r101["inliers"] == 101 and r101["final_step"] <= pf.ACCEPT_STEP and r101["tracking_ok"]


def rotation_cap(run):
    """testRotationCapRejectsUpdate (:331-387)."""
    big = pf.rotation_leak(100)
    assert len(big) == 200
    r = twice(run, big, dist=3.0)
    counters(r, 200, 200, 200, 0, 0)
    assert r["inliers"] == 0 and not r["converged"] and not r["tracking_ok"]
    assert math.isinf(r["final_step"]) and r["final_step"] > 0
    assert pose_is(r, I4), "the over-cap update is rejected, not clamped"
    vy = pf.row_vy()
    raw_rot = 200.0 * pf.HUBER_K * vy / (200.0 * pf.HUBER_K * vy * vy + pf.DAMP_ESC)
    assert raw_rot > pf.ROT_CAP + 0.3
    s = twice(run, pf.rotation_leak(10), dist=3.0)
    assert s["inliers"] == 20
    oracle_rot = 20.0 * pf.HUBER_K * vy / (20.0 * pf.HUBER_K * vy * vy + pf.DAMP_ESC)
    assert abs(float(s["final_step"]) - oracle_rot) <= pf.TOL_MATCH
    assert max(abs(float(s["pose"][i, 3])) for i in range(3)) <= pf.TOL_MATCH
    assert abs(float(s["pose"][2, 1])) >= 0.2


def translation_cap(run):
    """testTranslationCapRejectsUpdate (:398-427)."""
    r = twice(run, pf.grid(300, 339, 237, 242, 2.0, 1.0), dist=3.0)
    counters(r, 240, 240, 240, 0, 0)
    assert r["inliers"] == 0 and pose_is(r, I4)
    assert math.isinf(r["final_step"]) and not r["converged"] and not r["tracking_ok"]
    assert 0.02 * 240.0 / (0.02 * 240.0 + pf.DAMP_ESC) > pf.TRANS_CAP + 0.5
    s = twice(run, pf.grid(317, 322, 239, 240, 2.0, 1.0), dist=3.0)
    assert s["inliers"] == 12
    assert abs(float(s["pose"][2, 3]) - (-0.24 / 1.24)) <= pf.TOL_MATCH


def adaptive_damping(run):
    """testAdaptiveDamping (:469-532): escalation to 1.0, and level 0's own 0.01."""
    re = twice(run, pf.escalation())
    counters(re, 13, 13, 13, 0, 0)
    tz = float(re["pose"][2, 3])
    assert abs(tz - (-0.14 / 13.25)) <= pf.TOL_MATCH
    assert abs(tz - (-0.14 / 12.35)) >= pf.GAP_FLOOR and abs(tz - (-0.14 / 12.26)) >= pf.GAP_FLOOR
    rw = twice(run, pf.well_conditioned())
    assert rw["inliers"] == 3224 and rw["angle_filtered"] == 0 and rw["dist_filtered"] == 0
    tz = float(rw["pose"][2, 3])
    assert abs(tz - (-0.24 / 24.01)) <= pf.TOL_MATCH
    assert abs(tz - (-0.24 / 24.1)) >= pf.GAP_FLOOR_S and abs(tz - (-0.24 / 25.0)) >= pf.GAP_FLOOR_S
    assert abs(float(rw["final_step"]) - 0.24 / 24.01) <= pf.TOL_MATCH


def model_validity_gate(run):
    """testModelNormalGate (:556-611): squared-norm gates on the model vertex / normal."""
    base = twice(run, pf.model_normal((0, 0, -1)))
    assert base["inliers"] == 11, "an opposite-sign unit model normal is accepted"
    rc = twice(run, pf.model_normal(None))
    assert rc["inliers"] == 10 and abs(float(rc["pose"][2, 3]) - (-0.1 / 11.0)) <= pf.TOL_MATCH
    counters(rc, 10, 10, 10, 0, 0)
    for name, n, accepted in pf.MODEL_NORMAL_CASES:
        r = twice(run, pf.model_normal(n))
        counters(r, 11, 11 if accepted else 10, 11, 0, 0)
        assert r["inliers"] == (11 if accepted else 10), name
        if not accepted:
            assert pose_is(r, rc["pose"]), name + ": a rejected sample leaves the pose bit-identical"
            assert r["error"] == rc["error"], name
    nv = pf.model_normal((0, 0, -1))
    nv[-1].mv[2] = np.nan
    r = twice(run, nv)
    counters(r, 11, 10, 11, 0, 0)
    assert r["inliers"] == 10


def live_normal_gate(run):
    """testLiveNormalGate (:642-686): squared-norm live-normal gate after the distance
    gate and before the angle gate, with no counter of its own."""
    r = twice(run, pf.live_normal((1, 0, 0)))
    counters(r, 10, 10, 10, 0, 0)
    assert r["inliers"] == 10 and abs(float(r["pose"][0, 3])) > 1e-4
    assert abs(float(r["pose"][0, 3]) + 0.1 / 21.0) <= pf.TOL_MATCH
    for bad in ((0, 0, 0), (np.nan, 0, 0), (np.inf, 0, 0)):
        s = twice(run, pf.live_normal(bad))
        assert s["inliers"] == 0
        counters(s, 10, 10, 10, 0, 0)
        assert pose_is(s, I4) and math.isinf(s["final_step"]) and not s["converged"] and not s["tracking_ok"]
    o = twice(run, pf.live_normal((np.nan, 0, 0), 0.01, 0.5))
    counters(o, 10, 10, 10, 10, 0)
    assert o["inliers"] == 0


def angle_gate(run):
    """testAngleThresholdClampAndGate (:729-754): at the clamped 85 degrees orthogonal
    normals are still filtered; aligned ones are not."""
    r = twice(run, pf.orthogonal_normals(), angle=120.0)
    counters(r, 10, 10, 10, 0, 10)
    assert r["inliers"] == 0 and pose_is(r, I4)
    c = run(pf.grid(315, 324, 235, 244, 1.0, 0.01), angle=120.0)
    assert c["angle_filtered"] == 0 and c["inliers"] >= 10


def proper_rotation(run):
    """testSvdReturnsProperRotation (:775-821): a reflected estimate projects onto a
    proper rotation (det +1, orthonormal), translation untouched."""
    cs, est = pf.reflection()
    r = twice(run, cs, est=est)
    assert r["inliers"] == 12 and r["converged"] and abs(r["final_step"]) < 1e-20
    counters(r, 12, 12, 12, 0, 0)
    R = np.asarray(r["pose"], np.float64)[:3, :3]
    assert abs(np.linalg.det(R) - 1.0) <= 1e-5
    assert np.abs(R.T @ R - np.eye(3)).max() <= 1e-4
    assert np.allclose(np.asarray(r["pose"])[:, 3], [0, 0, 0, 1], atol=1e-6)


ALL = [acceptable_final_step, inlier_floor, rotation_cap, translation_cap, adaptive_damping,
       model_validity_gate, live_normal_gate, angle_gate, proper_rotation]
