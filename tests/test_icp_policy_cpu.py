"""The reference's ICP numeric-policy contract (tests/icp_numeric_policy_contract.cpp:
241-849) on the CPU oracle, FP32 and FP64, plus the parameter sanitiser and the
default result of the C ABI (host code, no GPU)."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from tests import policy_contract as pc
from tests import policy_fixtures as pf


def oracle_runner(dtype):
    def run(cs, dist=0.1, angle=30.0, est=None):
        lv, ln, mv, mn = pf.corr_arrays(cs)
        T = pc.I4 if est is None else est
        r = oracle.icp_track(lv, ln, mv, mn, pf.K640, T, pc.I4, iters=(1, 0, 0), dist=dist, angle=angle,
                             dtype=dtype)
        r["converged"] = bool(r["converged"])
        r["tracking_ok"] = bool(r["tracking_ok"])
        r["pose"] = np.asarray(r["pose"], np.float32) if dtype == np.float32 else np.asarray(r["pose"])
        return r
    return run


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("scenario", pc.ALL, ids=[f.__name__ for f in pc.ALL])
def test_icp_policy_oracle(scenario, dtype):
    if dtype == np.float64:
        # pose_is compares float32 bit patterns; the FP64 oracle's pose is compared in float32
        def run(cs, **kw):
            r = oracle_runner(np.float64)(cs, **kw)
            r["pose"] = np.asarray(r["pose"], np.float32)
            return r
        scenario(run)
    else:
        scenario(oracle_runner(np.float32))


def test_fixture_projection_lands_where_the_contract_says():
    """The fixture's own projection (projectedPixel, :116-121) puts the identity-pose
    live vertex of an on-grid pixel back on that pixel, so no correspondence is lost."""
    for c in pf.grid(300, 339, 237, 242, 1.0, 5e-4) + pf.well_conditioned():
        assert (c.mpx, c.mpy) == (c.lpx, c.lpy)


def test_icp_params_sanitize_abi():
    """ICPTracker::sanitizeParams (src/tracking/ICPTracker.cpp:24-34) through the C ABI:
    the angle threshold clamps to [0, 85] (non-finite -> 30) and no other field changes
    (icp_numeric_policy_contract.cpp:711-728)."""
    from azu_amd import native as N

    for a_in, a_out in pf.ANGLE_CLAMP_CASES:
        p = N.IcpParams()
        for i, it in enumerate((7, 3, 2)):
            p.max_iterations[i] = it
        p.dist_threshold, p.angle_threshold, p.min_depth, p.max_depth = 0.25, a_in, 0.4, 4.5
        o = N.IcpParams()
        N.check(N.lib.azu_icp_params_sanitize(ctypes.byref(p), ctypes.byref(o)))
        assert o.angle_threshold == np.float32(a_out), (a_in, o.angle_threshold)
        assert list(o.max_iterations) == [7, 3, 2]
        assert (o.dist_threshold, o.min_depth, o.max_depth) == (np.float32(0.25), np.float32(0.4), np.float32(4.5))


def test_default_result_abi():
    """testDefaultResultPoseIsIdentity (:825-836): a fresh result is the identity pose,
    final_step +inf, not converged, not ok, no inliers, zero error."""
    from azu_amd import native as N

    r = N.IcpResult()
    N.check(N.lib.azu_icp_result_default(ctypes.byref(r)))
    assert np.array_equal(np.frombuffer(r.pose, np.float32), np.eye(4, dtype=np.float32).reshape(16))
    assert math.isinf(r.final_step) and r.final_step > 0
    assert not r.converged and not r.tracking_ok and r.inliers == 0 and r.error == 0.0
