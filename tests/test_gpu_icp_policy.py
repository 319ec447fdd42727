"""The reference's ICP numeric-policy contract (tests/icp_numeric_policy_contract.cpp:
241-849) through the GPU tracker (C ABI): one level-0 iteration over correspondence
sets loaded with azu_tracker_load_level, the same closed forms as on the oracle."""
import numpy as np
import pytest

from azu_amd.fusion import ICPParams, ICPTracker
from tests import policy_contract as pc
from tests import policy_fixtures as pf

pytestmark = pytest.mark.gpu


def gpu_run(cs, dist=0.1, angle=30.0, est=None):
    lv, ln, mv, mn = pf.corr_arrays(cs)
    tr = ICPTracker(640, 480, pf.K640, ICPParams(max_iterations=(1, 0, 0), dist_threshold=dist,
                                                  angle_threshold=angle))
    try:
        for l in range(3):
            tr.load_level(l, lv[l], ln[l])
        return tr.track(mv, mn, pc.I4 if est is None else est, pc.I4)
    finally:
        tr.close()


@pytest.mark.parametrize("scenario", pc.ALL, ids=[f.__name__ for f in pc.ALL])
def test_icp_policy_gpu(scenario):
    print("icp policy scenario", scenario.__name__, flush=True)
    scenario(gpu_run)


@pytest.mark.parametrize("a_in,a_out", pf.ANGLE_CLAMP_CASES)
def test_icp_params_clamp_gpu(a_in, a_out):
    """The tracker's own parameters after creation and after setParams (:711-728)."""
    import ctypes
    from azu_amd import native as N

    p = ICPParams(max_iterations=(7, 3, 2), dist_threshold=0.25, angle_threshold=a_in, min_depth=0.4,
                  max_depth=4.5)
    for how in ("ctor", "setParams"):
        tr = ICPTracker(640, 480, pf.K640, p if how == "ctor" else ICPParams())
        if how == "setParams":
            tr.setParams(p)
        o = N.IcpParams()
        N.check(N.lib.azu_tracker_get_params(tr.h, ctypes.byref(o)))
        tr.close()
        assert o.angle_threshold == np.float32(a_out) and list(o.max_iterations) == [7, 3, 2]
        assert (o.dist_threshold, o.min_depth, o.max_depth) == (np.float32(0.25), np.float32(0.4), np.float32(4.5))
