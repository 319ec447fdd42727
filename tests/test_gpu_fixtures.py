"""The reference's closed-form fixtures run through the GPU path (C ABI)."""
import numpy as np
import pytest

from azu_amd import native as N
from azu_amd.fusion import ICPParams, ICPTracker, TSDFParams, TSDFVolume
from tests import fixtures as fx

pytestmark = pytest.mark.gpu


def _icp_fixture(corruption):
    lv, ln, mv, mn = fx.icp_weighting(corruption)
    tr = ICPTracker(640, 480, fx.K640, ICPParams(max_iterations=(1, 0, 0)))
    for l in range(3):
        tr.load_level(l, lv[l], ln[l])
    I4 = np.eye(4, dtype=np.float32)
    return tr.track(mv, mn, I4, I4)


def test_icp_weighting_canonical_gpu():
    """tests/icp_weighting_contract.cpp:163-216 through the GPU tracker."""
    r = _icp_fixture("clean")
    assert r["valid_live_points"] == 11 and r["valid_model_points"] == 11 and r["projected_points"] == 11
    assert r["inliers"] == 11 and r["dist_filtered"] == 0 and r["angle_filtered"] == 0
    tz = float(r["pose"][2, 3])
    assert abs(tz - fx.ICP_TZ) <= fx.ICP_TOL
    assert abs(tz - fx.ICP_TZ_UNWEIGHTED) >= fx.ICP_GAP and abs(tz - fx.ICP_TZ_W2) >= fx.ICP_GAP
    assert abs(r["error"] - fx.ICP_ERR) <= fx.ICP_TOL
    assert abs(float(r["pose"][0, 3])) <= fx.ICP_TOL and abs(float(r["pose"][1, 3])) <= fx.ICP_TOL
    r2 = _icp_fixture("clean")
    assert np.array_equal(r["pose"], r2["pose"]) and r["error"] == r2["error"]  # deterministic reduction


@pytest.mark.parametrize("corruption", ["inf_normal", "nan_depth"])
def test_icp_weighting_rejects_non_finite_gpu(corruption):
    clean = _icp_fixture("clean")
    r = _icp_fixture(corruption)
    assert r["projected_points"] == 12 and r["inliers"] == 11 and r["valid_model_points"] == 11
    assert np.isfinite(r["pose"]).all() and np.isfinite(r["error"])
    assert np.array_equal(r["pose"], clean["pose"]) and r["error"] == clean["error"]


def test_pyramid_jump_guard_gpu():
    """tests/icp_weighting_contract.cpp:277-391 through the GPU downsample."""
    d, v, n = fx.pyramid_field()
    tr = ICPTracker(16, 16, (1.0, 1.0, 0.0, 0.0))
    tr.load_level(0, v, n, depth=d)
    tr.downsample()
    d1, v1, n1 = tr.level_host(1)
    fx.check_pyramid_field((d, v, n), (d1, v1, n1))


@pytest.mark.parametrize("fmt,mode", [(N.FMT_F32, N.MODE_EXACT), (N.FMT_F32, N.MODE_FAST),
                                      (N.FMT_F16_U16, N.MODE_EXACT), (N.FMT_F16_U16, N.MODE_FAST)])
def test_integrate_hand_fold_gpu(fmt, mode):
    """tests/tsdf_integration_race_contract.cpp:232-324 for every storage form and kernel.
    FP16 tsdf storage rounds to 11 significant bits after every fold; two folds of values
    in [0.5, 1) can drift by one binary16 ulp there (2^-11), so its bar is 5e-4, not 2e-5."""
    g = TSDFVolume(TSDFParams(**fx.FOLD_PARAMS), fmt=fmt, mode=mode)
    I4 = np.eye(4, dtype=np.float32)
    for d in fx.fold_frames():
        g.integrate(d, fx.FOLD_K, I4, *fx.FOLD_BAND)
    t, w = g.voxelData()
    if fmt == N.FMT_F32:
        fx.check_fold(t, w)
    else:
        for k in range(1, 64):
            te, we = fx.fold_expect(k)
            assert w[k, 32, 32] == we and abs(float(t[k, 32, 32]) - te) < 5e-4
        assert w[50, 32, 32] == 2.0 and w[54, 32, 32] == 0.0 and w[50, 32, 33] == 0.0
