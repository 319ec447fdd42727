"""Closed-form pins of the GPU live frame geometry and constant-velocity estimate
(C ABI): the same closed forms as tests/test_geometry_pins_cpu.py, plus bit-equality
with the oracle on the same depth maps."""
import ctypes

import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd.fusion import ICPTracker, pose_cm, pose_rm
from tests import geometry_pins as gp

pytestmark = pytest.mark.gpu


def gpu_level0(depth):
    tr = ICPTracker(gp.W, gp.H, gp.K640)
    buf = N.device_array(np.ascontiguousarray(depth, np.float32))
    tr.prepare_m_device(buf.ptr)
    N.synchronize()
    _, v, n = tr.level_host(0)
    tr.close()
    return v, n


def oracle_level0(depth):
    _, vs, ns = oracle.pyramid(depth, gp.K640)
    return vs[0], ns[0]


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def test_tilted_plane_gpu():
    d = gp.plane_depth()
    v, n = gpu_level0(d)
    gp.check_plane(v, n, d)
    vo, no = oracle_level0(d)
    assert same_bits(v, vo) and same_bits(n, no)


@pytest.mark.parametrize("d0,step,zeroed", gp.STEP_CASES)
def test_normal_depth_jump_threshold_gpu(d0, step, zeroed):
    d = gp.step_depth(d0, step)
    v, n = gpu_level0(d)
    gp.check_step(n, zeroed)
    assert same_bits(n, oracle_level0(d)[1])


def test_normal_hole_gpu():
    v, n = gpu_level0(gp.hole_depth())
    gp.check_hole(v, n)


@pytest.mark.parametrize("kind", ["world", "body"])
def test_predict_pose_uniform_motion_gpu(kind):
    """azu_tracker_predict: last = the pose set, current = the tracked pose; the
    prediction is cur (last^-1 cur) = the next pose of a uniform motion, and the
    oracle's float evaluation of the same product bit for bit."""
    tr = ICPTracker(gp.W, gp.H, gp.K640)
    for k in (1, 5, 20):
        last, cur, nxt = (gp.uniform_motion(j, kind).astype(np.float32) for j in (k - 1, k, k + 1))
        N.check(N.lib.azu_tracker_set_pose(tr.h, N.fptr(pose_cm(last)), None))
        N.check(N.lib.azu_tracker_set_tracked_pose(tr.h, N.fptr(pose_cm(cur)), None))
        dp = ctypes.c_void_p()
        N.check(N.lib.azu_tracker_predict(tr.h, ctypes.byref(dp), None))
        out = np.empty(16, np.float32)
        N.synchronize()
        N.check(N.lib.azu_memcpy_d2h(out.ctypes.data, dp.value, 64, None))
        N.synchronize()
        p = pose_rm(out)
        assert np.abs(p.astype(np.float64) - gp.uniform_motion(k + 1, kind)).max() <= 2e-6
        assert same_bits(p, oracle.predict_pose(cur, last))
    tr.close()
