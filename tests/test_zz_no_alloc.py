"""azu_integrate_noalloc (the allocation-free integrate) against azu_integrate through
the C ABI: the same volume bit for bit, on two calls in a row, the second with a
smaller frame, for both storage forms; oversized frames are refused."""
import ctypes

import numpy as np
import pytest

from azu_amd import native as N
from azu_amd import scene as sc
from tests.gpu_util import gpu_volume, small_params

pytestmark = pytest.mark.gpu


def _frames():
    T0, T1 = sc.pose(0, 1), sc.pose(3, 1)
    big = sc.render(T0, sc.K640, 640, 480, 10, sc.SEED, True, 0).astype(np.float32) * np.float32(0.001)
    K320 = (262.5, 262.5, 159.75, 119.75)  # scaleIntrinsics of the 640 x 480 camera to 320 x 240
    small = sc.render(T1, K320, 320, 240, 10, sc.SEED, True, 3).astype(np.float32) * np.float32(0.001)
    return [(big, sc.K640, T0), (small, K320, T1)]


def _integrate(fn, vol, depth, K, T):
    from azu_amd.fusion import pose_cm
    H, W = depth.shape
    d = N.device_array(np.ascontiguousarray(depth))
    Kc = N.f32(K)
    N.check(fn(vol.h, d.ptr, W, H, N.fptr(Kc), N.fptr(pose_cm(T)), 0.3, 5.0, None))
    N.synchronize()
    d.free()


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16], ids=["f32", "f16"])
def test_integrate_noalloc_matches_integrate(fmt):
    p = small_params(256, 0.004)
    ref = gpu_volume(p, fmt)
    na = gpu_volume(p, fmt)
    for i, (depth, K, T) in enumerate(_frames()):
        print("noalloc call", i, depth.shape, flush=True)
        _integrate(N.lib.azu_integrate, ref, depth, K, T)
        _integrate(N.lib.azu_integrate_noalloc, na, depth, K, T)
        tr, wr = ref.voxelData()
        tn, wn = na.voxelData()
        assert np.array_equal(tr.view(np.uint32), tn.view(np.uint32)) and np.array_equal(wr, wn), i
        assert ref.counters() == na.counters()
    ref.close()
    na.close()


def test_integrate_noalloc_rejects_oversized_frame():
    p = small_params(64, 0.016)
    vol = gpu_volume(p)
    d = N.DeviceBuffer(4)
    Kc = N.f32(sc.K640)
    from azu_amd.fusion import pose_cm
    rc = N.lib.azu_integrate_noalloc(vol.h, d.ptr, 1600, 1200, N.fptr(Kc), N.fptr(pose_cm(np.eye(4))), 0.3, 5.0,
                                     None)
    assert rc == N.AZU_ERR_ARG and b"scratch" in N.lib.azu_last_error()
    vol.close()
