"""Global point cloud of a GPU volume (azu_vol_point_cloud; reference
TSDFVolume::extractGlobalPointCloud, src/tsdf/TSDFVolume.cpp:571-632) against the
oracle's serial restatement, bit for bit and in the same linear order:

  * uploaded fields with the predicate's edges (weight exactly 1, |tsdf| exactly 0.2,
    NaN) at a resolution that is not a multiple of the 64-voxel row chunk;
  * a fused 256^3 volume (sphere + plane sequence), both storage forms;
  * a capacity below the count writes exactly the first `capacity` points and reports
    the full count."""
import ctypes

import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.native import check, lib
from tests.gpu_util import gpu_volume, say, small_params

pytestmark = pytest.mark.gpu


def cloud_of(vol):
    t, w = vol.voxelData()
    p = vol.params
    return oracle.point_cloud(t, w, p.voxel_size, p.origin)


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16])
def test_point_cloud_edges_uploaded(fmt):
    res = 100
    rng = np.random.default_rng(3)
    t = rng.uniform(-0.5, 0.5, (res, res, res)).astype(np.float32)
    w = rng.integers(0, 4, (res, res, res)).astype(np.float32)
    t.reshape(-1)[::97] = np.float32(0.2)
    t.reshape(-1)[1::97] = np.float32(-0.2)
    t.reshape(-1)[2::89] = np.nan
    w.reshape(-1)[3::89] = 1.0
    vol = gpu_volume(small_params(res=res, vs=0.01), fmt=fmt)
    vol.upload(t, w)
    say("point cloud %d^3 fmt %d" % (res, fmt))
    got = vol.point_cloud()
    want = cloud_of(vol)
    vol.close()
    assert got.shape == want.shape and len(want) > 1000
    assert np.array_equal(got, want)


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16])
def test_point_cloud_fused_volume(fmt):
    frames, truth, K = sc.sequence("c1_256", frames=4)
    vol = gpu_volume(sc.volume_params("c1_256"), fmt=fmt)
    for f, T in zip(frames, truth):
        vol.integrate(f.astype(np.float32) * np.float32(0.001), K, T)
    say("point cloud of a fused 256^3 volume, fmt %d" % fmt)
    got = vol.point_cloud()
    want = cloud_of(vol)
    vol.close()
    say("  %d points" % len(want))
    assert len(want) > 10000
    assert np.array_equal(got, want)


def test_point_cloud_capacity_below_count():
    res = 64
    t = np.zeros((res, res, res), np.float32)
    w = np.full((res, res, res), 2.0, np.float32)
    vol = gpu_volume(small_params(res=res, vs=0.01))
    vol.upload(t, w)
    n = res ** 3
    cap = 1000
    buf = N.device_array(np.full(3 * (cap + 16), -7.0, np.float32))
    got_n = ctypes.c_ulonglong()
    check(lib.azu_vol_point_cloud(vol.h, buf.ptr, cap, ctypes.byref(got_n), None))
    out = buf.download(np.float32, (cap + 16, 3))
    want = cloud_of(vol)
    vol.close()
    assert got_n.value == n == len(want)
    assert np.array_equal(out[:cap], want[:cap])
    assert np.all(out[cap:] == -7.0)


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16])
def test_point_cloud_fused_512(fmt):
    """The 512^3 / 2 mm volume after 4 frames of the sphere + plane + 10 boxes sequence:
    bit-exact against the oracle at full size, and timed (both passes, the scan and the
    count read-back; event-timed over 3 calls)."""
    import time
    frames, truth, K = sc.sequence("c2_512", frames=4)
    vol = gpu_volume(sc.volume_params("c2_512"), fmt=fmt)
    for f, T in zip(frames, truth):
        vol.integrate(f.astype(np.float32) * np.float32(0.001), K, T)
    got = vol.point_cloud()
    t0 = time.perf_counter()
    for _ in range(3):
        vol.point_cloud()
    ms = (time.perf_counter() - t0) / 3 * 1e3
    want = cloud_of(vol)
    vol.close()
    say("point cloud 512^3 fmt %d: %d points, %.2f ms per extraction (host clock, incl. download)"
        % (fmt, len(want), ms))
    assert len(want) > 100000
    assert np.array_equal(got, want)
