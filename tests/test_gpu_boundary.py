"""Boundary behaviour of the C ABI on the GPU: host transfers and meshing after work
queued on a non-blocking stream, single-voxel reads and writes, the usage count, and
parameter changes (reference TSDFVolume.cpp:124-146, 563-569)."""
import ctypes

import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import TSDFParams, pose_cm
from tests.gpu_util import frames, gpu_volume, oracle_volume, small_params

pytestmark = pytest.mark.gpu


def _depths(n):
    return [(oracle.depth_mm(mm), T) for mm, T in frames(n)]


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16], ids=["f32", "f16"])
def test_download_and_mesh_wait_for_other_streams(fmt):
    """Integrate on a non-blocking stream, then download / mesh / read voxels on the null
    stream with no explicit synchronisation: each waits for the queued work."""
    p = small_params(256, 0.004)
    ref = gpu_volume(p, fmt)
    vol = gpu_volume(p, fmt)
    fr = _depths(4)
    for d, T in fr:
        ref.integrate(d, sc.K640, T)
    s = ctypes.c_void_p()
    N.check(N.lib.azu_stream_create(ctypes.byref(s)))
    bufs = [N.device_array(d) for d, _ in fr]
    Kc = N.f32(sc.K640)
    for b, (_, T) in zip(bufs, fr):
        N.check(N.lib.azu_integrate(vol.h, b.ptr, 640, 480, N.fptr(Kc), N.fptr(pose_cm(T)), 0.3, 5.0, s))
    t, w = vol.voxelData()  # azu_vol_download on the null stream, nothing synchronised before
    tr, wr = ref.voxelData()
    assert np.array_equal(t.view(np.uint32), tr.view(np.uint32)) and np.array_equal(w, wr)
    vol.reset(s)
    for b, (_, T) in zip(bufs, fr):
        N.check(N.lib.azu_integrate(vol.h, b.ptr, 640, 480, N.fptr(Kc), N.fptr(pose_cm(T)), 0.3, 5.0, s))
    pos, _, idx, _ = vol.marching_cubes()  # azu_mc_extract, null stream
    pr, _, ir, _ = ref.marching_cubes()
    assert np.array_equal(pos.view(np.uint32), pr.view(np.uint32)) and np.array_equal(idx, ir)
    N.check(N.lib.azu_stream_sync(s))
    N.check(N.lib.azu_stream_destroy(s))


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16], ids=["f32", "f16"])
def test_voxel_read_write_and_usage(fmt):
    p = small_params(64, 0.016)
    vol = gpu_volume(p, fmt)
    d, T = _depths(1)[0]
    vol.integrate(d, sc.K640, T)
    t, w = vol.voxelData()
    flat_t, flat_w = t.reshape(-1), w.reshape(-1)
    first, count = 64 * 64 * 30 + 64 * 20, 200
    rt = np.empty(count, np.float32)
    rw = np.empty(count, np.float32)
    N.check(N.lib.azu_vol_read_voxels(vol.h, first, count, rt.ctypes.data, rw.ctypes.data))
    assert np.array_equal(rt.view(np.uint32), flat_t[first:first + count].view(np.uint32))
    assert np.array_equal(rw, flat_w[first:first + count])
    used = ctypes.c_ulonglong()
    N.check(N.lib.azu_vol_usage(vol.h, ctypes.byref(used)))
    assert used.value == int((flat_w > 0).sum()) > 0
    # write a range, read it back; the whole-volume download agrees
    wt = np.linspace(-0.9, 0.9, count).astype(np.float32)
    ww = np.arange(count, dtype=np.float32) % 7
    N.check(N.lib.azu_vol_write_voxels(vol.h, first, count, wt.ctypes.data, ww.ctypes.data))
    N.check(N.lib.azu_vol_read_voxels(vol.h, first, count, rt.ctypes.data, rw.ctypes.data))
    want_t = wt if fmt == N.FMT_F32 else wt.astype(np.float16).astype(np.float32)
    assert np.array_equal(rt.view(np.uint32), want_t.view(np.uint32)) and np.array_equal(rw, ww)
    t2, w2 = vol.voxelData()
    assert np.array_equal(t2.reshape(-1)[first:first + count].view(np.uint32), want_t.view(np.uint32))
    # out-of-range and non-integer weights are refused
    assert N.lib.azu_vol_read_voxels(vol.h, 64 ** 3 - 10, 11, rt.ctypes.data, rw.ctypes.data) == N.AZU_ERR_ARG
    if fmt == N.FMT_F16_U16:
        bad = np.full(count, 0.5, np.float32)
        assert N.lib.azu_vol_write_voxels(vol.h, first, count, wt.ctypes.data, bad.ctypes.data) == N.AZU_ERR_ARG
    vol.close()


def test_voxel_write_reaches_raycast():
    """A voxel range written through azu_vol_write_voxels is what the next raycast sees
    (its empty-space map is rebuilt): a wall written into an empty volume is hit at the
    same place as the oracle's raycast of the same voxels."""
    p = dict(resolution=32, voxel_size=0.04, truncation=0.12, max_weight=128.0, origin=(-0.64, -0.64, 0.0),
             min_depth=0.3, max_depth=5.0)
    vol = gpu_volume(p)
    k = np.arange(32, dtype=np.float32)
    col = np.clip((np.float32(0.38) - k * np.float32(0.04)) / np.float32(0.12), -1, 1).astype(np.float32)
    t = np.broadcast_to(col[:, None, None], (32, 32, 32)).astype(np.float32)
    w = np.ones((32, 32, 32), np.float32)
    v0, _ = vol.raycast(np.eye(4), (50, 50, 31.5, 23.5), 64, 48)
    assert not v0.any()
    N.check(N.lib.azu_vol_write_voxels(vol.h, 0, 32 ** 3, t.ctypes.data, w.ctypes.data))
    v1, n1 = vol.raycast(np.eye(4), (50, 50, 31.5, 23.5), 64, 48)
    o = oracle_volume(p)
    o.set(t, w)
    vo, no = o.raycast(np.eye(4), (50, 50, 31.5, 23.5), 64, 48)
    assert np.array_equal(v1.view(np.uint32), vo.view(np.uint32)) and np.array_equal(n1.view(np.uint32),
                                                                                      no.view(np.uint32))
    assert abs(float(v1[24, 32, 2]) - 0.38) < 1e-5


def test_set_params_keeps_a_usable_volume():
    """setParams: equal parameters keep the voxels; a changed field clears them; a
    resolution change reallocates and the volume works at the new size; invalid
    parameters are refused and leave the volume as it was."""
    p = small_params(64, 0.016)
    vol = gpu_volume(p)
    d, T = _depths(1)[0]
    vol.integrate(d, sc.K640, T)
    before = vol.voxelData()
    vol.setParams(TSDFParams(**p))
    after = vol.voxelData()
    assert np.array_equal(before[0].view(np.uint32), after[0].view(np.uint32))
    bad = dict(p, max_weight=0.5)
    with pytest.raises(N.AzuError):
        vol.setParams(TSDFParams(**bad))
    same = vol.voxelData()
    assert np.array_equal(before[0].view(np.uint32), same[0].view(np.uint32))
    p2 = dict(p, resolution=96)
    vol.setParams(TSDFParams(**p2))
    t, w = vol.voxelData()
    assert t.shape == (96, 96, 96) and not w.any() and (t == 1).all()
    vol.integrate(d, sc.K640, T)
    o = oracle_volume(p2)
    o.integrate(d, sc.K640, T)
    to, wo = o.get()
    t, w = vol.voxelData()
    assert np.array_equal(t.view(np.uint32), to.view(np.uint32)) and np.array_equal(w, wo)
