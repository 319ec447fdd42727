"""The large-volume form: a 1024^3 FP16 + UInt16 TSDF (4 GiB, 1 mm voxels), 1280x960
depth from 2x bicubic upsampling of the 640x480 frames, 300 frames of the sphere + plane
+ boxes scene (orbit + jitter), marching cubes at the end compared bit for bit with the
CPU mesher run on the same (downloaded) volume.

There is no CPU tracking oracle at this size (minutes per frame), so the tracked poses
are held to the scene's true poses: within 2 mm / 0.25 deg over all 300 frames (measured
about 1 mm / 0.11 deg; the 512^3 sequence, with its FP64 oracle, stays within 0.6 mm of
truth). The mesh is compared twice: under the reference's 2 M triangle cap, and whole.
"""
import time

import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.pipeline import FrameLoop
from tests.gpu_util import pose_diff, say

pytestmark = pytest.mark.gpu


def test_large_volume_sequence_and_mesh():
    t0 = time.perf_counter()
    frames, truth, K = sc.sequence("c5_1024", frames=300)
    say("rendered 300 frames %dx%d in %.1f s" % (frames.shape[2], frames.shape[1], time.perf_counter() - t0))
    dev = [N.device_array(f) for f in frames]
    p = sc.volume_params("c5_1024")
    loop = FrameLoop(p, N.FMT_F16_U16, 1280, 960, K)
    loop.reset(truth[0])
    t0 = time.perf_counter()
    for i, d in enumerate(dev):
        loop.frame(d.ptr)
        if i % 100 == 99:
            N.synchronize()
            say("frame %d, %.2f ms/frame so far" % (i + 1, 1000 * (time.perf_counter() - t0) / (i + 1)))
    P = loop.poses()
    d = np.array([pose_diff(P[k], truth[k]) for k in range(300)])
    say("vs truth over 300 frames: max %.3f mm %.4f deg, median %.3f mm" % (
        1000 * d[:, 0].max(), np.degrees(d[:, 1].max()), 1000 * np.median(d[:, 0])))
    assert d[:, 0].max() <= 2e-3 and np.degrees(d[:, 1].max()) <= 0.25

    t, w = loop.vol.voxelData()
    for cap in (2000000, 200000000):
        t0 = time.perf_counter()
        pos, nrm, idx, tr = loop.vol.marching_cubes(cap)
        say("GPU mesh (cap %d): %d vertices, %d triangles, truncated %s (%.2f s)" % (
            cap, len(pos), len(idx) // 3, tr, time.perf_counter() - t0))
        t0 = time.perf_counter()
        rp, rn, ri, rtr = oracle.mc_extract(t, w, p["voxel_size"], p["origin"], cap)
        say("CPU mesher: %d vertices, %d triangles (%.1f s)" % (len(rp), len(ri) // 3, time.perf_counter() - t0))
        assert tr == rtr and tr == (cap == 2000000)
        assert pos.shape == rp.shape and np.array_equal(idx, ri)
        assert np.array_equal(pos.view(np.uint32), rp.view(np.uint32))
        assert np.array_equal(nrm.view(np.uint32), rn.view(np.uint32))
        del pos, nrm, idx, rp, rn, ri
    loop.close()
