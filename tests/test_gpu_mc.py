"""GPU marching cubes against the CPU mesher restatement: identical vertices, normals and
indices (bit for bit, same order) on identical volumes, with and without the triangle cap."""
import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import TSDFParams, TSDFVolume
from tests.gpu_util import frames, say

pytestmark = pytest.mark.gpu


def _fused(res, vs, fmt, nframes=3):
    p = TSDFParams(res, vs, 3 * vs, 128.0, sc.ORIGIN, 0.3, 5.0)
    v = TSDFVolume(p, fmt=fmt)
    for mm, T in frames(nframes):
        v.integrate(mm.astype(np.float32) * np.float32(0.001), sc.K640, T)
    return v


def _check(v, cap=2000000):
    pos, nrm, idx, tr = v.marching_cubes(cap)
    t, w = v.voxelData()
    rp, rn, ri, rtr = oracle.mc_extract(t, w, v.params.voxel_size, v.params.origin, cap)
    say("mesh: %d vertices, %d triangles, truncated %s (oracle %d, %d)" % (len(pos), len(idx) // 3, tr, len(rp),
                                                                           len(ri) // 3))
    assert tr == rtr
    assert pos.shape == rp.shape and idx.shape == ri.shape
    assert np.array_equal(idx, ri)
    assert np.array_equal(pos.view(np.uint32), rp.view(np.uint32))
    assert np.array_equal(nrm.view(np.uint32), rn.view(np.uint32))
    return len(idx) // 3


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16])
def test_mc_bitwise_fused(fmt):
    v = _fused(128, 0.008, fmt)
    assert _check(v) > 1000


def test_mc_bitwise_capped():
    v = _fused(96, 0.0107, N.FMT_F32)
    n = _check(v)
    _check(v, cap=n // 3)


def test_mc_empty_and_tiny():
    v = TSDFVolume(TSDFParams(2, 0.01, 0.03, 128.0, (0, 0, 0), 0.3, 5.0))
    pos, nrm, idx, tr = v.marching_cubes()
    assert len(pos) == 0 and len(idx) == 0 and not tr


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16])
def test_mc_bitwise_512(fmt):
    """The 512^3 / 2 mm volume after 4 frames of the sphere + plane + 10 boxes sequence:
    bit-exact against the CPU mesher at full size, and timed (host clock over 3
    extractions, allocation and download included)."""
    import time
    frames_, truth, K = sc.sequence("c2_512", frames=4)
    v = TSDFVolume(TSDFParams(**sc.volume_params("c2_512")), fmt=fmt)
    for mm, T in zip(frames_, truth):
        v.integrate(mm.astype(np.float32) * np.float32(0.001), K, T)
    n = _check(v)
    t0 = time.perf_counter()
    for _ in range(3):
        v.marching_cubes()
    say("mesh 512^3 fmt %d: %d triangles, %.2f ms per extraction" % (fmt, n, (time.perf_counter() - t0) / 3 * 1e3))
    v.close()
    assert n > 100000
