"""The reference's fusion-bias and raycast-accuracy contracts on the GPU, both storage
forms (tests/plane_contracts.py has the fixtures and bars); the FP32 form also equals
the FP32 oracle bit for bit on the same fixture."""
import ctypes

import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd.fusion import TSDFParams, TSDFVolume
from tests import plane_contracts as pc
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

FORMS = [(N.FMT_F32, "f32"), (N.FMT_F16_U16, "f16")]


def _integrate(v, n):
    d = pc.plane_depth()
    for _ in range(n):
        v.integrate(d, pc.K, np.eye(4, dtype=np.float32), *pc.BAND)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("fmt,name", FORMS)
def test_fusion_bias_gpu(fmt, name):
    v = TSDFVolume(TSDFParams(**pc.PARAMS), fmt=fmt)
    _integrate(v, 10)
    pos, _, _, trunc = v.marching_cubes()
    assert not trunc
    mean, p95 = pc.check_bias(pos)
    say("fusion bias %s: %d vertices, mean %.4f mm, p95 %.4f mm" % (name, len(pos), 1e3 * mean, 1e3 * p95))
    if fmt == N.FMT_F32:
        p = pc.PARAMS
        o = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                          p["min_depth"], p["max_depth"])
        _integrate(o, 10)
        t, w = v.voxelData()
        to, wo = o.get()
        assert np.array_equal(_bits(t), _bits(to)) and np.array_equal(_bits(w), _bits(wo))
        o.close()
    v.close()


@pytest.mark.parametrize("fmt,name", FORMS)
def test_weights_count_frames_gpu(fmt, name):
    for n in (1, 3, 70):
        v = TSDFVolume(TSDFParams(**pc.PARAMS), fmt=fmt)
        _integrate(v, n)
        used = ctypes.c_ulonglong()
        N.check(N.lib.azu_vol_usage(v.h, ctypes.byref(used)))
        pc.check_weights(v.voxelData()[1], n, used.value)
        v.close()


@pytest.mark.parametrize("fmt,name", FORMS)
def test_repeat_runs_bit_identical_gpu(fmt, name):
    vols = []
    for _ in range(2):
        v = TSDFVolume(TSDFParams(**pc.PARAMS), fmt=fmt)
        _integrate(v, 3)
        vols.append(v.voxelData())
        v.close()
    assert all(np.array_equal(_bits(a), _bits(b)) for a, b in zip(vols[0], vols[1]))


@pytest.mark.parametrize("fmt,name", FORMS)
def test_raycast_accuracy_gpu(fmt, name):
    v = TSDFVolume(TSDFParams(**pc.PARAMS), fmt=fmt)
    _integrate(v, 5)
    front = v.raycast(pc.POSE_FRONT, pc.K, pc.W, pc.H)
    behind = v.raycast(pc.POSE_BEHIND, pc.K, pc.W, pc.H)
    oblique = v.raycast(pc.POSE_OBLIQUE, pc.K, pc.W, pc.H)
    s = [pc.cast_stats(*r) for r in (front, behind, oblique)]
    say("raycast accuracy %s:" % name, s)
    pc.check_front(s[0])
    pc.check_behind(s[1])
    pc.check_oblique(s[2])
    if fmt == N.FMT_F32:
        p = pc.PARAMS
        o = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                          p["min_depth"], p["max_depth"])
        _integrate(o, 5)
        for pose, got in ((pc.POSE_FRONT, front), (pc.POSE_BEHIND, behind), (pc.POSE_OBLIQUE, oblique)):
            vo, no = o.raycast(pose, pc.K, pc.W, pc.H)
            assert np.array_equal(_bits(got[0]), _bits(vo)) and np.array_equal(_bits(got[1]), _bits(no))
        o.close()
    v.close()
