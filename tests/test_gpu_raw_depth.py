"""Raw Kinect v1 depth codes into the live pyramid (azu_tracker_prepare_raw; reference
sensor boundary cpuDepthMeters, include/sensor/DepthValidity.h:55-93, feeding
buildFrameData, src/sensor/FrameData.cpp:137-194), bit for bit against the oracle:

  * every one of the 65536 codes in one 256 x 256 frame, band [0.3, 2.5]: level-0 depth
    equals the oracle's conversion (accepted set raw 1..954, the pole codes rejected,
    invalid codes exactly +0.0);
  * the synthetic scene encoded as raw codes: all three pyramid levels (depth, vertices,
    normals) equal the oracle's pyramid of the oracle's conversion."""
import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import ICPParams, ICPTracker
from tests.gpu_util import frames, say

pytestmark = pytest.mark.gpu


def test_every_raw_code():
    raw = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    K = np.float32([200.0, 200.0, 127.5, 127.5])
    tr = ICPTracker(256, 256, K, ICPParams(min_depth=0.3, max_depth=2.5))
    d_raw = N.device_array(raw)
    say("prepare_raw over all 65536 codes")
    tr.prepare_raw_device(d_raw.ptr)
    d, _, _ = tr.level_host(0)
    tr.close()
    want = oracle.depth_raw(raw, 0.3, 2.5)
    assert np.array_equal(d.view(np.uint32), want.view(np.uint32))
    acc = np.nonzero(d.reshape(-1))[0]
    assert acc[0] == 1 and acc[-1] == 954 and len(acc) == 954


def test_raw_scene_pyramid_bitwise():
    mm, _ = frames(1)[0]
    m = mm.astype(np.float64) * 1e-3
    with np.errstate(divide="ignore"):
        code = (1.0 / np.where(m > 0, m, np.inf) - 3.3309495161) / -0.0030711016
    raw = np.where(mm > 0, np.clip(np.rint(code), 1, 2046), 0).astype(np.uint16)
    raw[::37, ::41] = 2047  # sensor sentinel pixels
    tr = ICPTracker()
    d_raw = N.device_array(raw)
    say("prepare_raw of the scene frame")
    tr.prepare_raw_device(d_raw.ptr)
    ds, vs_, ns = oracle.pyramid(oracle.depth_raw(raw, 0.3, 5.0), sc.K640)
    for l in range(3):
        d, v, n = tr.level_host(l)
        assert np.array_equal(d.view(np.uint32), ds[l].view(np.uint32)), l
        assert np.array_equal(v.view(np.uint32), vs_[l].view(np.uint32)), l
        assert np.array_equal(n.view(np.uint32), ns[l].view(np.uint32)), l
    tr.close()
    assert np.count_nonzero(ds[0]) > 100000
