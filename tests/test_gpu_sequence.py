"""The 100-frame 512^3 frame loop on the GPU against the oracle's tracked poses
(tests/golden/c2_512_poses.npz, written by tests/golden/make_sequence_poses.py).

Bars: every frame's pose within 1 mm / 0.1 deg of the FP64 oracle, for the FP32 form and
for the FP16 + UInt16 form; the FP32 form's updated-voxel count equals the FP32
oracle's on frame 0 (identical inputs, bit-exact integration)."""
import os

import numpy as np
import pytest

from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.pipeline import FrameLoop
from tests.gpu_util import pose_diff, say

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "c2_512_poses.npz")


@pytest.fixture(scope="module")
def sequence():
    frames, truth, K = sc.sequence("c2_512", frames=100)
    return frames, truth, K, [N.device_array(f) for f in frames]


@pytest.mark.parametrize("fmt,name", [(N.FMT_F32, "fp32"), (N.FMT_F16_U16, "fp16")])
def test_sequence_against_fp64_oracle(sequence, fmt, name):
    if not os.path.exists(GOLDEN):
        pytest.skip("golden poses not generated")
    g = np.load(GOLDEN)
    frames, truth, K, dev = sequence
    loop = FrameLoop(sc.volume_params("c2_512"), fmt, 640, 480, K)
    loop.reset(truth[0])
    if fmt == N.FMT_F32:
        loop.frame(dev[0].ptr)
        upd, _ = loop.vol.counters(reset=True)
        assert upd == int(g["fp32_n_upd"][0])
        rest = dev[1:]
    else:
        rest = dev
    for d in rest:
        loop.frame(d.ptr)
    P = loop.poses()
    loop.close()
    worst_t = worst_r = 0.0
    worst32 = 0.0
    for k in range(100):
        dt, ang = pose_diff(P[k], g["fp64"][k])
        worst_t, worst_r = max(worst_t, dt), max(worst_r, ang)
        worst32 = max(worst32, pose_diff(P[k], g["fp32"][k])[0])
    terr = np.abs(P[:, :3, 3] - truth[:, :3, 3]).max()
    say(name, "vs fp64 oracle: max %.3f mm, %.4f deg; vs fp32 oracle max %.3f mm; vs truth %.3f mm" % (
        1000 * worst_t, np.degrees(worst_r), 1000 * worst32, 1000 * terr))
    assert worst_t <= 1e-3 and np.degrees(worst_r) <= 0.1
