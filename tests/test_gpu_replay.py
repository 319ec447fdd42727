"""Lockstep replay of a fakenect recording on the GPU (azu_amd.replay; reference
tools/azu_replay.cpp): a recording written from the 512^3 scene sequence as raw Kinect
codes is replayed through the graded loop; every frame grades Good, the poses follow the
scene truth, and trace.csv has one "track" row per frame and one "integ" row per
integrated frame, in the FrameTrace schema."""
import csv

import numpy as np
import pytest

from azu_amd import native as N
from azu_amd import replay as rp
from azu_amd import scene as sc
from tests.gpu_util import pose_diff, say

pytestmark = pytest.mark.gpu


def mm_to_raw(mm):
    m = mm.astype(np.float64) * 1e-3
    with np.errstate(divide="ignore"):
        code = (1.0 / np.where(m > 0, m, np.inf) - 3.3309495161) / -0.0030711016
    return np.where(mm > 0, np.clip(np.rint(code), 1, 2046), 0).astype(np.uint16)


def test_replay_recording(tmp_path):
    nf = 8
    frames, truth, K = sc.sequence("c2_512", frames=nf)
    rec = tmp_path / "rec"
    rec.mkdir()
    names = []
    for k, mm in enumerate(frames):
        name = "d-%.6f-%d-%d.pgm" % (1700000000 + k / 30.0, 1000 + 2000000 * k, k)
        rp.write_depth(str(rec / name), mm_to_raw(mm))
        names.append(name)
    (rec / "INDEX.txt").write_text("\n".join(names) + "\n")
    say("replay of an %d-frame raw recording" % nf)
    records, poses = rp.replay(str(rec), str(tmp_path / "out"), sc.volume_params("c2_512"), K, pose0=truth[0])
    assert len(records) == nf and len(poses) == nf
    assert all(r["quality"] == N.TRACK_GOOD for r in records)
    err = max(pose_diff(poses[k], truth[k])[0] for k in range(nf))
    say("  all Good, max |t - t_true| %.2f mm" % (1000 * err))
    assert err < 2.5e-3  # 1.96 mm measured (FP16 + U16, the reference's degeneracy filter on)
    rows = list(csv.DictReader(open(tmp_path / "out" / "trace.csv")))
    track = [r for r in rows if r["kind"] == "track"]
    integ = [r for r in rows if r["kind"] == "integ"]
    assert len(track) == nf and len(integ) == nf
    assert [int(r["frame_id"]) for r in track] == list(range(nf))
    assert track[0]["quality"] == "-1" and all(r["quality"] == "0" for r in track[1:])
    for k, r in enumerate(track):
        assert abs(float(r["tx"]) - poses[k][0, 3]) < 1e-5
        assert int(r["inliers"]) > 10000 or k == 0
    assert abs(float(track[1]["t_ms"]) - 2000000 / 60000.0) < 1e-3
