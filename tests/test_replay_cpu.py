"""Replay harness host parts (azu_amd/replay.py): fakenect index and depth files, the
FrameTrace CSV schema of the reference (src/app/FrameTrace.cpp:23-29), pose quaternions."""
import numpy as np
from scipy.spatial.transform import Rotation

from azu_amd import replay as rp


def test_fakenect_files_round_trip(tmp_path):
    raw = (np.arange(480 * 640, dtype=np.uint32) % 2048).astype(np.uint16).reshape(480, 640)
    names = ["d-1700000000.250000-12345-0.pgm", "r-1700000000.260000-12345-0.ppm", "garbage",
             "a-1700000000.270000-12345-0.dump", "d-1700000000.283333-2012345-1.pgm"]
    (tmp_path / "INDEX.txt").write_text("\r\n".join(names) + "\n")
    for n in (names[0], names[4]):
        rp.write_depth(str(tmp_path / n), raw)
    idx = rp.read_index(str(tmp_path))
    assert [e[2] for e in idx] == [names[0], names[4]]
    assert idx[1][1] == 2012345 and abs(idx[0][0] - 1700000000.25) < 1e-6
    assert np.array_equal(rp.read_depth(str(tmp_path / names[0])), raw)
    assert (tmp_path / names[0]).read_bytes()[:17] == b"P5 640 480 65535\n"


def test_trace_header_and_quaternion():
    cols = rp.TRACE_HEADER.strip().split(",")
    assert len(cols) == 48 and cols[:6] == ["kind", "frame_id", "t_ms", "quality", "relocalizing", "queued"]
    assert cols[-2:] == ["keyframes", "fern_dissim"]
    rng = np.random.default_rng(1)
    for _ in range(200):
        R = Rotation.from_rotvec(rng.normal(size=3) * rng.uniform(0, 3.1)).as_matrix()
        w, x, y, z = rp._quat(R)
        q = Rotation.from_quat([x, y, z, w])
        assert np.allclose(q.as_matrix(), R, atol=1e-12)
        assert abs(w * w + x * x + y * y + z * z - 1) < 1e-12
