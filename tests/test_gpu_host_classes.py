"""The C++ host classes (include/tsdf/TSDFVolume.h, include/tracking/ICPTracker.h,
include/meshing/MarchingCubes.h) run on the GPU by tests/cpp/host_classes.cpp: a frame
loop in the reference pipeline's order compared bit for bit with the C ABI path
(azu_amd.pipeline.FrameLoop through ctypes), the reference's volume reset contract
(tests/tsdf_reset_contract.cpp), its ICP weighting fixture through the host
track(FramePyramid, ModelFrame, ...) entry, and host-mirror voxel writes."""
import os
import struct
import subprocess

import numpy as np
import pytest

from azu_amd import native as N
from azu_amd import scene as sc
from azu_amd.fusion import pose_cm
from azu_amd.pipeline import FrameLoop
from tests.gpu_util import frames, small_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "host_classes")


def run(*args):
    assert os.path.exists(EXE), "tests/cpp/bin/host_classes is not built: run __graft_entry__.build()"
    print("host_classes", *args, flush=True)
    p = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.parametrize("mode", ["reset", "track", "voxels"])
def test_cpp_contracts(mode):
    rc, out = run(mode)
    assert rc == 0 and "PASS" in out, out


@pytest.mark.parametrize("fmt", [N.FMT_F32, N.FMT_F16_U16], ids=["f32", "f16"])
def test_cpp_frame_loop_matches_abi(tmp_path, fmt):
    n = 6
    p = small_params(128, 0.008)
    fr = frames(n)
    mm = np.stack([f[0] for f in fr]).astype(np.uint16)
    pose0 = np.asarray(fr[0][1], np.float32)
    H, W = mm.shape[1:]
    hdr = struct.pack("5i", W, H, n, p["resolution"], 1 if fmt == N.FMT_F16_U16 else 0)
    hdr += struct.pack("5f", p["voxel_size"], p["truncation"], *p["origin"])
    hdr += struct.pack("4f", *sc.K640) + pose_cm(pose0).tobytes()
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    src.write_bytes(hdr + mm.tobytes())
    rc, out = run("loop", src, dst)
    assert rc == 0, out
    raw = dst.read_bytes()
    r3 = p["resolution"] ** 3
    poses_cpp = np.frombuffer(raw, np.float32, 16 * n).reshape(n, 4, 4).transpose(0, 2, 1)
    off = 64 * n
    t_cpp = np.frombuffer(raw, np.float32, r3, off)
    w_cpp = np.frombuffer(raw, np.float32, r3, off + 4 * r3)
    nv, nt, valid, progress = np.frombuffer(raw, np.int64, 4, off + 8 * r3)
    off2 = off + 8 * r3 + 32
    pos_cpp = np.frombuffer(raw, np.float32, 3 * nv, off2).reshape(nv, 3)
    idx_cpp = np.frombuffer(raw, np.uint32, 3 * nt, off2 + 12 * nv)
    assert valid == 1 and progress == 1 and nt > 1000

    # the same frames through the C ABI (device poses, no host round trip)
    loop = FrameLoop(p, fmt, W, H, sc.K640)
    loop.reset(pose0)
    dev = [N.device_array(f) for f in mm]
    for d in dev:
        loop.frame(d.ptr)
    poses_abi = loop.poses()
    t_abi, w_abi = loop.vol.voxelData()
    pos, _, idx, _ = loop.vol.marching_cubes()
    loop.close()
    assert np.array_equal(poses_cpp.view(np.uint32), poses_abi.view(np.uint32))
    assert np.array_equal(t_cpp.view(np.uint32), t_abi.reshape(-1).view(np.uint32))
    assert np.array_equal(w_cpp, w_abi.reshape(-1))
    assert np.array_equal(pos_cpp.view(np.uint32), pos.view(np.uint32)) and np.array_equal(idx_cpp, idx)
