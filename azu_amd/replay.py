"""Lockstep replay of a fakenect recording through the graded frame loop, with a
per-frame trace in the reference's FrameTrace CSV schema.

Reference: tools/azu_replay.cpp:1-40 (each depth frame is fully handled, graded and if
Good integrated and raycast, before the next), tools/FakenectRecording.h (INDEX.txt,
"d-<host s>-<ticks>-..." depth files: a one-line PGM header then 640x480 little-endian
uint16 raw 11-bit codes), include/app/FrameTrace.h:19-51 and src/app/FrameTrace.cpp:16-50
(columns, pose as translation + quaternion w x y z, 7 significant digits).

What the reference's replay adds and this one does not: RGB, the accelerometer and
gravity checks, relocalisation, keyframes, the mesh and summary files. Their trace
columns carry the reference's "not measured" values (tilt -1, relocalisation fields
0 / empty / -1, keyframes 0, fern_dissim -1)."""
import os
import time

import numpy as np

from . import native as N
from .pipeline import GradedFrameLoop

WIDTH, HEIGHT = 640, 480

TRACE_HEADER = ("kind,frame_id,t_ms,quality,relocalizing,queued,"
                "tx,ty,tz,qw,qx,qy,qz,ptx,pty,ptz,pqw,pqx,pqy,pqz,"
                "inliers,valid_live,projected,valid_model,dist_filtered,angle_filtered,"
                "rms_m,final_step,converged,model_frame_id,outside_volume,degenerate_dofs,"
                "ms_preprocess,ms_icp,ms_track,ms_integrate,ms_raycast,tilt_err_deg,"
                "reloc_solves,reloc_refines,reloc_source,reloc_reject,reloc_consistent,"
                "reloc_violation,reloc_coverage,reloc_eig,keyframes,fern_dissim\n")


def read_index(rec_dir):
    """Depth entries of <rec_dir>/INDEX.txt in file order: (host_s, ticks, file name)."""
    out = []
    with open(os.path.join(rec_dir, "INDEX.txt")) as f:
        for line in f:
            line = line.rstrip("\r\n")
            parts = line.split("-")
            if len(line) < 3 or len(parts) < 4 or parts[0] != "d":
                continue
            try:
                out.append((float(parts[1]), int(parts[2]), line))
            except ValueError:
                continue
    return out


def read_depth(path):
    """Raw 11-bit codes (480, 640) uint16 of one fakenect depth file."""
    with open(path, "rb") as f:
        f.readline()
        data = f.read(WIDTH * HEIGHT * 2)
    if len(data) != WIDTH * HEIGHT * 2:
        raise ValueError("short depth file: %s" % path)
    return np.frombuffer(data, dtype="<u2").reshape(HEIGHT, WIDTH).copy()


def write_depth(path, raw):
    """Writes raw codes as a fakenect depth file (the format read_depth reads)."""
    raw = np.ascontiguousarray(raw, dtype="<u2")
    with open(path, "wb") as f:
        f.write(b"P5 640 480 65535\n")
        f.write(raw.tobytes())


def _quat(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    R = np.asarray(R, np.float64)
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = (0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s)
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = ((R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s)
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = ((R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s)
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = ((R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s)
    return q


def _g(x):
    return "%.7g" % x


def _pose_fields(T):
    T = np.asarray(T, np.float64)
    return [_g(T[0, 3]), _g(T[1, 3]), _g(T[2, 3])] + [_g(v) for v in _quat(T[:3, :3])]


class FrameTraceWriter:
    """FrameTrace CSV (src/app/FrameTrace.cpp:16-50): one "track" row per graded frame,
    one "integ" row per integrated frame."""

    def __init__(self, path):
        self.f = open(path, "w")
        self.f.write(TRACE_HEADER)

    def row(self, kind, frame_id, t_ms, quality, queued, pose, predicted, r, model_frame_id, dofs, ms_track,
            ms_integ):
        f = [kind, str(frame_id), _g(t_ms), str(quality), "0", str(int(queued))]
        f += _pose_fields(pose) + _pose_fields(predicted)
        if r is not None:
            rms = float(np.sqrt(max(0.0, r.error)))
            f += [str(r.inliers), str(r.valid_live_points), str(r.projected_points), str(r.valid_model_points),
                  str(r.dist_filtered), str(r.angle_filtered), _g(rms), _g(r.final_step), str(int(r.converged))]
        else:
            f += ["0"] * 6 + ["0", "0", "0"]
        f += [str(model_frame_id), "0", str(dofs), "0", "0", _g(ms_track), _g(ms_integ), "0", "-1",
              "0", "0", "", "", "-1", "-1", "-1", "-1", "0", "-1"]
        self.f.write(",".join(f) + "\n")

    def close(self):
        self.f.close()


def replay(rec_dir, out_dir, volume_params, K, pose0=None, fmt=N.FMT_F16_U16, max_frames=None,
           motion_model="constant_velocity", degeneracy_rel=5e-3):
    """Replays the recording's depth frames in order through GradedFrameLoop (raw codes
    in), writing <out_dir>/trace.csv. degeneracy_rel is keepObservableMotion's ratio (the
    pipeline's default 5e-3; 0 turns the filter off). Returns the per-frame decision
    records and the tracked poses (row-major)."""
    os.makedirs(out_dir, exist_ok=True)
    entries = read_index(rec_dir)
    if max_frames is not None:
        entries = entries[:max_frames]
    loop = GradedFrameLoop(volume_params, fmt, WIDTH, HEIGHT, K, motion_model=motion_model,
                           max_frames=max(1, len(entries)), degeneracy_rel=degeneracy_rel)
    loop.reset(np.eye(4, dtype=np.float32) if pose0 is None else pose0)
    trace = FrameTraceWriter(os.path.join(out_dir, "trace.csv"))
    buf = N.DeviceBuffer(WIDTH * HEIGHT * 2)
    records = []
    model_frame = 0
    t0_ticks, wraps, last = None, 0, None
    try:
        for k, (_, ticks, name) in enumerate(entries):
            raw = read_depth(os.path.join(rec_dir, name))
            if last is not None and ticks < last:
                wraps += 1
            last = ticks
            if t0_ticks is None:
                t0_ticks = ticks
            t_ms = ((ticks + wraps * (1 << 32)) - t0_ticks) / 60000.0  # 60 MHz device clock
            buf.upload(raw)
            t = time.perf_counter()
            rec = loop.frame(buf.ptr, raw=True)
            N.synchronize(loop.stream)
            ms = 1e3 * (time.perf_counter() - t)
            pose = loop.poses()[-1]
            queued = rec["quality"] == N.TRACK_GOOD
            pred = rec["predicted"] if rec["predicted"] is not None else pose
            trace.row("track", k, t_ms, rec["quality"] if k > 0 else -1, queued, pose, pred, rec["result"],
                      model_frame, rec["degenerate_dofs"], ms, 0.0)
            if queued:
                trace.row("integ", k, t_ms, rec["quality"] if k > 0 else -1, True, pose, pred, rec["result"], k,
                          rec["degenerate_dofs"], 0.0, ms)
                model_frame = k
            records.append(rec)
        poses = loop.poses()
    finally:
        trace.close()
        buf.free()
        loop.close()
    return records, poses
