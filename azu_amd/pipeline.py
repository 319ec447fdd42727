"""The per-frame KinectFusion loop on one GPU, in the reference pipeline's order
(src/app/PipelineController.cpp: trackingLoopBody + integrationLoopBody):

  frame 0:   live pyramid; integrate at the initial pose; raycast the model there;
  frame k:   live pyramid; ICP against the model raycast at the previous tracked pose
             (reference pose = estimate = that pose); integrate and raycast at the
             tracked pose. The reference pipeline's constant-velocity estimate
             (PipelineController.cpp:1162-1187) is available as motion_model=
             "constant_velocity"; the benchmark loop does not use it.

The tracked pose never leaves the device inside the loop (the ICP state feeds
integrate and raycast through a device pointer); each frame's pose is copied to a
pinned host slot asynchronously, so frames are enqueued back to back with no host
synchronisation. Grading, retry and relocalisation are not part of the loop.
"""
import ctypes

import numpy as np

from . import native as N
from .fusion import ICPParams, ICPTracker, TSDFParams, TSDFVolume, pose_cm, pose_rm
from .native import check, lib


class FrameLoop:
    def __init__(self, tsdf_params, fmt, W, H, K, icp_params=None, mode=N.MODE_AUTO, stream=None, max_frames=4096,
                 motion_model="none"):
        """motion_model: "none" (the estimate is the previous tracked pose, the benchmark
        loop) or "constant_velocity" (the reference pipeline's cur * (last^-1 * cur))."""
        self.vol = TSDFVolume(tsdf_params if isinstance(tsdf_params, TSDFParams) else TSDFParams(**tsdf_params),
                              fmt=fmt, mode=mode)
        self.icp = icp_params or ICPParams(min_depth=self.vol.params.min_depth, max_depth=self.vol.params.max_depth)
        self.tr = ICPTracker(W, H, K, self.icp)
        self.W, self.H = W, H
        self.K = N.f32(K)
        self.stream = stream
        self.mv = N.DeviceBuffer(W * H * 12)
        self.mn = N.DeviceBuffer(W * H * 12)
        self.d_pose = self.tr.pose_device()
        self.d0 = self.tr.level(0)[0]
        self.max_frames = max_frames
        p = ctypes.c_void_p()
        check(lib.azu_host_alloc(ctypes.byref(p), 64 * max_frames))
        self._poses_host = p.value
        self.nframes = 0
        self.motion_model = motion_model

    def reset(self, pose0):
        self.vol.reset(self.stream)
        check(lib.azu_tracker_set_pose(self.tr.h, N.fptr(pose_cm(pose0)), self.stream))
        self.nframes = 0

    def frame(self, d_mm):
        """Enqueue one frame from device uint16 millimetres."""
        s = self.stream
        check(lib.azu_tracker_prepare_mm(self.tr.h, d_mm, s))
        if self.nframes > 0:
            # reference pose: the model's raycast pose (the previous tracked pose)
            est = self.d_pose
            if self.motion_model == "constant_velocity":
                pred = ctypes.c_void_p()
                check(lib.azu_tracker_predict(self.tr.h, ctypes.byref(pred), s))
                est = pred.value
            check(lib.azu_tracker_track_dpose(self.tr.h, self.mv.ptr, self.mn.ptr, self.W, self.H, est,
                                              self.d_pose, s))
        lo, hi = self.vol.params.min_depth, self.vol.params.max_depth
        check(lib.azu_integrate_dpose(self.vol.h, self.d0, self.W, self.H, N.fptr(self.K), self.d_pose, lo, hi, s))
        check(lib.azu_raycast_dpose(self.vol.h, self.d_pose, N.fptr(self.K), self.W, self.H, self.mv.ptr,
                                    self.mn.ptr, None, s))
        if self.nframes < self.max_frames:
            check(lib.azu_memcpy_d2h(self._poses_host + 64 * self.nframes, self.d_pose, 64, s))
        self.nframes += 1

    def poses(self):
        """Tracked poses (row-major 4x4) of the frames run since the last reset (synchronises)."""
        N.synchronize(self.stream)
        n = min(self.nframes, self.max_frames)
        buf = (ctypes.c_float * (16 * n)).from_address(self._poses_host)
        arr = np.frombuffer(buf, dtype=np.float32).reshape(n, 16)
        return np.stack([pose_rm(a) for a in arr]) if n else np.zeros((0, 4, 4), np.float32)

    def model(self):
        N.synchronize(self.stream)
        return (self.mv.download(np.float32, (self.H, self.W, 3)), self.mn.download(np.float32, (self.H, self.W, 3)))

    def close(self):
        if self._poses_host:
            lib.azu_host_free(self._poses_host)
            self._poses_host = None
        self.tr.close()
        self.vol.close()
