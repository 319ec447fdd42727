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
synchronisation. Grading and retry are GradedFrameLoop's (one host synchronisation per
track); relocalisation is not part of either loop.
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


class GradedFrameLoop(FrameLoop):
    """The frame loop with the reference pipeline's tracking decision
    (src/app/PipelineController.cpp:1160-1187, 1329-1406, 1500-1590), graded on the host
    by azu_track_grade (include/tracking/TrackingPolicy.h):

      estimate  the previous pose, or the constant-velocity prediction;
      track     against the model image with the model's raycast pose as reference; when
                the estimate was a prediction and the result grades Failed, track the
                same prepared live frame again from the previous pose (the retry);
      grade     keepObservableMotion (degeneracy_rel > 0), then classifyTracking;
      Good      continue from the graded pose, integrate, raycast the model there;
      Poor      continue from the graded pose; volume and model image unchanged;
      Failed    keep the previous pose; failures_before_lost consecutive failures set
                `lost` (relocalisation is not part of this loop).

    Each frame synchronises once per track to read the ICP result, as the reference's
    tracking thread does; FrameLoop is the unsynchronised loop without grading."""

    def __init__(self, *args, policy=None, degeneracy_rel=5e-3, **kw):
        super().__init__(*args, **kw)
        self.policy = policy or N.TrackingPolicy()
        if policy is None:
            check(lib.azu_tracking_policy_defaults(ctypes.byref(self.policy)))
        self.degeneracy_rel = float(degeneracy_rel)
        self.d_model_pose = N.DeviceBuffer(64)
        self.d_prev = N.DeviceBuffer(64)
        self.decisions = []
        self.consecutive_failures = 0
        self.lost = False
        self._prev = None

    def reset(self, pose0):
        super().reset(pose0)
        self._prev = (ctypes.c_float * 16)(*N.f32(pose_cm(pose0)).tolist())
        self.decisions = []
        self.consecutive_failures = 0
        self.lost = False

    def _track(self, d_est):
        check(lib.azu_tracker_track_dpose(self.tr.h, self.mv.ptr, self.mn.ptr, self.W, self.H, d_est,
                                          self.d_model_pose.ptr, self.stream))
        r = N.IcpResult()
        check(lib.azu_tracker_result(self.tr.h, ctypes.byref(r), self.stream))
        return r

    def _grade(self, r, rel):
        d = N.TrackDecision()
        check(lib.azu_track_grade(ctypes.byref(r), self._prev, ctypes.byref(self.policy), rel, ctypes.byref(d)))
        return d

    def frame(self, d_src, raw=False):
        """Run one frame from device uint16 millimetres (raw=True: raw Kinect codes);
        returns its decision record (quality, degenerate_dofs, retried, inliers,
        weakest_ratio, result: the graded IcpResult, predicted: the estimate's pose)."""
        s = self.stream
        check((lib.azu_tracker_prepare_raw if raw else lib.azu_tracker_prepare_mm)(self.tr.h, d_src, s))
        rec = dict(quality=N.TRACK_GOOD, degenerate_dofs=0, retried=False, inliers=0, weakest_ratio=None,
                   result=None, predicted=None)
        integrate = True
        if self.nframes > 0:
            check(lib.azu_memcpy_d2d(self.d_prev.ptr, self.d_pose, 64, s))
            est, predicted = self.d_pose, False
            if self.motion_model == "constant_velocity":
                pred = ctypes.c_void_p()
                check(lib.azu_tracker_predict(self.tr.h, ctypes.byref(pred), s))
                est, predicted = pred.value, True
            pose_est = np.empty(16, np.float32)
            check(lib.azu_memcpy_d2h(pose_est.ctypes.data, est, 64, s))
            r = self._track(est)
            rec["predicted"] = pose_rm(pose_est)
            if predicted and self._grade(r, 0.0).quality == N.TRACK_FAILED:
                r = self._track(self.d_prev.ptr)
                rec["retried"] = True
            d = self._grade(r, self.degeneracy_rel)
            check(lib.azu_tracker_set_tracked_pose(self.tr.h, d.pose, s))
            self._prev = (ctypes.c_float * 16)(*d.pose)
            ratio = ctypes.c_float()
            check(lib.azu_track_weakest_ratio(r.information, ctypes.byref(ratio)))
            r.pose[:] = d.pose
            rec.update(quality=d.quality, degenerate_dofs=d.degenerate_dofs, inliers=r.inliers,
                       weakest_ratio=ratio.value, result=r)
            integrate = bool(d.integrate)
            if d.quality == N.TRACK_FAILED:
                self.consecutive_failures += 1
                if self.consecutive_failures >= self.policy.failures_before_lost:
                    self.lost = True
            else:
                self.consecutive_failures = 0
        if integrate:
            lo, hi = self.vol.params.min_depth, self.vol.params.max_depth
            check(lib.azu_integrate_dpose(self.vol.h, self.d0, self.W, self.H, N.fptr(self.K), self.d_pose, lo, hi,
                                          s))
            check(lib.azu_raycast_dpose(self.vol.h, self.d_pose, N.fptr(self.K), self.W, self.H, self.mv.ptr,
                                        self.mn.ptr, None, s))
            check(lib.azu_memcpy_d2d(self.d_model_pose.ptr, self.d_pose, 64, s))
        if self.nframes < self.max_frames:
            check(lib.azu_memcpy_d2h(self._poses_host + 64 * self.nframes, self.d_pose, 64, s))
        self.nframes += 1
        self.decisions.append(rec)
        return rec

    def close(self):
        for b in (self.d_model_pose, self.d_prev):
            if b is not None:
                b.free()
        self.d_model_pose = self.d_prev = None
        super().close()
