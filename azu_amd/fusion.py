"""Python mirror of the reference's TSDFVolume / ICPTracker interfaces over the C ABI.

Method names and argument meaning follow include/tsdf/TSDFVolume.h:69-190 and
include/tracking/ICPTracker.h:87-134 of the reference; poses are 4x4 world-from-camera
matrices (numpy, row-major here, converted to the ABI's column-major float[16]).
Everything runs on the GPU through libazu_hip.so.
"""
import ctypes

import numpy as np

from . import native as N
from .native import check, lib

FRAME_W, FRAME_H = 640, 480
KINECT_K = (525.0, 525.0, 319.5, 239.5)


def pose_cm(T):
    """Row-major 4x4 -> column-major float32[16] (Eigen storage)."""
    return np.ascontiguousarray(np.asarray(T, dtype=np.float32).T.reshape(16))


def pose_rm(cm16):
    return np.asarray(cm16, dtype=np.float32).reshape(4, 4).T.copy()


class TSDFParams:
    """kfusion::tsdf::TSDFParams (include/tsdf/TSDFVolume.h:39-53)."""

    def __init__(self, resolution=256, voxel_size=0.010, truncation=0.030, max_weight=128.0,
                 origin=(-1.28, -1.28, 0.0), min_depth=0.30, max_depth=5.00):
        self.resolution = int(resolution)
        self.voxel_size = float(voxel_size)
        self.truncation = float(truncation)
        self.max_weight = float(max_weight)
        self.origin = tuple(float(o) for o in origin)
        self.min_depth = float(min_depth)
        self.max_depth = float(max_depth)

    def c(self):
        p = N.TsdfParams()
        p.resolution = self.resolution
        p.voxel_size = self.voxel_size
        p.truncation = self.truncation
        p.max_weight = self.max_weight
        for i in range(3):
            p.origin[i] = self.origin[i]
        p.min_depth = self.min_depth
        p.max_depth = self.max_depth
        return p


class TSDFVolume:
    """GPU TSDF volume (TSDFVolume.h:69-190) in F32 or F16_U16 storage."""

    def __init__(self, params=None, fmt=N.FMT_F32, mode=N.MODE_AUTO):
        self.params = params or TSDFParams()
        self.fmt = fmt
        h = ctypes.c_void_p()
        cp = self.params.c()
        check(lib.azu_vol_create(ctypes.byref(cp), fmt, ctypes.byref(h)))
        self.h = h.value
        if mode != N.MODE_AUTO:
            self.set_mode(mode)

    def set_mode(self, mode):
        check(lib.azu_vol_set_mode(self.h, mode))

    def set_skip(self, on):
        """Empty-space skipping in the raycast (on by default; same results either way)."""
        check(lib.azu_vol_set_skip(self.h, 1 if on else 0))

    def setParams(self, params):
        cp = params.c()
        check(lib.azu_vol_set_params(self.h, ctypes.byref(cp)))
        self.params = params

    def division_check(self):
        """(voxel_size, truncation, reciprocal): True where the exact kernels use the
        verified fast form instead of the IEEE division."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib.azu_vol_division_check(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return bool(a.value), bool(b.value), bool(c.value)

    def reset(self, stream=None):
        check(lib.azu_vol_reset(self.h, stream))

    @property
    def res(self):
        return self.params.resolution

    def voxelData(self):
        """(tsdf, weight) float32 arrays of shape (res, res, res) indexed [z, y, x]."""
        n = self.res ** 3
        t = np.empty(n, np.float32)
        w = np.empty(n, np.float32)
        check(lib.azu_vol_download(self.h, t.ctypes.data, w.ctypes.data))
        r = self.res
        return t.reshape(r, r, r), w.reshape(r, r, r)

    def upload(self, tsdf, weight):
        t = N.f32(tsdf).reshape(-1)
        w = N.f32(weight).reshape(-1)
        assert t.size == self.res ** 3 and w.size == t.size
        check(lib.azu_vol_upload(self.h, t.ctypes.data, w.ctypes.data))

    def counters(self, reset=False):
        u = ctypes.c_ulonglong()
        n = ctypes.c_ulonglong()
        check(lib.azu_vol_counters(self.h, ctypes.byref(u), ctypes.byref(n), 1 if reset else 0))
        return u.value, n.value

    def integrate_device(self, d_depth_m, K, pose, W, H, min_depth=0.3, max_depth=5.0, stream=None):
        Kc = N.f32(K)
        check(lib.azu_integrate(self.h, d_depth_m, W, H, N.fptr(Kc), N.fptr(pose_cm(pose)), min_depth, max_depth,
                                stream))

    def integrate(self, depth_meters, K, pose, min_depth=0.3, max_depth=5.0):
        """TSDFVolume::integrate with a host float metre frame (H, W)."""
        d = N.f32(depth_meters)
        H, W = d.shape
        buf = N.device_array(d)
        self.integrate_device(buf.ptr, K, pose, W, H, min_depth, max_depth)
        N.synchronize()
        buf.free()

    def raycast(self, pose, K, W=FRAME_W, H=FRAME_H, with_t=False):
        """TSDFVolume::raycast -> (vertices, normals) float32 (H, W, 3) (+ t_hit (H, W))."""
        dv = N.DeviceBuffer(W * H * 12)
        dn = N.DeviceBuffer(W * H * 12)
        Kc = N.f32(K)
        if with_t:
            dt = N.DeviceBuffer(W * H * 4)
            dp = N.device_array(pose_cm(pose))
            check(lib.azu_raycast_dpose(self.h, dp.ptr, N.fptr(Kc), W, H, dv.ptr, dn.ptr, dt.ptr, None))
        else:
            check(lib.azu_raycast(self.h, N.fptr(pose_cm(pose)), N.fptr(Kc), W, H, dv.ptr, dn.ptr, None))
        v = dv.download(np.float32, (H, W, 3))
        n = dn.download(np.float32, (H, W, 3))
        if with_t:
            return v, n, dt.download(np.float32, (H, W))
        return v, n

    def marching_cubes(self, max_triangles=2000000):
        """MarchingCubes::extract on the device -> (positions (V, 3), normals (V, 3),
        indices (3T,) uint32, truncated)."""
        m = ctypes.c_void_p()
        check(lib.azu_mc_extract(self.h, max_triangles, ctypes.byref(m), None))
        try:
            nv = ctypes.c_ulonglong()
            nt = ctypes.c_ulonglong()
            tr = ctypes.c_int()
            check(lib.azu_mesh_info(m.value, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(tr)))
            pos = np.empty((nv.value, 3), np.float32)
            nrm = np.empty((nv.value, 3), np.float32)
            idx = np.empty(3 * nt.value, np.uint32)
            check(lib.azu_mesh_download(m.value, pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data))
        finally:
            lib.azu_mesh_destroy(m.value)
        return pos, nrm, idx, bool(tr.value)

    def point_cloud(self):
        """TSDFVolume::extractGlobalPointCloud on the device -> points (P, 3) float32: the
        voxels with weight > 1 and |tsdf| < 0.2 at origin + (x, y, z) * voxel_size, in
        linear voxel order (x fastest)."""
        n = ctypes.c_ulonglong()
        check(lib.azu_vol_point_cloud(self.h, None, 0, ctypes.byref(n), None))
        if n.value == 0:
            return np.zeros((0, 3), np.float32)
        buf = N.DeviceBuffer(n.value * 12)
        try:
            m = ctypes.c_ulonglong()
            check(lib.azu_vol_point_cloud(self.h, buf.ptr, n.value, ctypes.byref(m), None))
            assert m.value == n.value
            return buf.download(np.float32, (n.value, 3))
        finally:
            buf.free()

    def storage(self):
        p = ctypes.c_void_p()
        b = ctypes.c_size_t()
        check(lib.azu_vol_storage(self.h, ctypes.byref(p), ctypes.byref(b)))
        return p.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib.azu_vol_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ICPParams:
    """kfusion::tracking::ICPParams (include/tracking/ICPTracker.h:19-25)."""

    def __init__(self, max_iterations=(10, 5, 4), dist_threshold=0.1, angle_threshold=30.0, min_depth=0.3,
                 max_depth=5.0):
        self.max_iterations = tuple(int(i) for i in max_iterations)
        self.dist_threshold = float(dist_threshold)
        self.angle_threshold = float(angle_threshold)
        self.min_depth = float(min_depth)
        self.max_depth = float(max_depth)

    def c(self):
        p = N.IcpParams()
        for i in range(3):
            p.max_iterations[i] = self.max_iterations[i]
        p.dist_threshold = self.dist_threshold
        p.angle_threshold = self.angle_threshold
        p.min_depth = self.min_depth
        p.max_depth = self.max_depth
        return p


class ICPTracker:
    """GPU ICP tracker (ICPTracker.h:87-134): live pyramid + 3-level point-to-plane ICP."""

    def __init__(self, W=FRAME_W, H=FRAME_H, K=KINECT_K, params=None):
        self.W, self.H = W, H
        self.K = N.f32(K)
        self.params = params or ICPParams()
        h = ctypes.c_void_p()
        cp = self.params.c()
        check(lib.azu_tracker_create(W, H, N.fptr(self.K), ctypes.byref(cp), ctypes.byref(h)))
        self.h = h.value

    def setParams(self, params):
        cp = params.c()
        check(lib.azu_tracker_set_params(self.h, ctypes.byref(cp)))
        self.params = params

    def prepare_mm_device(self, d_mm, stream=None):
        check(lib.azu_tracker_prepare_mm(self.h, d_mm, stream))

    def prepare_raw_device(self, d_raw, stream=None):
        """Live pyramid from device uint16 raw Kinect codes (azu_tracker_prepare_raw)."""
        check(lib.azu_tracker_prepare_raw(self.h, d_raw, stream))

    def prepare_m_device(self, d_m, stream=None):
        check(lib.azu_tracker_prepare_m(self.h, d_m, stream))

    def load_level(self, l, vertices, normals, depth=None):
        """Load one live pyramid level from host arrays (the reference's FramePyramid)."""
        bufs = [N.device_array(N.f32(vertices)), N.device_array(N.f32(normals))]
        dd = N.device_array(N.f32(depth)) if depth is not None else None
        check(lib.azu_tracker_load_level(self.h, l, dd.ptr if dd else None, bufs[0].ptr, bufs[1].ptr, None))
        N.synchronize()

    def downsample(self):
        check(lib.azu_tracker_downsample(self.h, None))

    def level(self, l):
        d = ctypes.c_void_p()
        v = ctypes.c_void_p()
        n = ctypes.c_void_p()
        W = ctypes.c_int()
        H = ctypes.c_int()
        check(lib.azu_tracker_level(self.h, l, ctypes.byref(d), ctypes.byref(v), ctypes.byref(n), ctypes.byref(W),
                                    ctypes.byref(H)))
        return d.value, v.value, n.value, W.value, H.value

    def level_host(self, l):
        d, v, n, W, H = self.level(l)
        N.synchronize()
        out = []
        for ptr, shape in ((d, (H, W)), (v, (H, W, 3)), (n, (H, W, 3))):
            a = np.empty(shape, np.float32)
            check(lib.azu_memcpy_d2h(a.ctypes.data, ptr, a.nbytes, None))
            out.append(a)
        N.synchronize()
        return out

    def track_device(self, d_mv, d_mn, MW, MH, est, ref, stream=None):
        check(lib.azu_tracker_track(self.h, d_mv, d_mn, MW, MH, N.fptr(pose_cm(est)), N.fptr(pose_cm(ref)),
                                    stream))

    def result(self, stream=None):
        r = N.IcpResult()
        check(lib.azu_tracker_result(self.h, ctypes.byref(r), stream))
        return {
            "pose": pose_rm(np.frombuffer(r.pose, dtype=np.float32)),
            "error": r.error, "inliers": r.inliers, "converged": bool(r.converged),
            "tracking_ok": bool(r.tracking_ok), "final_step": r.final_step,
            "information": np.frombuffer(r.information, dtype=np.float32).reshape(6, 6).copy(),
            "valid_live_points": r.valid_live_points, "valid_model_points": r.valid_model_points,
            "projected_points": r.projected_points, "dist_filtered": r.dist_filtered,
            "angle_filtered": r.angle_filtered, "iterations": r.iterations,
        }

    def track(self, model_v, model_n, est, ref):
        """ICPTracker::track against a host model image (MH, MW, 3) after a prepare call."""
        mv = N.device_array(N.f32(model_v))
        mn = N.device_array(N.f32(model_n))
        MH, MW = model_v.shape[:2]
        self.track_device(mv.ptr, mn.ptr, MW, MH, est, ref)
        return self.result()

    def pose_device(self):
        p = ctypes.c_void_p()
        check(lib.azu_tracker_pose_device(self.h, ctypes.byref(p)))
        return p.value

    def close(self):
        if getattr(self, "h", None):
            lib.azu_tracker_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
