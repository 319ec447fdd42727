"""Synthetic depth frames (libazu_scene.so): the analytic sphere / plane / box scene,
orbit poses and the 2x bicubic upsampler used by the frame-loop workloads.

Workload constants (sizes, seeds) live in WORKLOADS; every choice the configuration
text leaves open is fixed in azu_amd/csrc/scene.cpp's header comment.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "libazu_scene.so")
if not os.path.exists(_LIB):
    raise ImportError("libazu_scene.so is not built: run __graft_entry__.build() or azu_amd/build.py")
_lib = ctypes.CDLL(_LIB)
_lib.azu_scene_pose.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
_lib.azu_scene_pose_d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
_lib.azu_scene_render.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
_lib.azu_scene_upsample2x.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]

K640 = (525.0, 525.0, 319.5, 239.5)

# name -> resolution, voxel size, truncation, frames, boxes, jitter, image size
WORKLOADS = {
    # 256^3 @ 4 mm, the reference CPU configuration
    "c1_256": dict(res=256, vs=0.004, trunc=0.012, frames=30, boxes=0, jitter=0, W=640, H=480),
    # 512^3 @ 2 mm with 10 boxes and pose jitter (FP32 and FP16/U16 forms)
    "c2_512": dict(res=512, vs=0.002, trunc=0.006, frames=100, boxes=10, jitter=1, W=640, H=480),
    # 1024^3 @ 1 mm, 1280x960 from 2x bicubic upsampling, 300 frames
    "c5_1024": dict(res=1024, vs=0.001, trunc=0.003, frames=300, boxes=10, jitter=1, W=1280, H=960),
}
SEED = 20240601
ORIGIN = (-0.512, -0.512, 0.288)  # the 1.024 m cube centred on the sphere (0, 0, 0.8)
MAX_WEIGHT = 128.0
BAND = (0.3, 5.0)


def pose(frame, jitter=0, seed=SEED):
    """World-from-camera 4x4 (row-major float32) of a frame."""
    out = np.empty(16, np.float32)
    _lib.azu_scene_pose(int(frame), int(jitter), int(seed), out.ctypes.data)
    return out.reshape(4, 4).T.copy()


def pose64(frame, jitter=0, seed=SEED):
    out = np.empty(16, np.float64)
    _lib.azu_scene_pose_d(int(frame), int(jitter), int(seed), out.ctypes.data)
    return out.reshape(4, 4).T.copy()


def render(T, K=K640, W=640, H=480, boxes=0, seed=SEED, noise=True, frame=0, with_true=False):
    """uint16 millimetre depth (H, W) of the scene seen from pose T."""
    cm = np.ascontiguousarray(np.asarray(T, np.float32).T.reshape(16))
    Kc = np.asarray(K, np.float32)
    out = np.empty((H, W), np.uint16)
    true = np.empty((H, W), np.float64) if with_true else None
    _lib.azu_scene_render(int(boxes), int(seed), cm.ctypes.data, Kc.ctypes.data, W, H,
                          int(seed) * 1000003 + int(frame), 1 if noise else 0, out.ctypes.data,
                          true.ctypes.data if with_true else None)
    return (out, true) if with_true else out


def upsample2x(mm):
    mm = np.ascontiguousarray(mm, dtype=np.uint16)
    H, W = mm.shape
    out = np.empty((2 * H, 2 * W), np.uint16)
    _lib.azu_scene_upsample2x(mm.ctypes.data, W, H, out.ctypes.data)
    return out


def scale_intrinsics(K, rw, rh, w, h):
    """scaleIntrinsics (include/sensor/CameraIntrinsics.h:33-40), in float32."""
    K = np.asarray(K, np.float32)
    if w == rw and h == rh:
        return K.copy()
    sx = np.float32(w) / np.float32(rw)
    sy = np.float32(h) / np.float32(rh)
    return np.array([K[0] * sx, K[1] * sy, (K[2] + np.float32(0.5)) * sx - np.float32(0.5),
                     (K[3] + np.float32(0.5)) * sy - np.float32(0.5)], np.float32)


def sequence(name, frames=None, seed=SEED):
    """(frames_mm (F, H, W) uint16, poses (F, 4, 4), K float32[4]) of a workload."""
    w = WORKLOADS[name]
    nf = w["frames"] if frames is None else frames
    W0, H0 = 640, 480
    out = []
    poses = []
    for f in range(nf):
        T = pose(f, w["jitter"], seed)
        mm = render(T, K640, W0, H0, w["boxes"], seed, True, f)
        if w["W"] != W0:
            mm = upsample2x(mm)
        out.append(mm)
        poses.append(T)
    K = scale_intrinsics(K640, W0, H0, w["W"], w["H"])
    return np.stack(out), np.stack(poses), K


def volume_params(name):
    w = WORKLOADS[name]
    return dict(resolution=w["res"], voxel_size=w["vs"], truncation=w["trunc"], max_weight=MAX_WEIGHT,
                origin=ORIGIN, min_depth=BAND[0], max_depth=BAND[1])
