"""ctypes binding of the C ABI in include/azu_hip.h (libazu_hip.so, built in-tree).

Importing this module loads the gfx950 library and fails loudly when it is missing:
there is no CPU fallback for any entry point.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libazu_hip.so")

AZU_OK, AZU_ERR_ARG, AZU_ERR_HIP, AZU_ERR_ALLOC, AZU_ERR_STATE = 0, 1, 2, 3, 4
FMT_F32, FMT_F16_U16 = 0, 1
MODE_AUTO, MODE_EXACT, MODE_FAST = 0, 1, 2

c_float_p = ctypes.POINTER(ctypes.c_float)
c_void_p = ctypes.c_void_p


class TsdfParams(ctypes.Structure):
    _fields_ = [
        ("resolution", ctypes.c_int),
        ("voxel_size", ctypes.c_float),
        ("truncation", ctypes.c_float),
        ("max_weight", ctypes.c_float),
        ("origin", ctypes.c_float * 3),
        ("min_depth", ctypes.c_float),
        ("max_depth", ctypes.c_float),
    ]


class IcpParams(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int * 3),
        ("dist_threshold", ctypes.c_float),
        ("angle_threshold", ctypes.c_float),
        ("min_depth", ctypes.c_float),
        ("max_depth", ctypes.c_float),
    ]


class IcpResult(ctypes.Structure):
    _fields_ = [
        ("pose", ctypes.c_float * 16),
        ("error", ctypes.c_float),
        ("inliers", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("tracking_ok", ctypes.c_int),
        ("final_step", ctypes.c_float),
        ("information", ctypes.c_float * 36),
        ("valid_live_points", ctypes.c_int),
        ("valid_model_points", ctypes.c_int),
        ("projected_points", ctypes.c_int),
        ("dist_filtered", ctypes.c_int),
        ("angle_filtered", ctypes.c_int),
        ("iterations", ctypes.c_int),
    ]


class TrackingPolicy(ctypes.Structure):
    _fields_ = [
        ("min_inliers", ctypes.c_int),
        ("min_good_inliers", ctypes.c_int),
        ("min_fit_ratio", ctypes.c_float),
        ("max_rms_m", ctypes.c_float),
        ("max_frame_translation", ctypes.c_float),
        ("max_frame_rotation", ctypes.c_float),
        ("failures_before_lost", ctypes.c_int),
    ]


class TrackDecision(ctypes.Structure):
    _fields_ = [
        ("quality", ctypes.c_int),
        ("degenerate_dofs", ctypes.c_int),
        ("integrate", ctypes.c_int),
        ("pose", ctypes.c_float * 16),
    ]


TRACK_GOOD, TRACK_POOR, TRACK_FAILED = 0, 1, 2


class AzuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("azu status %d: %s" % (code, msg))
        self.code = code


# Every exported entry point of include/azu_hip.h with its ctypes signature.
_SIGS = {
    "azu_last_error": (ctypes.c_char_p, []),
    "azu_abi_version": (ctypes.c_int, []),
    "azu_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "azu_malloc": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_size_t]),
    "azu_free": (ctypes.c_int, [c_void_p]),
    "azu_host_alloc": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_size_t]),
    "azu_host_free": (ctypes.c_int, [c_void_p]),
    "azu_memcpy_h2d": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "azu_memcpy_d2h": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "azu_memcpy_d2d": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "azu_stream_create": (ctypes.c_int, [ctypes.POINTER(c_void_p)]),
    "azu_stream_destroy": (ctypes.c_int, [c_void_p]),
    "azu_stream_sync": (ctypes.c_int, [c_void_p]),
    "azu_vol_create": (ctypes.c_int, [ctypes.POINTER(TsdfParams), ctypes.c_int, ctypes.POINTER(c_void_p)]),
    "azu_vol_destroy": (ctypes.c_int, [c_void_p]),
    "azu_vol_reset": (ctypes.c_int, [c_void_p, c_void_p]),
    "azu_vol_set_params": (ctypes.c_int, [c_void_p, ctypes.POINTER(TsdfParams)]),
    "azu_vol_get_params": (ctypes.c_int, [c_void_p, ctypes.POINTER(TsdfParams), ctypes.POINTER(ctypes.c_int)]),
    "azu_vol_set_mode": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "azu_vol_set_skip": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "azu_vol_division_check": (ctypes.c_int, [c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "azu_vol_download": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_vol_upload": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_vol_read_voxels": (ctypes.c_int, [c_void_p, ctypes.c_size_t, ctypes.c_size_t, c_void_p, c_void_p]),
    "azu_vol_write_voxels": (ctypes.c_int, [c_void_p, ctypes.c_size_t, ctypes.c_size_t, c_void_p, c_void_p]),
    "azu_vol_usage": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]),
    "azu_vol_storage": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(ctypes.c_size_t)]),
    "azu_vol_counters": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_ulonglong),
                                        ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]),
    "azu_vol_integrated_frames": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "azu_integrate": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p,
                                     ctypes.c_float, ctypes.c_float, c_void_p]),
    "azu_integrate_noalloc": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p,
                                             ctypes.c_float, ctypes.c_float, c_void_p]),
    "azu_integrate_dpose": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_float_p, c_void_p,
                                           ctypes.c_float, ctypes.c_float, c_void_p]),
    "azu_raycast": (ctypes.c_int, [c_void_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p,
                                   c_void_p]),
    "azu_raycast_dpose": (ctypes.c_int, [c_void_p, c_void_p, c_float_p, ctypes.c_int, ctypes.c_int, c_void_p,
                                         c_void_p, c_void_p, c_void_p]),
    "azu_raycast_depth_loss": (ctypes.c_int, [c_void_p, c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, c_void_p,
                                              ctypes.c_float, ctypes.c_int, c_void_p, c_void_p,
                                              ctypes.POINTER(ctypes.c_double), c_void_p]),
    "azu_mc_extract": (ctypes.c_int, [c_void_p, ctypes.c_ulonglong, ctypes.POINTER(c_void_p), c_void_p]),
    "azu_mesh_info": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong),
                                     ctypes.POINTER(ctypes.c_int)]),
    "azu_mesh_download": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "azu_mesh_device": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                       ctypes.POINTER(c_void_p)]),
    "azu_mesh_destroy": (ctypes.c_int, [c_void_p]),
    "azu_tracker_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_float_p, ctypes.POINTER(IcpParams),
                                          ctypes.POINTER(c_void_p)]),
    "azu_tracker_destroy": (ctypes.c_int, [c_void_p]),
    "azu_tracker_set_params": (ctypes.c_int, [c_void_p, ctypes.POINTER(IcpParams)]),
    "azu_tracker_get_params": (ctypes.c_int, [c_void_p, ctypes.POINTER(IcpParams)]),
    "azu_icp_params_sanitize": (ctypes.c_int, [ctypes.POINTER(IcpParams), ctypes.POINTER(IcpParams)]),
    "azu_icp_result_default": (ctypes.c_int, [ctypes.POINTER(IcpResult)]),
    "azu_tracker_prepare_mm": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_tracker_prepare_raw": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_cond_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(c_void_p)]),
    "azu_cond_destroy": (ctypes.c_int, [c_void_p]),
    "azu_cond_reset_ema": (ctypes.c_int, [c_void_p, c_void_p]),
    "azu_cond_set_minimal_depth": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "azu_cond_process": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                                        c_void_p]),
    "azu_cond_run_stages": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_int, c_void_p]),
    "azu_cond_read_ema": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_cond_write_ema": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_tracker_prepare_m": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "azu_tracker_load_level": (ctypes.c_int, [c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "azu_tracker_downsample": (ctypes.c_int, [c_void_p, c_void_p]),
    "azu_tracker_level": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                         ctypes.POINTER(c_void_p), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "azu_tracker_track": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_float_p,
                                         c_float_p, c_void_p]),
    "azu_tracker_track_dpose": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_void_p,
                                               c_void_p, c_void_p]),
    "azu_tracker_set_pose": (ctypes.c_int, [c_void_p, c_float_p, c_void_p]),
    "azu_tracker_predict": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p), c_void_p]),
    "azu_tracker_set_tracked_pose": (ctypes.c_int, [c_void_p, c_float_p, c_void_p]),
    "azu_vol_point_cloud": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_ulonglong,
                                           ctypes.POINTER(ctypes.c_ulonglong), c_void_p]),
    "azu_tracking_policy_defaults": (ctypes.c_int, [ctypes.POINTER(TrackingPolicy)]),
    "azu_track_grade": (ctypes.c_int, [ctypes.POINTER(IcpResult), c_float_p, ctypes.POINTER(TrackingPolicy),
                                       ctypes.c_float, ctypes.POINTER(TrackDecision)]),
    "azu_track_weakest_ratio": (ctypes.c_int, [c_float_p, ctypes.POINTER(ctypes.c_float)]),
    "azu_tracker_pose_device": (ctypes.c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "azu_tracker_result": (ctypes.c_int, [c_void_p, ctypes.POINTER(IcpResult), c_void_p]),
    "azu_tracker_phase_times": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "azu_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "azu_profile_reset": (ctypes.c_int, []),
    "azu_profile_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
}

# Entry points added on top of the base set (declared in include/azu_hip.h as well).
_EXTRA_SIGS = {}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libazu_hip.so is not built (%s): run __graft_entry__.build() or azu_amd/build.py"
                          % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in list(_SIGS.items()) + list(_EXTRA_SIGS.items()):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc != AZU_OK:
        raise AzuError(rc, lib.azu_last_error().decode(errors="replace"))
    return rc


def fptr(a):
    """float* of a C-contiguous float32 numpy array."""
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_float_p)


def f32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


class DeviceBuffer:
    """A device allocation (azu_malloc) with blocking host copies."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = c_void_p()
        check(lib.azu_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        check(lib.azu_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes, stream))
        check(lib.azu_stream_sync(stream))
        return self

    def download(self, dtype, shape, stream=None):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib.azu_stream_sync(stream))
        check(lib.azu_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream))
        check(lib.azu_stream_sync(stream))
        return out

    def free(self):
        if self.ptr:
            lib.azu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_array(arr, stream=None):
    arr = np.ascontiguousarray(arr)
    return DeviceBuffer(arr.nbytes).upload(arr, stream)


def synchronize(stream=None):
    check(lib.azu_stream_sync(stream))


def device_info():
    dev = ctypes.c_int()
    name = ctypes.create_string_buffer(256)
    hbm = ctypes.c_size_t()
    check(lib.azu_device_info(ctypes.byref(dev), name, 256, ctypes.byref(hbm)))
    return dev.value, name.value.decode(errors="replace"), hbm.value


def profile_enable(on=True):
    check(lib.azu_profile_enable(1 if on else 0))


def profile_reset():
    check(lib.azu_profile_reset())


def profile_read():
    n_max = 64
    names = ctypes.create_string_buffer(48 * n_max)
    launches = (ctypes.c_longlong * n_max)()
    total = (ctypes.c_double * n_max)()
    n = ctypes.c_int()
    check(lib.azu_profile_read(names, n_max, launches, total, ctypes.byref(n)))
    out = {}
    for i in range(n.value):
        nm = names.raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (launches[i], total[i])
    return out
