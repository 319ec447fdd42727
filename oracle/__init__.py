"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker or the timed CPU baseline; the product (azu_amd) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("oracle/liboracle.so is not built: run make -C oracle")
lib = ctypes.CDLL(LIB_PATH)
P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
LL = ctypes.c_longlong

for sfx, T in (("f32", F), ("f64", D)):
    getattr(lib, "orc_vol_create_" + sfx).restype = P
    getattr(lib, "orc_vol_create_" + sfx).argtypes = [I, F, F, F, P, F, F]
    getattr(lib, "orc_vol_destroy_" + sfx).argtypes = [P]
    getattr(lib, "orc_vol_reset_" + sfx).argtypes = [P]
    getattr(lib, "orc_vol_get_" + sfx).argtypes = [P, P, P]
    getattr(lib, "orc_vol_set_" + sfx).argtypes = [P, P, P]
    getattr(lib, "orc_depth_mm_" + sfx).argtypes = [P, I, T, T, P]
    getattr(lib, "orc_integrate_" + sfx).restype = LL
    getattr(lib, "orc_integrate_" + sfx).argtypes = [P, P, I, I, P, P, T, T, P]
    getattr(lib, "orc_raycast_" + sfx).argtypes = [P, P, P, I, I, P, P, P]
    getattr(lib, "orc_raycast_touched_" + sfx).restype = LL
    getattr(lib, "orc_raycast_touched_" + sfx).argtypes = [P, P, P, I, I]
    getattr(lib, "orc_pyramid_" + sfx).argtypes = [P, I, I, P] + [P] * 8
    getattr(lib, "orc_icp_system_" + sfx).restype = I
    getattr(lib, "orc_icp_system_" + sfx).argtypes = [P, P, I, I, P, P, I, I, P, P, P, F, F, P, P, P, P]
    getattr(lib, "orc_icp_track_" + sfx).argtypes = [P] * 6 + [I, I, P, P, I, I, P, P, P, P, F, F, P, P, P, P]
    getattr(lib, "orc_run_sequence_" + sfx).restype = I
    getattr(lib, "orc_run_sequence_" + sfx).argtypes = [P, P, I, I, I, P, P, P, F, F, P, P, P, P, I, P]
    getattr(lib, "orc_run_sequence_m_" + sfx).restype = I
    getattr(lib, "orc_run_sequence_m_" + sfx).argtypes = [P, P, I, I, I, P, P, P, F, F, P, P, I, P]
lib.orc_jacobi_eigenvalues.argtypes = [P, I, P]
lib.orc_ldlt_solve_f32.restype = I
lib.orc_ldlt_solve_f32.argtypes = [P, P, P]
lib.orc_ldlt_solve_f64.restype = I
lib.orc_ldlt_solve_f64.argtypes = [P, P, P]
lib.orc_project_so3.argtypes = [P, P]
lib.orc_damping_f32.restype = F
lib.orc_damping_f32.argtypes = [P, I]
lib.orc_floor_to_int.restype = I
lib.orc_floor_to_int.argtypes = [F, P]
lib.orc_scale_intrinsics.argtypes = [P, I, I, I, I, P]
lib.orc_num_threads.restype = I
lib.orc_set_icp_chunks.argtypes = [I]
lib.orc_predict_pose_f32.argtypes = [P, P, P]
lib.orc_predict_pose_f64.argtypes = [P, P, P]
lib.orc_mc_extract.restype = LL
lib.orc_cond_median.argtypes = [P, P, I, I, F, F]
lib.orc_cond_fill_holes.argtypes = [P, P, I, I, F, F]
lib.orc_cond_guided.argtypes = [P, P, P, I, I, F, F, I]
lib.orc_cond_ema.argtypes = [P, P, P, I, I, F, F]
lib.orc_cond_process.argtypes = [P, P, P, P, I, I, F, F, I, I]
lib.orc_mc_extract.argtypes = [P, P, I, F, P, LL, P, P, P, P, P, P]


def _d(a):
    """Address of a numpy array's data. The caller keeps the array referenced until the
    C call returns: the address of a temporary is freed (and may be reused) at once."""
    return a.ctypes.data


def cm(T, dtype=np.float32):
    """Row-major 4x4 -> column-major 16 (Eigen storage)."""
    return np.ascontiguousarray(np.asarray(T, dtype=dtype).T.reshape(16))


def rm(c):
    return np.asarray(c).reshape(4, 4).T.copy()


class Volume:
    """Oracle TSDF volume (reference 20 B AoS voxel layout), f32 or f64."""

    def __init__(self, resolution, voxel_size, truncation, max_weight=128.0, origin=(0, 0, 0), min_depth=0.3,
                 max_depth=5.0, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.sfx = "f32" if self.dtype == np.float32 else "f64"
        self.res = int(resolution)
        o = np.asarray(origin, np.float32)
        self.params = dict(resolution=resolution, voxel_size=voxel_size, truncation=truncation,
                           max_weight=max_weight, origin=tuple(float(x) for x in origin), min_depth=min_depth,
                           max_depth=max_depth)
        self.h = getattr(lib, "orc_vol_create_" + self.sfx)(self.res, voxel_size, truncation, max_weight, _d(o),
                                                             min_depth, max_depth)

    def fn(self, name):
        return getattr(lib, "orc_%s_%s" % (name, self.sfx))

    def get(self):
        n = self.res ** 3
        t = np.empty(n, self.dtype)
        w = np.empty(n, self.dtype)
        self.fn("vol_get")(self.h, _d(t), _d(w))
        r = self.res
        return t.reshape(r, r, r), w.reshape(r, r, r)

    def set(self, tsdf, weight):
        t = np.ascontiguousarray(tsdf, self.dtype).reshape(-1)
        w = np.ascontiguousarray(weight, self.dtype).reshape(-1)
        self.fn("vol_set")(self.h, _d(t), _d(w))

    def reset(self):
        self.fn("vol_reset")(self.h)

    def integrate(self, depth_m, K, T, min_depth=0.3, max_depth=5.0):
        d = np.ascontiguousarray(depth_m, self.dtype)
        H, W = d.shape
        Kc = np.asarray(K, np.float32)
        newly = ctypes.c_longlong()
        Tc = cm(T)
        n = self.fn("integrate")(self.h, _d(d), W, H, _d(Kc), _d(Tc), min_depth, max_depth, ctypes.byref(newly))
        return n, newly.value

    def raycast(self, T, K, W, H):
        v = np.empty((H, W, 3), self.dtype)
        nn = np.empty((H, W, 3), self.dtype)
        ns = ctypes.c_longlong()
        Kc = np.asarray(K, np.float32)
        Tc = cm(T)
        self.fn("raycast")(self.h, _d(Tc), _d(Kc), W, H, _d(v), _d(nn), ctypes.byref(ns))
        return v, nn

    def raycast_touched(self, T, K, W, H):
        Kc = np.asarray(K, np.float32)
        Tc = cm(T)
        return self.fn("raycast_touched")(self.h, _d(Tc), _d(Kc), W, H)

    def run_sequence(self, frames_mm, K, pose0, iters=(10, 5, 4), dist=0.1, angle=30.0, poses_in=None,
                     velocity=False):
        fr = np.ascontiguousarray(frames_mm, np.uint16)
        F_, H, W = fr.shape
        Kc = np.asarray(K, np.float32)
        it = np.asarray(iters, np.int32)
        poses_out = np.empty((F_, 16), np.float32)
        ms = np.empty((F_, 5), np.float64)
        nupd = np.empty(F_, np.int64)
        pin = None
        if poses_in is not None:
            pin = np.ascontiguousarray(np.stack([cm(T) for T in poses_in]), np.float32)
        icp = np.zeros((F_, 4), np.float64)  # converged, final_step, inliers, packed steps per level
        p0 = cm(pose0)
        self.fn("run_sequence")(self.h, _d(fr), F_, W, H, _d(Kc), _d(p0), _d(it), dist, angle,
                                _d(pin) if pin is not None else None, _d(poses_out), _d(ms), _d(nupd),
                                1 if velocity else 0, _d(icp))
        self.last_icp = icp
        return np.stack([rm(p) for p in poses_out]), ms, nupd

    def run_sequence_m(self, frames_m, K, pose0, iters=(10, 5, 4), dist=0.1, angle=30.0, velocity=False):
        """The same loop over float metre frames (F, H, W), used as given (already
        band-gated by the caller's decoding). Returns (poses (F, 4, 4), updated-voxel
        counts, ICP outcome per frame: converged, final step, inliers, packed steps)."""
        fr = np.ascontiguousarray(frames_m, np.float32)
        F_, H, W = fr.shape
        Kc = np.asarray(K, np.float32)
        it = np.asarray(iters, np.int32)
        p0 = cm(pose0)
        poses_out = np.empty((F_, 16), np.float32)
        nupd = np.empty(F_, np.int64)
        icp = np.zeros((F_, 4), np.float64)
        self.fn("run_sequence_m")(self.h, _d(fr), F_, W, H, _d(Kc), _d(p0), _d(it), dist, angle, _d(poses_out),
                                  _d(nupd), 1 if velocity else 0, _d(icp))
        return np.stack([rm(p) for p in poses_out]), nupd, icp

    def close(self):
        if self.h:
            self.fn("vol_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def predict_pose(cur, last, dtype=np.float32):
    """The pipeline's constant-velocity estimate cur * (last^-1 * cur), 4x4 row-major."""
    c, l = cm(cur, dtype), cm(last, dtype)
    out = np.empty(16, dtype)
    getattr(lib, "orc_predict_pose_" + ("f32" if np.dtype(dtype) == np.float32 else "f64"))(_d(c), _d(l), _d(out))
    return rm(out)


def depth_mm(mm, min_d=0.3, max_d=5.0, dtype=np.float32):
    mm = np.ascontiguousarray(mm, np.uint16)
    out = np.empty(mm.shape, dtype)
    getattr(lib, "orc_depth_mm_" + ("f32" if np.dtype(dtype) == np.float32 else "f64"))(
        _d(mm), mm.size, min_d, max_d, _d(out))
    return out


def pyramid(depth_m, K, dtype=np.float32):
    d = np.ascontiguousarray(depth_m, dtype)
    H, W = d.shape
    sfx = "f32" if np.dtype(dtype) == np.float32 else "f64"
    Kc = np.asarray(K, np.float32)
    shapes = [(H, W), (H // 2, W // 2), (H // 4, W // 4)]
    d1 = np.empty(shapes[1], dtype)
    d2 = np.empty(shapes[2], dtype)
    vs = [np.empty(s + (3,), dtype) for s in shapes]
    ns = [np.empty(s + (3,), dtype) for s in shapes]
    getattr(lib, "orc_pyramid_" + sfx)(_d(d), W, H, _d(Kc), _d(d1), _d(d2), _d(vs[0]), _d(ns[0]), _d(vs[1]),
                                        _d(ns[1]), _d(vs[2]), _d(ns[2]))
    return [d, d1, d2], vs, ns


def icp_system(lv, ln, mv, mn, K, T, ref, dist=0.1, angle=30.0, dtype=np.float32):
    sfx = "f32" if np.dtype(dtype) == np.float32 else "f64"
    lv = np.ascontiguousarray(lv, dtype)
    ln = np.ascontiguousarray(ln, dtype)
    mv = np.ascontiguousarray(mv, dtype)
    mn = np.ascontiguousarray(mn, dtype)
    H, W = lv.shape[:2]
    MH, MW = mv.shape[:2]
    A = np.empty(36, dtype)
    b = np.empty(6, dtype)
    res = np.empty(1, dtype)
    cnt = np.empty(6, np.int32)
    Kc = np.asarray(K, np.float32)
    Tc, Rc = cm(T), cm(ref)
    ok = getattr(lib, "orc_icp_system_" + sfx)(_d(lv), _d(ln), W, H, _d(mv), _d(mn), MW, MH, _d(Kc), _d(Tc),
                                                 _d(Rc), dist, angle, _d(A), _d(b), _d(res), _d(cnt))
    return bool(ok), A.reshape(6, 6), b, float(res[0]), dict(zip(
        ["valid_live", "valid_model", "projected", "dist_filtered", "angle_filtered", "inliers"], cnt.tolist()))


def icp_track(levels_v, levels_n, mv, mn, K, est, ref, iters=(10, 5, 4), dist=0.1, angle=30.0, dtype=np.float32):
    sfx = "f32" if np.dtype(dtype) == np.float32 else "f64"
    lv = [np.ascontiguousarray(v, dtype) for v in levels_v]
    ln = [np.ascontiguousarray(n, dtype) for n in levels_n]
    mv = np.ascontiguousarray(mv, dtype)
    mn = np.ascontiguousarray(mn, dtype)
    H, W = lv[0].shape[:2]
    MH, MW = mv.shape[:2]
    Kc = np.asarray(K, np.float32)
    it = np.asarray(iters, np.int32)
    pose = np.empty(16, dtype)
    sc = np.empty(2, dtype)
    ints = np.empty(9, np.int32)
    info = np.empty(36, dtype)
    Ec, Rc = cm(est), cm(ref)  # kept alive across the call: _d() takes the address only
    getattr(lib, "orc_icp_track_" + sfx)(_d(lv[0]), _d(ln[0]), _d(lv[1]), _d(ln[1]), _d(lv[2]), _d(ln[2]), W, H,
                                          _d(mv), _d(mn), MW, MH, _d(Kc), _d(Ec), _d(Rc), _d(it), dist,
                                          angle, _d(pose), _d(sc), _d(ints), _d(info))
    keys = ["inliers", "converged", "tracking_ok", "valid_live_points", "valid_model_points", "projected_points",
            "dist_filtered", "angle_filtered", "steps"]
    r = dict(zip(keys, ints.tolist()))
    r.update(pose=rm(pose), error=float(sc[0]), final_step=float(sc[1]), information=info.reshape(6, 6))
    return r


def mc_extract(tsdf, weight, voxel_size, origin, max_tris=2000000):
    t = np.ascontiguousarray(tsdf, np.float32)
    w = np.ascontiguousarray(weight, np.float32)
    res = t.shape[0]
    o = np.asarray(origin, np.float32)
    nv = ctypes.c_longlong()
    ni = ctypes.c_longlong()
    tr = ctypes.c_int()
    lib.orc_mc_extract(_d(t), _d(w), res, voxel_size, _d(o), max_tris, None, None, None, ctypes.byref(nv),
                       ctypes.byref(ni), ctypes.byref(tr))
    pos = np.empty((nv.value, 3), np.float32)
    nrm = np.empty((nv.value, 3), np.float32)
    idx = np.empty(ni.value, np.uint32)
    lib.orc_mc_extract(_d(t), _d(w), res, voxel_size, _d(o), max_tris, _d(pos), _d(nrm), _d(idx), ctypes.byref(nv),
                       ctypes.byref(ni), ctypes.byref(tr))
    return pos, nrm, idx, bool(tr.value)


def point_cloud(tsdf, weight, voxel_size, origin):
    """TSDFVolume::extractGlobalPointCloud (src/tsdf/TSDFVolume.cpp:571-632) restated
    serially: every voxel with weight > 1 and |tsdf| < 0.2, at voxelToWorld(x, y, z) =
    origin + (x, y, z) * voxel_size (TSDFVolume.cpp:559-561) in float32, in linear voxel
    order (z, y, x with x fastest; the reference's CPU form concatenates per-thread lists,
    its GPU form compacts in this order). tsdf / weight are (res, res, res) arrays indexed
    [z, y, x]."""
    t = np.asarray(tsdf, np.float32)
    w = np.asarray(weight, np.float32)
    keep = (w > np.float32(1.0)) & (np.abs(t) < np.float32(0.2))
    z, y, x = np.nonzero(keep)
    vs = np.float32(voxel_size)
    o = np.asarray(origin, np.float32)
    return np.stack([o[0] + x.astype(np.float32) * vs, o[1] + y.astype(np.float32) * vs,
                     o[2] + z.astype(np.float32) * vs], axis=1).astype(np.float32)


def depth_raw(raw, min_d=0.3, max_d=5.0):
    """cpuDepthMeters (include/sensor/DepthValidity.h:55-93) restated in float32: raw 0
    and raw >= 2047 -> 0; else 1 / (raw * -0.0030711016 + 3.3309495161), kept when finite
    and inside [min_d, max_d], else 0 (no clamping)."""
    r = np.asarray(raw, np.uint16)
    valid = (r != 0) & (r < 2047)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        m = np.float32(1.0) / (r.astype(np.float32) * np.float32(-0.0030711016) + np.float32(3.3309495161))
    keep = valid & np.isfinite(m) & (m >= np.float32(min_d)) & (m <= np.float32(max_d))
    return np.where(keep, m, np.float32(0.0)).astype(np.float32)


# Depth conditioning (src/sensor/SignalConditioner_omp.cpp:337-716), restated in
# oracle_condition.hpp. Stage bits, applied in this order.
COND_MEDIAN, COND_HOLES, COND_GUIDED, COND_EMA = 1, 2, 4, 8
COND_FULL = 15


def _u16(a):
    return np.ascontiguousarray(a, np.uint16)


def cond_median(raw, lo=0.3, hi=2.5):
    r = _u16(raw)
    out = np.empty_like(r)
    lib.orc_cond_median(_d(r), _d(out), r.shape[1], r.shape[0], lo, hi)
    return out


def cond_fill_holes(raw, lo=0.3, hi=2.5):
    r = _u16(raw)
    out = np.empty_like(r)
    lib.orc_cond_fill_holes(_d(r), _d(out), r.shape[1], r.shape[0], lo, hi)
    return out


def cond_guided(raw, luma=None, lo=0.3, hi=2.5, exp_mode=0):
    """exp_mode 0: the C library's expf, as the reference; 1: the double exponential
    rounded to float (correctly rounded in practice)."""
    r = _u16(raw)
    lu = np.zeros(r.shape, np.float32) if luma is None else np.ascontiguousarray(luma, np.float32)
    out = np.empty_like(r)
    lib.orc_cond_guided(_d(r), _d(lu), _d(out), r.shape[1], r.shape[0], lo, hi, exp_mode)
    return out


def cond_ema(raw, state, lo=0.3, hi=2.5):
    """One EMA pass; state (float32, metres, 0 = no history) is updated in place."""
    r = _u16(raw)
    assert state.dtype == np.float32 and state.flags.c_contiguous and state.shape == r.shape
    out = np.empty_like(r)
    lib.orc_cond_ema(_d(r), _d(out), _d(state), r.shape[1], r.shape[0], lo, hi)
    return out


def cond_process(raw, state, luma=None, lo=0.3, hi=2.5, stages=COND_FULL, exp_mode=0):
    r = _u16(raw)
    lu = np.zeros(r.shape, np.float32) if luma is None else np.ascontiguousarray(luma, np.float32)
    assert state.dtype == np.float32 and state.flags.c_contiguous and state.shape == r.shape
    out = np.empty_like(r)
    lib.orc_cond_process(_d(r), _d(lu), _d(out), _d(state), r.shape[1], r.shape[0], lo, hi, stages, exp_mode)
    return out
