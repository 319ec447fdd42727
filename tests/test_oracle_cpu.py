"""The CPU oracle pinned against the reference's own fixtures and closed forms.

Every check here is one the reference's CPU contracts make (same numbers), or a
property the mathematics fixes (closed forms, library routines on small inputs), so a
dropped term, a wrong sign or a transposed operand in the restatement fails one of them.
"""
import numpy as np
import pytest

import oracle
from tests import fixtures as fx

F32 = np.float32


# ---- ICP: tests/icp_weighting_contract.cpp -------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_icp_weighting_canonical(dtype):
    lv, ln, mv, mn = fx.icp_weighting("clean")
    I4 = np.eye(4, dtype=np.float32)
    r = oracle.icp_track(lv, ln, mv, mn, fx.K640, I4, I4, iters=(1, 0, 0), dtype=dtype)
    assert r["valid_live_points"] == 11 and r["valid_model_points"] == 11 and r["projected_points"] == 11
    assert r["inliers"] == 11 and r["dist_filtered"] == 0 and r["angle_filtered"] == 0
    tz = float(r["pose"][2, 3])
    assert abs(tz - fx.ICP_TZ) <= fx.ICP_TOL
    assert abs(tz - fx.ICP_TZ_UNWEIGHTED) >= fx.ICP_GAP and abs(tz - fx.ICP_TZ_W2) >= fx.ICP_GAP
    assert abs(r["error"] - fx.ICP_ERR) <= fx.ICP_TOL and abs(r["error"] - fx.ICP_ERR_WE2) >= fx.ICP_GAP
    assert abs(float(r["pose"][0, 3])) <= fx.ICP_TOL and abs(float(r["pose"][1, 3])) <= fx.ICP_TOL
    r2 = oracle.icp_track(lv, ln, mv, mn, fx.K640, I4, I4, iters=(1, 0, 0), dtype=dtype)
    assert np.array_equal(r["pose"], r2["pose"]) and r["error"] == r2["error"]


@pytest.mark.parametrize("corruption", ["inf_normal", "nan_depth"])
def test_icp_weighting_rejects_non_finite(corruption):
    I4 = np.eye(4, dtype=np.float32)
    clean = oracle.icp_track(*fx.icp_weighting("clean"), fx.K640, I4, I4, iters=(1, 0, 0))
    r = oracle.icp_track(*fx.icp_weighting(corruption), fx.K640, I4, I4, iters=(1, 0, 0))
    assert r["projected_points"] == 12 and r["inliers"] == 11 and r["valid_model_points"] == 11
    assert np.isfinite(r["pose"]).all() and np.isfinite(r["error"])
    assert np.array_equal(r["pose"], clean["pose"]) and r["error"] == clean["error"]


def test_icp_system_matches_hand_sums():
    """The 6x6 system of the weighting fixture against sums written out in double."""
    lv, ln, mv, mn = fx.icp_weighting("clean")
    I4 = np.eye(4, dtype=np.float32)
    ok, A, b, res, cnt = oracle.icp_system(lv[0], ln[0], mv, mn, fx.K640, I4, I4, dtype=np.float64)
    assert ok and cnt["inliers"] == 11
    Ar = np.zeros((6, 6))
    br = np.zeros(6)
    rr = 0.0
    k = float(np.float32(0.02))  # kHuberK is the float constant 0.02f (ICPShared.h:34)
    for (py, px) in np.argwhere(mn[..., 2] == 1):
        v = lv[0][py, px].astype(np.float64)
        e = float(v[2] - mv[py, px, 2])
        J = np.concatenate([[0, 0, 1], np.cross(v, [0, 0, 1])])
        w = 1.0 if abs(e) <= k else k / abs(e)
        Ar += w * np.outer(J, J)
        br -= J * w * e
        rr += e * e if abs(e) <= k else 2 * k * abs(e) - k ** 2
    np.testing.assert_allclose(A, Ar, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b, br, rtol=1e-12, atol=1e-12)
    assert abs(res - rr) < 1e-12


# ---- pyramid: tests/icp_weighting_contract.cpp:277-391 ---------------------------
def test_pyramid_jump_guard():
    d, v, n = fx.pyramid_field()
    # the oracle's pyramid back-projects level 0 from depth; the fixture's vertices are
    # arbitrary, so drive downsample through a 16x16 level-0 identical to the fixture
    ds, vs, ns = oracle.pyramid(d, (1.0, 1.0, 0.0, 0.0))
    assert ds[1].shape == (8, 8)
    # depths follow from the fixture directly
    assert ds[1][0, 0] == 1.0 and ds[1][0, 1] == 1.0 and ds[1][2, 2] == 0.0
    assert abs(float(ds[1][1, 0]) - 1.00625) <= 1e-6
    lib_d1, lib_v1, lib_n1 = _downsample_fixture(d, v, n)
    fx.check_pyramid_field((d, v, n), (lib_d1, lib_v1, lib_n1))


def _downsample_fixture(d, v, n):
    """Level 1 of the oracle pyramid built from the fixture's own level-0 arrays."""
    import ctypes
    H, W = d.shape
    d1 = np.empty((H // 2, W // 2), np.float32)
    v1 = np.empty((H // 2, W // 2, 3), np.float32)
    n1 = np.empty_like(v1)
    lib = oracle.lib
    fn = lib.orc_downsample_f32
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 3
    dd = np.ascontiguousarray(d, np.float32)
    vv = np.ascontiguousarray(v, np.float32)
    nn = np.ascontiguousarray(n, np.float32)
    fn(dd.ctypes.data, vv.ctypes.data, nn.ctypes.data, W, H, d1.ctypes.data, v1.ctypes.data, n1.ctypes.data)
    return d1, v1, n1


# ---- integration: tests/tsdf_integration_race_contract.cpp:232-324 ----------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_integrate_hand_fold(dtype):
    p = fx.FOLD_PARAMS
    v = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                      p["min_depth"], p["max_depth"], dtype=dtype)
    I4 = np.eye(4, dtype=np.float32)
    for d in fx.fold_frames():
        v.integrate(d, fx.FOLD_K, I4, *fx.FOLD_BAND)
    t, w = v.get()
    fx.check_fold(t, w)


def test_integrate_thread_count_invariant():
    """One writer per voxel: the volume is byte-identical whatever the OpenMP team size
    (tsdf_integration_race_contract.cpp:121-230)."""
    import subprocess
    import sys
    code = r"""
import sys, hashlib; sys.path.insert(0, %r)
import numpy as np, oracle
from azu_amd import scene as sc
v = oracle.Volume(96, 0.0107, 0.032, origin=sc.ORIGIN)
for f in range(3):
    T = sc.pose(f, 1); d = oracle.depth_mm(sc.render(T, boxes=10, frame=f))
    v.integrate(d, sc.K640, T)
t, w = v.get(); print(hashlib.sha256(t.tobytes() + w.tobytes()).hexdigest())
""" % (fx.__file__.rsplit("/tests/", 1)[0],)
    hashes = set()
    for n in ("1", "3", "8"):
        env = dict(__import__("os").environ, OMP_NUM_THREADS=n)
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True)
        hashes.add(out.stdout.strip())
    assert len(hashes) == 1


# ---- small dense algebra against library routines ----------------------------------
def test_jacobi_eigenvalues_vs_numpy():
    rng = np.random.default_rng(1)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = np.ascontiguousarray(M @ M.T + np.diag(rng.uniform(0, 1, 6)))
        eig = np.empty(6)
        oracle.lib.orc_jacobi_eigenvalues(A.ctypes.data, 6, eig.ctypes.data)
        np.testing.assert_allclose(eig, np.linalg.eigvalsh(A), rtol=1e-10, atol=1e-10)


def test_ldlt_solve_vs_numpy():
    rng = np.random.default_rng(2)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = np.ascontiguousarray(M @ M.T + 0.1 * np.eye(6))
        b = rng.normal(size=6)
        x = np.empty(6)
        assert oracle.lib.orc_ldlt_solve_f64(A.ctypes.data, b.ctypes.data, x.ctypes.data) == 1
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-9)


def test_project_so3_vs_svd():
    rng = np.random.default_rng(3)
    for k in range(30):
        M = np.ascontiguousarray(rng.normal(size=(3, 3)))
        R = np.empty(9)
        oracle.lib.orc_project_so3(M.ctypes.data, R.ctypes.data)
        R = R.reshape(3, 3)
        U, s, Vt = np.linalg.svd(M)
        if np.linalg.det(U @ Vt) < 0:
            U[:, 2] = -U[:, 2]
        np.testing.assert_allclose(R, U @ Vt, atol=1e-10)
        assert abs(np.linalg.det(R) - 1) < 1e-12


def test_damping_policy():
    """icp_numeric_policy_contract: 0.01 at level 0, 0.1 above, 1.0 when singular or
    ill-conditioned (ICPShared.h:86-100)."""
    A = np.ascontiguousarray(np.diag([1.0, 2, 3, 4, 5, 6]).astype(np.float32))
    assert oracle.lib.orc_damping_f32(A.ctypes.data, 0) == np.float32(0.01)
    assert oracle.lib.orc_damping_f32(A.ctypes.data, 1) == np.float32(0.1)
    S = np.ascontiguousarray(np.diag([0.0, 2, 3, 4, 5, 6]).astype(np.float32))
    assert oracle.lib.orc_damping_f32(S.ctypes.data, 0) == np.float32(1.0)
    C = np.ascontiguousarray(np.diag([1e-3, 2, 3, 4, 5, 2e5]).astype(np.float32))
    assert oracle.lib.orc_damping_f32(C.ctypes.data, 0) == np.float32(1.0)  # cond 2e8 > 1e7


def test_floor_to_int_and_scale_intrinsics():
    """coordinate_rounding_contract / icp_lowres_model_contract closed forms."""
    import ctypes
    out = ctypes.c_int()
    for v, want in ((2.5, 2), (-0.5, -1), (-2.0, -2), (0.0, 0)):
        assert oracle.lib.orc_floor_to_int(v, ctypes.byref(out)) == 1 and out.value == want
    for v in (float("nan"), float("inf"), -float("inf"), 3e9, -3e9):
        assert oracle.lib.orc_floor_to_int(v, ctypes.byref(out)) == 0
    K = np.array(fx.K640, np.float32)
    o = np.empty(4, np.float32)
    oracle.lib.orc_scale_intrinsics(K.ctypes.data, 640, 480, 320, 240, o.ctypes.data)
    assert o.tolist() == [262.5, 262.5, 159.5, 119.5]
    oracle.lib.orc_scale_intrinsics(K.ctypes.data, 640, 480, 640, 480, o.ctypes.data)
    assert o.tolist() == list(K)
    oracle.lib.orc_scale_intrinsics(K.ctypes.data, 640, 480, 1280, 960, o.ctypes.data)
    assert o.tolist() == [1050.0, 1050.0, 639.5, 479.5]


def planar_field(res, vs, trunc, z0):
    """A planar TSDF written on paper: voxel (x, y, z) at depth origin_z + z * vs holds
    clip((z0 - (origin_z + z * vs)) / trunc, -1, 1) with weight 2, the rest weight 0."""
    zs = np.arange(res, dtype=np.float64) * vs
    t = np.clip((z0 - zs) / trunc, -1.0, 1.0).astype(np.float32)
    tsdf = np.broadcast_to(t[:, None, None], (res, res, res)).copy()
    weight = np.full((res, res, res), 2.0, np.float32)
    return tsdf, weight


def test_point_cloud_oracle_closed_forms():
    """The point cloud of a planar field is whole x-y layers: the layers whose |tsdf| is
    below 0.2, i.e. |z0 - z vs| < 0.2 trunc, so res^2 points per layer; the predicate is
    strict on both bounds (weight exactly 1 and |tsdf| exactly 0.2 are out), and NaN
    fails it."""
    res, vs, trunc = 16, 0.01, 0.05
    origin = (-0.08, -0.08, 0.0)
    tsdf, weight = planar_field(res, vs, trunc, z0=0.075)
    pts = oracle.point_cloud(tsdf, weight, vs, origin)
    layers = [z for z in range(res) if abs(0.075 - z * vs) < 0.2 * trunc - 1e-9]
    assert layers == [7, 8]
    assert pts.shape == (len(layers) * res * res, 3)
    # linear order: z, then y, then x fastest
    assert np.all(pts[:res, 0] == np.float32(-0.08) + np.arange(res, dtype=np.float32) * np.float32(vs))
    assert np.all(pts[:, 2][: res * res] == np.float32(7) * np.float32(vs))
    # strict bounds and NaN
    t = np.zeros((2, 2, 2), np.float32)
    w = np.full((2, 2, 2), 2.0, np.float32)
    w[0, 0, 0] = 1.0                      # weight exactly 1: out
    w[0, 0, 1] = np.nextafter(np.float32(1), np.float32(2))  # just above: in
    t[0, 1, 0] = np.float32(0.2)          # |tsdf| exactly 0.2: out
    t[0, 1, 1] = np.float32(-0.2)         # out
    t[1, 0, 0] = np.nextafter(np.float32(0.2), np.float32(0))  # in
    t[1, 0, 1] = np.nan                   # out
    w[1, 1, 0] = np.nan                   # out
    pts = oracle.point_cloud(t, w, 1.0, (0, 0, 0))
    # (x, y, z) of the kept voxels, in linear order
    assert [tuple(int(v) for v in p) for p in pts] == [(1, 0, 0), (0, 0, 1), (1, 1, 1)]


def test_depth_raw_oracle_golden_facts():
    """tests/depth_domain_contract.cpp:140-230 and tests/cpu_smoke_harness.cpp:48-57 of the
    reference: the predicate over all 65536 codes, invalid -> exactly +0.0, the default
    band [0.3, 2.5] accepts exactly raw 1..954, the pole codes 1084 / 1085 and 1500 / 2046
    are rejected, nothing negative / NaN / Inf comes out, raw 700 lands in the band, and
    every accepted value matches the double-precision curve."""
    raw = np.arange(65536, dtype=np.uint16)
    m = oracle.depth_raw(raw, 0.3, 2.5)
    assert np.all(np.isfinite(m)) and np.all(m >= 0)
    zero = m == 0
    assert np.all(np.signbit(m[zero]) == False)  # noqa: E712  (+0.0, not -0.0)
    accepted = np.nonzero(~zero)[0]
    assert accepted[0] == 1 and accepted[-1] == 954 and len(accepted) == 954
    for r in (0, 955, 1084, 1085, 1500, 2046, 2047, 2048, 65535):
        assert m[r] == 0.0, r
    d = 1.0 / (accepted.astype(np.float64) * np.float64(np.float32(-0.0030711016)) +
               np.float64(np.float32(3.3309495161)))
    assert np.max(np.abs(m[accepted] - d)) < 1e-6
    assert 0.3 < oracle.depth_raw(np.uint16(700), 0.3, 2.5) < 2.5
    assert oracle.depth_raw(np.uint16(500), 0.3, 5.0) < oracle.depth_raw(np.uint16(700), 0.3, 5.0)
