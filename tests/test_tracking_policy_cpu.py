"""The frame-level tracking policy (include/tracking/TrackingPolicy.h, host C++ over the
ICP result; reference include/tracking/TrackingPolicy.h:36-161).

Pinned three ways: the known-answer grades of the reference's
tests/tracking_policy_contract.cpp:51-105 (restated in tests/tracking_policy_contract.cpp,
with closed-form degeneracy cases on diagonal information matrices), and on seeded random
inputs against independent library routines: numpy.linalg.eigh for the 6x6 eigen-solve,
scipy's Rotation (as_rotvec / from_rotvec) for the rotation vector and Rodrigues' formula
of keepObservableMotion, numpy's 4x4 inverse for the motion gate."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

SHIM = r"""
#include "tracking/TrackingPolicy.h"
using namespace kfusion::tracking;
extern "C" int tp_eig(const double* a, double* ev, double* vec) { return policy_detail::symEig6(a, ev, vec) ? 1 : 0; }
extern "C" float tp_ratio(const float* info) { return weakestDirectionRatio(info); }
extern "C" int tp_within(const float* pose, const float* prev) { return withinFrameMotion(pose, prev) ? 1 : 0; }
extern "C" int tp_keep(const float* prev, const float* est, const float* info, float thr, float* out) {
    ObservableMotion m = keepObservableMotion(prev, est, info, thr);
    for (int i = 0; i < 16; ++i) out[i] = m.pose.data()[i];
    return m.degenerate_dofs;
}
"""


def test_contract_cases(tmp_path):
    exe = tmp_path / "contract"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", INC,
                    os.path.join(ROOT, "tests", "tracking_policy_contract.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "PASS" in out.stdout


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("tp")
    src, so = d / "shim.cpp", d / "libtp.so"
    src.write_text(SHIM)
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I", INC, str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.tp_ratio.restype = ctypes.c_float
    lib.tp_keep.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_float, ctypes.c_void_p]
    return lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def col_major(T):
    return np.ascontiguousarray(T.T.reshape(16)).astype(np.float32)


def from_col_major(a):
    return np.asarray(a, np.float64).reshape(4, 4).T


def random_pose(rng, max_angle, max_t):
    T = np.eye(4)
    rv = rng.normal(size=3)
    rv *= rng.uniform(0, max_angle) / np.linalg.norm(rv)
    T[:3, :3] = Rotation.from_rotvec(rv).as_matrix()
    T[:3, 3] = rng.uniform(-max_t, max_t, 3)
    return T


def spd(rng, eigenvalues):
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    return (Q * np.asarray(eigenvalues)) @ Q.T


def test_eigen_solve_matches_numpy(shim):
    rng = np.random.default_rng(7)
    for case in range(200):
        scale = 10.0 ** rng.uniform(-3, 6)
        ev_true = scale * 10.0 ** rng.uniform(-8, 0, 6)
        if case % 5 == 0:
            ev_true[rng.integers(6)] = 0.0
        if case % 7 == 0:
            ev_true[1] = ev_true[2]  # a repeated eigenvalue
        A = spd(rng, ev_true)
        A = 0.5 * (A + A.T)
        ev, vec = np.zeros(6), np.zeros(36)
        assert shim.tp_eig(ptr(A), ptr(ev), ptr(vec)) == 1
        want = np.linalg.eigh(A)[0]
        np.testing.assert_allclose(ev, want, rtol=0, atol=1e-12 * np.abs(want).max())
        V = vec.reshape(6, 6)
        np.testing.assert_allclose(V.T @ V, np.eye(6), atol=1e-12)
        np.testing.assert_allclose(V.T @ A @ V, np.diag(ev), atol=1e-11 * np.abs(want).max())


def test_weakest_direction_ratio_matches_numpy(shim):
    rng = np.random.default_rng(11)
    for _ in range(200):
        info = spd(rng, 10.0 ** rng.uniform(-4, 5, 6)).astype(np.float32)
        info = 0.5 * (info + info.T)
        w = np.linalg.eigvalsh(info.astype(np.float64))
        want = max(0.0, w[0]) / w[-1]
        got = shim.tp_ratio(ptr(np.ascontiguousarray(info)))
        assert abs(got - want) <= 1e-6 * max(want, 1e-3), (got, want)


def test_within_frame_motion_matches_numpy(shim):
    rng = np.random.default_rng(3)
    n_in = n_out = 0
    for _ in range(400):
        prev = random_pose(rng, np.pi, 2.0)
        step = random_pose(rng, 0.9, 0.25)
        pose = prev @ step
        pf, qf = col_major(prev), col_major(pose)
        d = np.linalg.inv(from_col_major(pf)) @ from_col_major(qf)
        trans = np.linalg.norm(d[:3, 3])
        angle = np.arccos(np.clip((np.trace(d[:3, :3]) - 1.0) * 0.5, -1.0, 1.0))
        if abs(trans - 0.15) < 1e-5 or abs(angle - 0.52) < 1e-5:
            continue  # within rounding of a gate edge
        want = trans <= 0.15 and angle <= 0.52
        assert shim.tp_within(ptr(qf), ptr(pf)) == int(want)
        n_in += want
        n_out += not want
    assert n_in > 20 and n_out > 20


def observable_numpy(prev, est, info, thr):
    """keepObservableMotion stated with numpy.linalg.eigh and scipy's rotation vector."""
    w, V = np.linalg.eigh(info)
    if not np.all(np.isfinite(info)) or not w[-1] > 0:
        return est, 0
    d = np.linalg.inv(prev) @ est
    xi = np.concatenate([d[:3, 3], Rotation.from_matrix(d[:3, :3]).as_rotvec()])
    keep = w >= thr * w[-1]
    dofs = int((~keep).sum())
    if dofs == 0:
        return est, 0
    kept = V[:, keep] @ (V[:, keep].T @ xi)
    step = np.eye(4)
    step[:3, :3] = Rotation.from_rotvec(kept[3:]).as_matrix()
    step[:3, 3] = kept[:3]
    return prev @ step, dofs


def test_keep_observable_motion_matches_numpy(shim):
    rng = np.random.default_rng(5)
    seen = set()
    for case in range(300):
        prev = random_pose(rng, np.pi * 0.95, 2.0)
        step = random_pose(rng, 3.0 if case % 10 == 0 else 0.4, 0.2)
        est = prev @ step
        # Eigenvalues clear of the threshold, so the kept set is not decided by rounding.
        n_low = case % 6
        ev = np.concatenate([10.0 ** rng.uniform(-8, -4, n_low), 10.0 ** rng.uniform(-1, 0, 6 - n_low)])
        info = spd(rng, 1e4 * ev).astype(np.float32)
        info = 0.5 * (info + info.T)
        pf, ef = col_major(prev), col_major(est)
        out = np.zeros(16, np.float32)
        dofs = shim.tp_keep(ptr(pf), ptr(ef), ptr(np.ascontiguousarray(info)), 5e-3, ptr(out))
        want, want_dofs = observable_numpy(from_col_major(pf), from_col_major(ef), info.astype(np.float64), 5e-3)
        assert dofs == want_dofs == n_low
        np.testing.assert_allclose(from_col_major(out), want, rtol=0, atol=5e-6)
        if dofs == 0:
            assert np.array_equal(out, ef)
        seen.add(dofs)
    assert seen == set(range(6))


def _yaw_move(tx, yaw):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec([0.0, yaw, 0.0]).as_matrix()
    T[0, 3] = tx
    return T


def _grade(pose, prev, inliers, valid_model, error, information=None, rel=0.0, policy=None):
    from azu_amd import native as N
    r = N.IcpResult()
    r.pose[:] = col_major(pose).tolist()
    r.inliers, r.valid_model_points, r.valid_live_points, r.error = inliers, valid_model, 300000, error
    if information is not None:
        r.information[:] = np.asarray(information, np.float32).reshape(36).tolist()
    pv = (ctypes.c_float * 16)(*col_major(prev).tolist())
    out = N.TrackDecision()
    assert N.lib.azu_track_grade(ctypes.byref(r), pv, policy, rel, ctypes.byref(out)) == N.AZU_OK
    return out


def test_c_abi_grades_like_the_reference_cases():
    """azu_track_grade (the C ABI over the policy header) on the reference contract's cases."""
    from azu_amd import native as N
    prev = _yaw_move(0.3, 0.2)
    near = prev @ _yaw_move(0.02, 0.03)
    far = prev @ _yaw_move(0.45, 0.7)
    cases = [
        (near, 150000, 200000, 4e-6, N.TRACK_GOOD),
        (near, 30000, 200000, 4e-6, N.TRACK_POOR),
        (near, 40000, 45000, 4e-6, N.TRACK_GOOD),
        (near, 1500, 1600, 4e-6, N.TRACK_POOR),
        (near, 150000, 200000, 9e-4, N.TRACK_POOR),
        (prev @ _yaw_move(0.2, 0.0), 150000, 200000, 4e-6, N.TRACK_FAILED),
        (prev @ _yaw_move(0.0, 0.7), 150000, 200000, 4e-6, N.TRACK_FAILED),
        (near, 99, 200000, 4e-6, N.TRACK_FAILED),
        (near, 150000, 200000, float("inf"), N.TRACK_FAILED),
        (far, 150000, 200000, 4e-6, N.TRACK_FAILED),
    ]
    for pose, inl, vm, err, want in cases:
        d = _grade(pose, prev, inl, vm, err)
        assert d.quality == want, (inl, vm, err, want)
        assert d.integrate == (want == N.TRACK_GOOD)
        # Failed keeps the previous pose; otherwise the graded pose is continued from.
        keep = prev if want == N.TRACK_FAILED else pose
        np.testing.assert_array_equal(np.array(d.pose, np.float32), col_major(keep))
        assert d.degenerate_dofs == 0
    bad = near.copy()
    bad[0, 3] = np.nan
    assert _grade(bad, prev, 150000, 200000, 4e-6).quality == N.TRACK_FAILED


def test_c_abi_policy_defaults_and_custom_policy():
    from azu_amd import native as N
    p = N.TrackingPolicy()
    assert N.lib.azu_tracking_policy_defaults(ctypes.byref(p)) == N.AZU_OK
    assert (p.min_inliers, p.min_good_inliers, p.failures_before_lost) == (100, 2000, 3)
    np.testing.assert_allclose([p.min_fit_ratio, p.max_rms_m, p.max_frame_translation, p.max_frame_rotation],
                               [0.40, 0.015, 0.15, 0.52], rtol=1e-7)
    prev = _yaw_move(0.3, 0.2)
    pose = prev @ _yaw_move(0.2, 0.0)
    assert _grade(pose, prev, 150000, 200000, 4e-6).quality == N.TRACK_FAILED
    p.max_frame_translation = 0.25
    assert _grade(pose, prev, 150000, 200000, 4e-6, policy=ctypes.byref(p)).quality == N.TRACK_GOOD
    r = N.IcpResult()
    out = N.TrackDecision()
    pv = (ctypes.c_float * 16)()
    assert N.lib.azu_track_grade(ctypes.byref(r), pv, None, float("nan"), ctypes.byref(out)) == N.AZU_ERR_ARG
    assert N.lib.azu_track_grade(None, pv, None, 0.0, ctypes.byref(out)) == N.AZU_ERR_ARG


def test_c_abi_degeneracy_applies_before_grading():
    """The pipeline order: keepObservableMotion first, then classifyTracking on its pose. A
    0.2 m slide along an unobserved x fails the motion gate as solved, and grades Good once
    the unobserved slide is removed."""
    from azu_amd import native as N
    prev = _yaw_move(0.3, 0.2)
    pose = prev @ _yaw_move(0.2, 0.0)
    info = np.diag([1e-3, 100, 100, 100, 100, 100])
    assert _grade(pose, prev, 150000, 200000, 4e-6, info, rel=0.0).quality == N.TRACK_FAILED
    d = _grade(pose, prev, 150000, 200000, 4e-6, info, rel=5e-3)
    assert d.quality == N.TRACK_GOOD and d.degenerate_dofs == 1
    np.testing.assert_allclose(from_col_major(np.array(d.pose)), prev, atol=2e-6)
    ratio = ctypes.c_float()
    assert N.lib.azu_track_weakest_ratio((ctypes.c_float * 36)(*np.float32(info).reshape(36)),
                                         ctypes.byref(ratio)) == N.AZU_OK
    assert abs(ratio.value - 1e-5) < 1e-10
