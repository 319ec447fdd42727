"""FP32 GPU path vs the FP32 CPU oracle on identical seeded frames.

Bars: integration, raycast and the live pyramid are bit-identical (the exact kernels
run the oracle's float operation sequence); the ICP pose agrees within 1e-5 m /
1e-5 rad (FP32 terms and per-wave FP32 sums, FP64 across waves and blocks, on the
GPU against the reference's float accumulation in the oracle).
"""
import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from tests.gpu_util import (frames, gpu_volume, icp_pose_parity, oracle_volume, pose_diff, say, small_params,
                            steps_per_level)

pytestmark = pytest.mark.gpu


def test_device_present():
    dev, name, hbm = N.device_info()
    say("device", dev, name, hbm)
    assert "gfx950" in name


@pytest.mark.parametrize("res,vs", [(64, 0.016), (128, 0.008), (256, 0.004)])
def test_integrate_exact_bitwise(res, vs):
    p = small_params(res, vs)
    g = gpu_volume(p)
    say("integrate exact", res, "verified fixed-divisor division (voxel size, truncation):", g.division_check())
    o = oracle_volume(p)
    upd_total = 0
    for mm, T in frames(3):
        d = oracle.depth_mm(mm)
        n, _ = o.integrate(d, sc.K640, T)
        upd_total += n
        g.integrate(d, sc.K640, T)
    tg, wg = g.voxelData()
    to, wo = o.get()
    assert np.array_equal(wg, wo)
    assert np.array_equal(tg.view(np.uint32), to.view(np.uint32))
    upd, _ = g.counters()
    assert upd == upd_total


def test_raycast_exact_bitwise():
    p = small_params(128, 0.008)
    g = gpu_volume(p)
    o = oracle_volume(p)
    fr = frames(3)
    for mm, T in fr:
        d = oracle.depth_mm(mm)
        o.integrate(d, sc.K640, T)
        g.integrate(d, sc.K640, T)
    for _, T in fr:
        vg, ng = g.raycast(T, sc.K640)
        vo, no = o.raycast(T, sc.K640, 640, 480)
        say("raycast hits", (vo[..., 2] != 0).mean())
        assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32))
        assert np.array_equal(ng.view(np.uint32), no.view(np.uint32))


def test_raycast_degenerate_poses_bitwise():
    """Poses whose rays never reach a sample (a non-finite rotation or translation, a
    camera far outside the volume) and a camera inside the volume: the FP32 raycast
    against the oracle bit for bit (every non-finite pose is an all-miss image there), and
    the FP16 + U16 form gives the same all-miss images for the non-finite poses."""
    p = small_params(64, 0.016)
    g = gpu_volume(p)
    h = gpu_volume(p, fmt=N.FMT_F16_U16)
    o = oracle_volume(p)
    fr = frames(2)
    for mm, T in fr:
        d = oracle.depth_mm(mm)
        o.integrate(d, sc.K640, T)
        g.integrate(d, sc.K640, T)
        h.integrate(d, sc.K640, T)
    T0 = np.asarray(fr[1][1], np.float32)
    cases = {}
    for name, (r, c, val) in {"rot_nan": (0, 1, np.nan), "rot_inf": (2, 2, np.inf), "t_nan": (0, 3, np.nan),
                              "t_inf": (2, 3, -np.inf), "t_far": (0, 3, 1.0e6)}.items():
        T = T0.copy()
        T[r, c] = val
        cases[name] = T
    inside = T0.copy()
    inside[:3, 3] = [0.0, 0.0, 0.8]  # the sphere's centre, inside the volume
    cases["inside"] = inside
    for name, T in cases.items():
        vg, ng = g.raycast(T, sc.K640)
        vo, no = o.raycast(T, sc.K640, 640, 480)
        say("degenerate pose", name, "oracle hits", int((vo[..., 2] != 0).sum()))
        assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), name
        assert np.array_equal(ng.view(np.uint32), no.view(np.uint32)), name
        if name not in ("t_far", "inside"):
            assert not vo.any(), name
            vh, nh = h.raycast(T, sc.K640)
            assert not vh.any() and not nh.any(), name


def test_integrate_degenerate_poses_bitwise():
    """A frame integrated at a degenerate pose (non-finite rotation or translation, a
    camera far outside the volume) leaves the volume as the oracle leaves it, bit for bit
    (FP32), and the FP16 + U16 volume unchanged where the pose is non-finite."""
    p = small_params(64, 0.016)
    g = gpu_volume(p)
    h = gpu_volume(p, fmt=N.FMT_F16_U16)
    o = oracle_volume(p)
    mm, T = frames(1)[0]
    d = oracle.depth_mm(mm)
    o.integrate(d, sc.K640, T)
    g.integrate(d, sc.K640, T)
    h.integrate(d, sc.K640, T)
    T0 = np.asarray(T, np.float32)
    for name, (r, c, val) in {"rot_nan": (1, 0, np.nan), "rot_inf": (0, 0, -np.inf), "t_nan": (1, 3, np.nan),
                              "t_inf": (0, 3, np.inf), "t_far": (2, 3, 1.0e6)}.items():
        Tb = T0.copy()
        Tb[r, c] = val
        th0, wh0 = h.voxelData()
        n_o, _ = o.integrate(d, sc.K640, Tb)
        g.integrate(d, sc.K640, Tb)
        h.integrate(d, sc.K640, Tb)
        tg, wg = g.voxelData()
        to, wo = o.get()
        say("degenerate pose", name, "oracle updates", n_o)
        assert np.array_equal(wg, wo), name
        assert np.array_equal(tg.view(np.uint32), to.view(np.uint32)), name
        if name != "t_far":
            th1, wh1 = h.voxelData()
            assert np.array_equal(wh1, wh0) and np.array_equal(th1.view(np.uint16), th0.view(np.uint16)), name


def test_live_pyramid_bitwise():
    from azu_amd.fusion import ICPTracker
    mm, T = frames(1)[0]
    tr = ICPTracker()
    dmm = N.device_array(mm)
    tr.prepare_mm_device(dmm.ptr)
    ds, vs_, ns = oracle.pyramid(oracle.depth_mm(mm), sc.K640)
    for l in range(3):
        d, v, n = tr.level_host(l)
        assert np.array_equal(d.view(np.uint32), ds[l].view(np.uint32)), l
        assert np.array_equal(v.view(np.uint32), vs_[l].view(np.uint32)), l
        assert np.array_equal(n.view(np.uint32), ns[l].view(np.uint32)), l


def test_icp_single_frame_pose():
    from azu_amd.fusion import ICPTracker
    p = small_params(256, 0.004)
    o = oracle_volume(p)
    fr = frames(2)
    (mm0, T0), (mm1, T1) = fr
    o.integrate(oracle.depth_mm(mm0), sc.K640, T0)
    mv, mn = o.raycast(T0, sc.K640, 640, 480)
    d1 = oracle.depth_mm(mm1)
    ds, vs_, ns = oracle.pyramid(d1, sc.K640)
    ro = oracle.icp_track(vs_, ns, mv, mn, sc.K640, T0, T0)
    tr = ICPTracker()
    dmm = N.device_array(mm1)
    tr.prepare_mm_device(dmm.ptr)
    rg = tr.track(mv, mn, T0, T0)
    say("oracle", ro["inliers"], ro["converged"], ro["final_step"], "gpu", rg["inliers"], rg["converged"],
        rg["final_step"], rg["iterations"])
    ok, dt, ang, same = icp_pose_parity(rg["pose"], rg["iterations"], ro["pose"], ro["steps"],
                                        rg["converged"] and bool(ro["converged"]))
    say("pose diff", dt, ang, "steps gpu", steps_per_level(rg["iterations"]), "oracle",
        steps_per_level(ro["steps"]), "true err", np.abs(rg["pose"][:3, 3] - T1[:3, 3]).max())
    assert ok, (dt, ang, same)
    # counts are taken at the two final poses, which differ by a few micrometres
    assert abs(rg["inliers"] - ro["inliers"]) <= 1e-3 * ro["inliers"]
    assert rg["tracking_ok"] == bool(ro["tracking_ok"])


def test_icp_deterministic():
    """Every block of a launch sums the previous launch's partials in one fixed order, so
    the whole track is deterministic: two tracks of the same input agree bit for bit
    (pose, information matrix, counts)."""
    from azu_amd.fusion import ICPTracker
    p = small_params(256, 0.004)
    o = oracle_volume(p)
    fr = frames(2)
    (mm0, T0), (mm1, _) = fr
    o.integrate(oracle.depth_mm(mm0), sc.K640, T0)
    mv, mn = o.raycast(T0, sc.K640, 640, 480)
    tr = ICPTracker()
    dmm = N.device_array(mm1)
    runs = []
    for _ in range(2):
        tr.prepare_mm_device(dmm.ptr)
        runs.append(tr.track(mv, mn, T0, T0))
    a, b = runs
    say("two tracks: inliers", a["inliers"], b["inliers"], "steps", a["iterations"], b["iterations"])
    assert np.array_equal(np.asarray(a["pose"], np.float32).view(np.uint32),
                          np.asarray(b["pose"], np.float32).view(np.uint32))
    assert np.array_equal(np.asarray(a["information"], np.float32).view(np.uint32),
                          np.asarray(b["information"], np.float32).view(np.uint32))
    assert a["inliers"] == b["inliers"] and a["iterations"] == b["iterations"]
