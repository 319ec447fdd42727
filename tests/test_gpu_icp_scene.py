"""The reference's reference-pose and low-resolution-model ICP contracts
(tests/icp_refpose_contract.cpp:33-90, tests/icp_lowres_model_contract.cpp:40-136) on the
GPU: the same closed-loop checks against the synthetic room's true poses, and every
track compared with the oracle's on the same inputs by the ICP parity rule, including
models of 320 x 240 and 160 x 120 (the tracker projects into them with scaleIntrinsics,
src/tracking/ICPTracker.cpp:158-159)."""
import numpy as np
import pytest

from azu_amd import native as N
from azu_amd.fusion import ICPParams, ICPTracker, TSDFParams, TSDFVolume
from tests import icp_scene_contract as ic
from tests import room_scene as rs
from tests.gpu_util import icp_pose_parity, say
from tests.test_icp_scene_cpu import OracleOps

pytestmark = pytest.mark.gpu


class GpuOps:
    def __init__(self):
        self.vol = TSDFVolume(TSDFParams(**rs.VOLUME), fmt=N.FMT_F32)

    def fuse(self, depth, pose, n):
        for _ in range(n):
            self.vol.integrate(depth, ic.K, pose, 0.3, 3.0)

    def model(self, pose, w, h, K):
        return self.vol.raycast(pose, K, w, h)

    def track(self, depth, model, est, ref, iters=(10, 5, 4), dist=0.1, angle=30.0):
        tr = ICPTracker(640, 480, ic.K, ICPParams(max_iterations=iters, dist_threshold=dist, angle_threshold=angle,
                                                  min_depth=0.3, max_depth=3.0))
        d = N.device_array(np.asarray(depth, np.float32))
        tr.prepare_m_device(d.ptr)
        r = tr.track(model[0], model[1], est, ref)
        tr.close()
        return r


class Paired:
    """Runs every operation on the GPU and on the oracle; models and poses must agree."""

    def __init__(self):
        self.g, self.o = GpuOps(), OracleOps()
        self.pairs = []

    def fuse(self, depth, pose, n):
        self.g.fuse(depth, pose, n)
        self.o.fuse(depth, pose, n)

    def model(self, pose, w, h, K):
        mg, mo = self.g.model(pose, w, h, K), self.o.model(pose, w, h, K)
        assert np.array_equal(mg[0].view(np.uint32), mo[0].view(np.uint32)), (w, h)
        assert np.array_equal(mg[1].view(np.uint32), mo[1].view(np.uint32)), (w, h)
        return mg

    def track(self, depth, model, est, ref, **kw):
        rg = self.g.track(depth, model, est, ref, **kw)
        ro = self.o.track(depth, model, est, ref, **kw)
        ok, dt, ang, same = icp_pose_parity(rg["pose"], rg["iterations"], ro["pose"], ro["steps"],
                                            rg["converged"] and bool(ro["converged"]))
        say("model %dx%d: GPU vs oracle %.2e m %.2e rad, same steps %s, inliers %d / %d" % (
            model[0].shape[1], model[0].shape[0], dt, ang, same, rg["inliers"], ro["inliers"]))
        self.pairs.append(ok)
        assert ok, (dt, ang, same)
        assert rg["tracking_ok"] == ro["tracking_ok"]
        return rg


def test_refpose_contract_gpu():
    p = Paired()
    ic.refpose(p)
    assert len(p.pairs) == 2


def test_lowres_model_contract_gpu():
    p = Paired()
    ic.lowres(p)
    assert len(p.pairs) == 4
