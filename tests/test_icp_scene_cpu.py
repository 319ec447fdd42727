"""The reference's reference-pose and low-resolution-model ICP contracts on the oracle
(closed-loop checks against the synthetic room's true poses), and scaleIntrinsics."""
import numpy as np
import pytest

import oracle
from tests import icp_scene_contract as ic
from tests import room_scene as rs


class OracleOps:
    def __init__(self):
        p = rs.VOLUME
        self.vol = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                                 p["min_depth"], p["max_depth"])

    def fuse(self, depth, pose, n):
        for _ in range(n):
            self.vol.integrate(depth, ic.K, pose, 0.3, 3.0)

    def model(self, pose, w, h, K):
        return self.vol.raycast(pose, K, w, h)

    def track(self, depth, model, est, ref, iters=(10, 5, 4), dist=0.1, angle=30.0):
        _, vs, ns = oracle.pyramid(np.asarray(depth, np.float32), ic.K)
        r = oracle.icp_track(vs, ns, model[0], model[1], ic.K, est, ref, iters=iters, dist=dist, angle=angle)
        r["tracking_ok"] = bool(r["tracking_ok"])
        return r


@pytest.mark.slow
def test_refpose_contract_oracle():
    ic.refpose(OracleOps())


@pytest.mark.slow
def test_lowres_model_contract_oracle():
    ic.lowres(OracleOps())


def test_scale_intrinsics_contract():
    """icp_lowres_model_contract.cpp:47-56: identical at the same size; half and quarter."""
    K = ic.K
    assert rs.scale_intrinsics(K, 640, 480, 640, 480) == K
    h = rs.scale_intrinsics(K, 640, 480, 320, 240)
    f = np.float32
    assert h[0] == f(K[0]) * f(0.5) and h[2] == (f(K[2]) + f(0.5)) * f(0.5) - f(0.5)
    assert h[3] == (f(K[3]) + f(0.5)) * f(0.5) - f(0.5)
    q = rs.scale_intrinsics(K, 640, 480, 160, 120)
    assert q[1] == f(K[1]) * f(0.25) and q[2] == (f(K[2]) + f(0.5)) * f(0.25) - f(0.5)
