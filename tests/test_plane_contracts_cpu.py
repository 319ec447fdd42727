"""The reference's fusion-bias and raycast-accuracy contracts on the FP32 oracle
(tests/plane_contracts.py has the fixtures and bars)."""
import numpy as np

import oracle
from tests import plane_contracts as pc


def _volume():
    p = pc.PARAMS
    return oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                         p["min_depth"], p["max_depth"])


def _integrate(v, n):
    d = pc.plane_depth()
    for _ in range(n):
        v.integrate(d, pc.K, np.eye(4), *pc.BAND)


def test_fusion_bias_oracle():
    v = _volume()
    _integrate(v, 10)
    t, w = v.get()
    pos, _, _, _ = oracle.mc_extract(t, w, pc.PARAMS["voxel_size"], pc.PARAMS["origin"])
    pc.check_bias(pos)
    v.close()


def test_weights_count_frames_oracle():
    for n in (1, 3, 70):
        v = _volume()
        _integrate(v, n)
        pc.check_weights(v.get()[1], n)
        v.close()


def test_raycast_accuracy_oracle():
    v = _volume()
    _integrate(v, 5)
    pc.check_front(pc.cast_stats(*v.raycast(pc.POSE_FRONT, pc.K, pc.W, pc.H)))
    pc.check_behind(pc.cast_stats(*v.raycast(pc.POSE_BEHIND, pc.K, pc.W, pc.H)))
    pc.check_oblique(pc.cast_stats(*v.raycast(pc.POSE_OBLIQUE, pc.K, pc.W, pc.H)))
    v.close()
