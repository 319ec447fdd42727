"""Closed-form pins of the oracle's live frame geometry (back-projection, normals and
their sign, borders, holes, the depth-jump threshold) and of its constant-velocity
estimate. The GPU versions are in tests/test_gpu_geometry_pins.py."""
import numpy as np
import pytest

import oracle
from tests import geometry_pins as gp


def level0(depth, dtype=np.float32):
    _, vs, ns = oracle.pyramid(depth, gp.K640, dtype=dtype)
    return vs[0], ns[0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_tilted_plane_vertices_and_normals(dtype):
    d = gp.plane_depth()
    v, n = level0(d, dtype)
    gp.check_plane(v, n, d)


@pytest.mark.parametrize("d0,step,zeroed", gp.STEP_CASES)
def test_normal_depth_jump_threshold(d0, step, zeroed):
    _, n = level0(gp.step_depth(d0, step))
    gp.check_step(n, zeroed)


def test_normal_hole():
    v, n = level0(gp.hole_depth())
    gp.check_hole(v, n)


def test_check_plane_rejects_wrong_sign_and_transpose():
    """The closed form discriminates: a flipped cross product or swapped image axes
    fail it (the mistakes it is there to catch)."""
    d = gp.plane_depth()
    v, n = level0(d)
    with pytest.raises(AssertionError):
        gp.check_plane(v, -n, d)
    d_t = gp.plane_depth(a=gp.PLANE[2], b=gp.PLANE[1])
    _, n_t = level0(d_t)
    with pytest.raises(AssertionError):
        gp.check_plane(v, n_t, d)


@pytest.mark.parametrize("kind", ["world", "body"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_predict_pose_uniform_motion(kind, dtype):
    tol = 2e-6 if dtype == np.float32 else 1e-12
    for k in (1, 5, 20):
        last, cur, nxt = (gp.uniform_motion(j, kind) for j in (k - 1, k, k + 1))
        p = oracle.predict_pose(cur.astype(dtype), last.astype(dtype), dtype)
        assert np.abs(p - nxt).max() <= tol, (k, np.abs(p - nxt).max())
        # the reversed order (last cur^-1 as the step) misses it by far more
        wrong = cur @ (last @ np.linalg.inv(cur))
        assert np.abs(wrong - nxt).max() > 1e-4
