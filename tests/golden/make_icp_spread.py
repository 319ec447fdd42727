"""Writes the FP32 oracle's own ICP spread over summation orders for the 512^3 / 2 mm
sequence, open loop, as a small fixture:

  c2_512_icp_spread.npz: spread (100, 2) = per frame k >= 1, the largest (translation m,
  rotation rad) difference between the oracle's ICP result summed with 8 ordered chunks
  (the golden poses' order) and with 1, 2, 4, 16, 32 or 64 chunks, all from the same
  inputs: the live frame k, and the model raycast at the oracle's pose of frame k - 1 from
  a volume integrated at the oracle's poses of frames 0..k-1 (tests/golden/
  c2_512_poses.npz). The reference folds its OpenMP threads' partial sums in order, so
  these are the results it gives with that many threads: the frame's intrinsic
  sensitivity to the summation order. Row 0 (no ICP) is zero.

Regenerate with: python tests/golden/make_icp_spread.py (about two minutes on 8 cores).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from azu_amd import scene as sc  # noqa: E402
from tests.gpu_util import pose_diff  # noqa: E402

CHUNKS = (1, 2, 4, 16, 32, 64)


def main(nframes=100):
    g = np.load(os.path.join(HERE, "c2_512_poses.npz"))
    frames, _, K = sc.sequence("c2_512", frames=nframes)
    p = sc.volume_params("c2_512")
    o = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                      p["min_depth"], p["max_depth"])
    P = g["fp32"].astype(np.float32)
    spread = np.zeros((nframes, 2))
    mv = mn = None
    for k in range(nframes):
        d = oracle.depth_mm(frames[k])
        if k > 0:
            _, vs, ns = oracle.pyramid(d, K)
            oracle.lib.orc_set_icp_chunks(8)
            r8 = oracle.icp_track(vs, ns, mv, mn, K, P[k - 1], P[k - 1])
            assert np.array_equal(np.asarray(r8["pose"], np.float32), P[k]), k  # the golden pose, reproduced
            for c in CHUNKS:
                oracle.lib.orc_set_icp_chunks(c)
                dt, ang = pose_diff(oracle.icp_track(vs, ns, mv, mn, K, P[k - 1], P[k - 1])["pose"], r8["pose"])
                spread[k] = np.maximum(spread[k], (dt, ang))
            oracle.lib.orc_set_icp_chunks(8)
            print(k, "spread %.2e m %.2e rad" % tuple(spread[k]), flush=True)
        o.integrate(d, K, P[k])
        mv, mn = o.raycast(P[k], K, 640, 480)
    o.close()
    np.savez_compressed(os.path.join(HERE, "c2_512_icp_spread.npz"), spread=spread)


if __name__ == "__main__":
    main()
