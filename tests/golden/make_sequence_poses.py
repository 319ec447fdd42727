"""Writes the oracle's tracked poses for the 512^3 / 2 mm frame-loop sequence (100 frames,
sphere + plane + 10 boxes, orbit + jitter) as a small fixture:

  c2_512_poses.npz: fp64 (100, 4, 4) poses of the FP64 oracle, fp32 (100, 4, 4) poses
  of the FP32 oracle, truth (100, 4, 4) scene poses, and per frame for both oracles the
  updated-voxel count and the ICP outcome (converged, final step, inliers, steps).

The GPU tests compare the GPU loop's poses against these without running the oracle
(minutes of CPU at 512^3). Regenerate with: python tests/golden/make_sequence_poses.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from azu_amd import scene as sc  # noqa: E402


def main(nframes=100):
    frames, truth, K = sc.sequence("c2_512", frames=nframes)
    p = sc.volume_params("c2_512")
    out = {"truth": truth.astype(np.float64)}
    for name, dt in (("fp32", np.float32), ("fp64", np.float64)):
        v = oracle.Volume(p["resolution"], p["voxel_size"], p["truncation"], p["max_weight"], p["origin"],
                          p["min_depth"], p["max_depth"], dtype=dt)
        t0 = time.time()
        poses, ms, nupd = v.run_sequence(frames, K, truth[0])
        print(name, "%.1f s" % (time.time() - t0), flush=True)
        out[name] = poses.astype(np.float64)
        out[name + "_n_upd"] = nupd
        out[name + "_icp"] = v.last_icp  # per frame: converged, final step, inliers, steps per level
        v.close()
    np.savez_compressed(os.path.join(HERE, "c2_512_poses.npz"), **out)


if __name__ == "__main__":
    main()
