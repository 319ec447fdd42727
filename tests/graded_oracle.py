"""The reference pipeline's graded tracking loop (src/app/PipelineController.cpp:
1160-1187, 1329-1406, 1500-1590) run on the CPU oracle, as the checker of the GPU's
GradedFrameLoop / replay: constant-velocity estimate, ICP against the model raycast at
its own pose, a retry from the previous pose when the predicted estimate grades Failed,
then Good -> update, integrate and raycast; Poor -> update only; Failed -> hold.

The grade is the reference's TrackingPolicy (include/tracking/TrackingPolicy.h:40-99),
restated here in Python with its default thresholds (independent of the product's C++
restatement, which its own contract cases pin): Failed when the pose or error is not
finite, fewer than 100 inliers, or the motion from the previous pose exceeds 0.15 m or
0.52 rad; Good with at least 2000 inliers, a fit ratio inliers / valid model points of at
least 0.40 and an RMS residual sqrt(error) of at most 1.5 cm; else Poor. The degeneracy
filter (keepObservableMotion) is not part of this checker: compare with the GPU loop run
with it off."""
import numpy as np

import oracle

GOOD, POOR, FAILED = 0, 1, 2


def classify(r, prev):
    pose = np.asarray(r["pose"], np.float64)
    if not np.isfinite(pose).all() or not np.isfinite(r["error"]) or r["inliers"] < 100:
        return FAILED
    d = np.linalg.inv(np.asarray(prev, np.float64)) @ pose
    trans = np.linalg.norm(d[:3, 3])
    angle = np.arccos(np.clip((np.trace(d[:3, :3]) - 1.0) * 0.5, -1.0, 1.0))
    if not (trans <= 0.15 and angle <= 0.52):
        return FAILED
    ratio = r["inliers"] / r["valid_model_points"] if r["valid_model_points"] > 0 else 0.0
    if r["inliers"] >= 2000 and ratio >= 0.40 and np.sqrt(max(0.0, r["error"])) <= 0.015:
        return GOOD
    return POOR


def run(vol, metres, K, pose0, velocity=True):
    """Runs the graded loop over float metre frames (F, H, W) on an oracle Volume.
    Returns per frame (pose 4x4, quality, retried, inliers); frame 0 is integrated at
    pose0 and has quality -1."""
    H, W = metres.shape[1:]
    pose = np.asarray(pose0, np.float32)
    last = pose
    out = []
    mv = mn = model_pose = None
    for k, d in enumerate(metres):
        if k == 0:
            vol.integrate(d, K, pose)
            mv, mn = vol.raycast(pose, K, W, H)
            model_pose = pose
            out.append((pose, -1, False, 0))
            continue
        _, vs, ns = oracle.pyramid(d, K)
        prev = pose
        est = oracle.predict_pose(pose, last) if velocity else pose
        last = pose
        r = oracle.icp_track(vs, ns, mv, mn, K, est, model_pose)
        q = classify(r, prev)
        retried = False
        if velocity and q == FAILED:
            r = oracle.icp_track(vs, ns, mv, mn, K, prev, model_pose)
            q = classify(r, prev)
            retried = True
        if q != FAILED:
            pose = np.asarray(r["pose"], np.float32)
        if q == GOOD:
            vol.integrate(d, K, pose)
            mv, mn = vol.raycast(pose, K, W, H)
            model_pose = pose
        out.append((pose, q, retried, r["inliers"]))
    return out
