// kfusion::tracking frame-level tracking decision for the MI355X backend: header-only host
// C++ over the ICPResult that ICPTracker::trackGPU reads back once per frame, shaped like
// the reference's include/tracking/TrackingPolicy.h:36-161 (iphands/azu): same names, same
// defaults, same grading. Poses are Eigen::Matrix4f (kfusion::Mat4f) or column-major 4x4
// float arrays; the information matrix is ICPResult::information (6x6, symmetric,
// [t; w] order).
//
//   Good   : finite, enough inliers, tight fit on the model correspondences, motion inside
//            the per-frame gate -> update the pose and integrate
//   Poor   : finite and inside the motion gate, weak fit -> update the pose, do not integrate
//   Failed : anything else -> keep the previous pose
//
// The 6x6 symmetric eigen-decomposition (weakestDirectionRatio, keepObservableMotion) is a
// cyclic Jacobi sweep in FP64; the rotation vector of a rotation matrix goes through the
// unit quaternion (Shepperd's method). Both are host work on 36 numbers per frame.
#pragma once

#include <algorithm>
#include <cmath>
#include <utility>

#include "ICPTracker.h"

namespace kfusion {
namespace tracking {

// Fewest inliers a usable solve has (reference include/tracking/ICPShared.h:76).
inline constexpr int kMinInliersForOk = 100;

enum class TrackQuality { Good, Poor, Failed };

// TrackingPolicy (TrackingPolicy.h:40-48), same defaults.
struct TrackingPolicy {
    int min_inliers = kMinInliersForOk;  // below this: Failed
    int min_good_inliers = 2000;         // Good needs this much support
    float min_fit_ratio = 0.40f;         // inliers / valid model correspondences
    float max_rms_m = 0.015f;            // metres
    float max_frame_translation = 0.15f; // metres between consecutive frames
    float max_frame_rotation = 0.52f;    // radians
    int failures_before_lost = 3;
};

namespace policy_detail {

inline bool allFinite(const float* a, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

// Rigid inverse of a column-major 4x4 [R t; 0 1], in FP64: [R^T, -R^T t].
inline void rigidInverse(const float* m, double* out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[c * 4 + r] = m[r * 4 + c];
    for (int r = 0; r < 3; ++r)
        out[12 + r] = -(out[r] * m[12] + out[4 + r] * m[13] + out[8 + r] * m[14]);
    out[3] = out[7] = out[11] = 0.0;
    out[15] = 1.0;
}

// c = a * b, column-major 4x4, FP64.
inline void mul4(const double* a, const double* b, double* c) {
    for (int col = 0; col < 4; ++col)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[col * 4 + k];
            c[col * 4 + r] = s;
        }
}

// prev^-1 * pose: the frame-to-frame increment, column-major, FP64.
inline void relative(const float* prev, const float* pose, double* d) {
    double pinv[16], p[16];
    rigidInverse(prev, pinv);
    for (int i = 0; i < 16; ++i) p[i] = pose[i];
    mul4(pinv, p, d);
}

// Eigen-decomposition of a symmetric 6x6 (row-major input a) by cyclic Jacobi rotations:
// ev[i] ascending, column i of vec (vec[r * 6 + i]) its unit eigenvector. Returns false
// when the sweeps do not converge.
inline bool symEig6(const double* a_in, double* ev, double* vec) {
    double a[36];
    for (int i = 0; i < 36; ++i) a[i] = a_in[i];
    for (int i = 0; i < 36; ++i) vec[i] = (i % 7 == 0) ? 1.0 : 0.0;
    bool done = false;
    for (int sweep = 0; sweep < 64 && !done; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < 6; ++p) {
            diag += a[p * 6 + p] * a[p * 6 + p];
            for (int q = p + 1; q < 6; ++q) off += a[p * 6 + q] * a[p * 6 + q];
        }
        if (off == 0.0 || off <= 1e-30 * diag) {
            done = true;
            break;
        }
        for (int p = 0; p < 5; ++p)
            for (int q = p + 1; q < 6; ++q) {
                const double apq = a[p * 6 + q];
                if (apq == 0.0) continue;
                // Rotation angle that zeroes a[p][q] (Golub & Van Loan, symmetric Schur).
                const double theta = (a[q * 6 + q] - a[p * 6 + p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; ++k) {
                    const double akp = a[k * 6 + p], akq = a[k * 6 + q];
                    a[k * 6 + p] = c * akp - s * akq;
                    a[k * 6 + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; ++k) {
                    const double apk = a[p * 6 + k], aqk = a[q * 6 + k];
                    a[p * 6 + k] = c * apk - s * aqk;
                    a[q * 6 + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 6; ++k) {
                    const double vkp = vec[k * 6 + p], vkq = vec[k * 6 + q];
                    vec[k * 6 + p] = c * vkp - s * vkq;
                    vec[k * 6 + q] = s * vkp + c * vkq;
                }
            }
    }
    if (!done) return false;
    int order[6] = {0, 1, 2, 3, 4, 5};
    std::sort(order, order + 6, [&](int x, int y) { return a[x * 6 + x] < a[y * 6 + y]; });
    double v2[36];
    for (int i = 0; i < 6; ++i) {
        ev[i] = a[order[i] * 6 + order[i]];
        for (int r = 0; r < 6; ++r) v2[r * 6 + i] = vec[r * 6 + order[i]];
    }
    for (int i = 0; i < 36; ++i) vec[i] = v2[i];
    return true;
}

// Rotation vector (axis * angle, angle in [0, pi]) of the rotation block of a
// column-major 4x4, through the unit quaternion (Shepperd's method).
inline void rotationVector(const double* m, double* w) {
    const double r00 = m[0], r11 = m[5], r22 = m[10];
    const double r10 = m[1], r01 = m[4], r20 = m[2], r02 = m[8], r21 = m[6], r12 = m[9];
    const double tr = r00 + r11 + r22;
    double qw, qx, qy, qz;
    if (tr > 0.0) {
        const double s = std::sqrt(tr + 1.0) * 2.0;
        qw = 0.25 * s;
        qx = (r21 - r12) / s;
        qy = (r02 - r20) / s;
        qz = (r10 - r01) / s;
    } else if (r00 > r11 && r00 > r22) {
        const double s = std::sqrt(1.0 + r00 - r11 - r22) * 2.0;
        qw = (r21 - r12) / s;
        qx = 0.25 * s;
        qy = (r01 + r10) / s;
        qz = (r02 + r20) / s;
    } else if (r11 > r22) {
        const double s = std::sqrt(1.0 + r11 - r00 - r22) * 2.0;
        qw = (r02 - r20) / s;
        qx = (r01 + r10) / s;
        qy = 0.25 * s;
        qz = (r12 + r21) / s;
    } else {
        const double s = std::sqrt(1.0 + r22 - r00 - r11) * 2.0;
        qw = (r10 - r01) / s;
        qx = (r02 + r20) / s;
        qy = (r12 + r21) / s;
        qz = 0.25 * s;
    }
    const double n = std::sqrt(qx * qx + qy * qy + qz * qz);
    if (n < 1e-300) {
        w[0] = w[1] = w[2] = 0.0;
        return;
    }
    const double angle = 2.0 * std::atan2(n, std::fabs(qw));
    const double sgn = qw < 0.0 ? -1.0 : 1.0;
    w[0] = sgn * qx / n * angle;
    w[1] = sgn * qy / n * angle;
    w[2] = sgn * qz / n * angle;
}

}  // namespace policy_detail

// RMS point-to-plane residual implied by ICPResult::error, the mean Huber loss
// (TrackingPolicy.h:52-54).
inline float icpRmsMeters(const ICPResult& r) { return std::sqrt(std::max(0.0f, r.error)); }

// inliers / valid_model_points: the share of live points that landed on a valid model
// pixel and passed the gates (TrackingPolicy.h:58-62).
inline float icpFitRatio(const ICPResult& r) {
    return r.valid_model_points > 0 ? static_cast<float>(r.inliers) / static_cast<float>(r.valid_model_points)
                                    : 0.0f;
}

// The fit alone, without the motion gate (TrackingPolicy.h:69-78).
inline TrackQuality classifyFit(const ICPResult& r, const TrackingPolicy& p = TrackingPolicy{}) {
    if (!policy_detail::allFinite(r.pose.data(), 16) || !std::isfinite(r.error) || r.inliers < p.min_inliers)
        return TrackQuality::Failed;
    if (r.inliers >= p.min_good_inliers && icpFitRatio(r) >= p.min_fit_ratio && icpRmsMeters(r) <= p.max_rms_m)
        return TrackQuality::Good;
    return TrackQuality::Poor;
}

// At most max_frame_translation / max_frame_rotation from prev_pose (TrackingPolicy.h:82-90):
// the increment prev^-1 * pose, its translation norm and the angle acos((trace R - 1) / 2).
inline bool withinFrameMotion(const float* pose, const float* prev_pose, const TrackingPolicy& p = TrackingPolicy{}) {
    double d[16];
    policy_detail::relative(prev_pose, pose, d);
    const double trans = std::sqrt(d[12] * d[12] + d[13] * d[13] + d[14] * d[14]);
    const double c = std::max(-1.0, std::min(1.0, (d[0] + d[5] + d[10] - 1.0) * 0.5));
    const double angle = std::acos(c);
    return trans <= p.max_frame_translation && angle <= p.max_frame_rotation;
}

// The fit, then the per-frame motion gate (TrackingPolicy.h:93-99).
inline TrackQuality classifyTracking(const ICPResult& r, const float* prev_pose,
                                     const TrackingPolicy& p = TrackingPolicy{}) {
    const TrackQuality fit = classifyFit(r, p);
    if (fit == TrackQuality::Failed) return fit;
    if (!withinFrameMotion(r.pose.data(), prev_pose, p)) return TrackQuality::Failed;
    return fit;
}

// Smallest over largest eigenvalue of the information matrix (row-major 6x6, symmetric):
// 0 for an empty, non-finite or unobserved-direction system (TrackingPolicy.h:104-112).
inline float weakestDirectionRatio(const float* information) {
    if (!policy_detail::allFinite(information, 36)) return 0.0f;
    double a[36], ev[6], v[36];
    for (int i = 0; i < 36; ++i) a[i] = information[i];
    if (!policy_detail::symEig6(a, ev, v)) return 0.0f;
    const double max_ev = ev[5];
    if (!(max_ev > 0.0)) return 0.0f;
    return static_cast<float>(std::max(0.0, ev[0]) / max_ev);
}

// ObservableMotion (TrackingPolicy.h:121-124): the pose with unobserved motion removed.
struct ObservableMotion {
    Mat4f pose = Mat4f::Identity();
    int degenerate_dofs = 0;
};

// Keeps only the part of the frame-to-frame motion prev -> est that the information
// matrix observes (TrackingPolicy.h:126-161): the increment as a twist xi = [t; w]
// (translation of prev^-1 * est, rotation vector of its rotation) is projected onto the
// eigenvectors whose eigenvalue is at least rel_threshold times the largest; when any
// direction is dropped the result is prev * exp(kept) (rotation by Rodrigues' formula,
// translation = kept t), otherwise est unchanged.
inline ObservableMotion keepObservableMotion(const float* prev, const float* est, const float* information,
                                             float rel_threshold = 5e-3f) {
    ObservableMotion out;
    for (int i = 0; i < 16; ++i) out.pose.data()[i] = est[i];
    if (!policy_detail::allFinite(information, 36)) return out;
    double a[36], ev[6], v[36];
    for (int i = 0; i < 36; ++i) a[i] = information[i];
    if (!policy_detail::symEig6(a, ev, v)) return out;
    const double max_ev = ev[5];
    if (!(max_ev > 0.0) || !std::isfinite(max_ev)) return out;

    double d[16];
    policy_detail::relative(prev, est, d);
    double xi[6] = {d[12], d[13], d[14], 0.0, 0.0, 0.0};
    policy_detail::rotationVector(d, xi + 3);

    double kept[6] = {};
    for (int i = 0; i < 6; ++i) {
        if (ev[i] >= static_cast<double>(rel_threshold) * max_ev) {
            double dot = 0.0;
            for (int r = 0; r < 6; ++r) dot += v[r * 6 + i] * xi[r];
            for (int r = 0; r < 6; ++r) kept[r] += v[r * 6 + i] * dot;
        } else {
            ++out.degenerate_dofs;
        }
    }
    if (out.degenerate_dofs == 0) return out;

    double step[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const double angle = std::sqrt(kept[3] * kept[3] + kept[4] * kept[4] + kept[5] * kept[5]);
    if (angle > 1e-9) {
        const double kx = kept[3] / angle, ky = kept[4] / angle, kz = kept[5] / angle;
        const double c = std::cos(angle), s = std::sin(angle), t = 1.0 - c;
        step[0] = c + kx * kx * t;
        step[1] = ky * kx * t + kz * s;
        step[2] = kz * kx * t - ky * s;
        step[4] = kx * ky * t - kz * s;
        step[5] = c + ky * ky * t;
        step[6] = kz * ky * t + kx * s;
        step[8] = kx * kz * t + ky * s;
        step[9] = ky * kz * t - kx * s;
        step[10] = c + kz * kz * t;
    }
    step[12] = kept[0];
    step[13] = kept[1];
    step[14] = kept[2];
    double pv[16], res[16];
    for (int i = 0; i < 16; ++i) pv[i] = prev[i];
    policy_detail::mul4(pv, step, res);
    for (int i = 0; i < 16; ++i) out.pose.data()[i] = static_cast<float>(res[i]);
    return out;
}

// The same calls for matrix types with data() giving the column-major array
// (Eigen::Matrix4f, Eigen::Matrix<float, 6, 6>), as the reference's signatures take them.
template <class A, class B, class = decltype(std::declval<const A&>().data()),
          class = decltype(std::declval<const B&>().data())>
inline bool withinFrameMotion(const A& pose, const B& prev_pose, const TrackingPolicy& p = TrackingPolicy{}) {
    return withinFrameMotion(pose.data(), prev_pose.data(), p);
}
template <class M, class = decltype(std::declval<const M&>().data())>
inline TrackQuality classifyTracking(const ICPResult& r, const M& prev_pose, const TrackingPolicy& p = TrackingPolicy{}) {
    return classifyTracking(r, prev_pose.data(), p);
}
template <class M, class = decltype(std::declval<const M&>().data())>
inline float weakestDirectionRatio(const M& information) {
    return weakestDirectionRatio(information.data());
}
template <class P, class E, class M6, class = decltype(std::declval<const P&>().data()),
          class = decltype(std::declval<const E&>().data()), class = decltype(std::declval<const M6&>().data())>
inline ObservableMotion keepObservableMotion(const P& prev, const E& est, const M6& information,
                                             float rel_threshold = 5e-3f) {
    return keepObservableMotion(prev.data(), est.data(), information.data(), rel_threshold);
}

}  // namespace tracking
}  // namespace kfusion
