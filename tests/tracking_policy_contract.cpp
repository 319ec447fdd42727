// Known-answer cases of the frame-level tracking policy (include/tracking/TrackingPolicy.h),
// the cases the reference's tests/tracking_policy_contract.cpp:51-105 grades, plus the
// degeneracy filter on diagonal information matrices whose kept motion is known in closed
// form. Built and run by tests/test_tracking_policy_cpu.py; prints one line per failure
// and exits non-zero on any.
#include "tracking/TrackingPolicy.h"

#include <cmath>
#include <cstdio>
#include <limits>
#include <string>

namespace {

int g_checks = 0, g_failures = 0;

void check(bool ok, const std::string& what, int line) {
    ++g_checks;
    if (!ok) {
        std::printf("FAIL line %d: %s\n", line, what.c_str());
        ++g_failures;
    }
}
#define EXPECT(cond, what) check((cond), (what), __LINE__)

using namespace kfusion::tracking;

// Column-major 4x4 with data(), standing in for Eigen::Matrix4f.
struct Mat4 {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float* data() const { return m; }
    float& at(int r, int c) { return m[c * 4 + r]; }
    float at(int r, int c) const { return m[c * 4 + r]; }
};

Mat4 operator*(const Mat4& a, const Mat4& b) {
    Mat4 c;
    for (int r = 0; r < 4; ++r)
        for (int col = 0; col < 4; ++col) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a.at(r, k) * b.at(k, col);
            c.at(r, col) = s;
        }
    return c;
}

// Rotation by yaw about the camera y axis, then translation tx along x.
Mat4 yawMove(float tx, float yaw) {
    Mat4 t;
    const float c = std::cos(yaw), s = std::sin(yaw);
    t.at(0, 0) = c;
    t.at(0, 2) = s;
    t.at(2, 0) = -s;
    t.at(2, 2) = c;
    t.at(0, 3) = tx;
    return t;
}

Mat4 rollMove(float tx, float ty, float tz, float roll) {
    Mat4 t;
    const float c = std::cos(roll), s = std::sin(roll);
    t.at(0, 0) = c;
    t.at(0, 1) = -s;
    t.at(1, 0) = s;
    t.at(1, 1) = c;
    t.at(0, 3) = tx;
    t.at(1, 3) = ty;
    t.at(2, 3) = tz;
    return t;
}

ICPResult result(const Mat4& pose, int inliers, int valid_model, float error, int valid_live = 300000) {
    ICPResult r;
    for (int i = 0; i < 16; ++i) r.pose.data()[i] = pose.m[i];
    r.inliers = inliers;
    r.valid_model_points = valid_model;
    r.valid_live_points = valid_live;
    r.error = error;
    r.converged = false;  // grading ignores convergence
    r.tracking_ok = false;
    return r;
}

struct Mat6 {
    float m[36] = {};
    const float* data() const { return m; }
};

Mat6 diag6(float a, float b, float c, float d, float e, float f) {
    Mat6 i;
    const float v[6] = {a, b, c, d, e, f};
    for (int k = 0; k < 6; ++k) i.m[k * 7] = v[k];
    return i;
}

float maxAbsDiff(const float* a, const float* b) {
    float m = 0.0f;
    for (int i = 0; i < 16; ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return m;
}

}  // namespace

int main() {
    const Mat4 prev = yawMove(0.3f, 0.2f);
    const Mat4 near = prev * yawMove(0.02f, 0.03f);

    // classifyTracking: fit grades and the motion gate.
    EXPECT(classifyTracking(result(near, 150000, 200000, 4e-6f), prev) == TrackQuality::Good,
           "tight fit, small motion, not converged: Good");
    EXPECT(classifyTracking(result(near, 30000, 200000, 4e-6f), prev) == TrackQuality::Poor,
           "fit ratio 0.15 < 0.40: Poor");
    EXPECT(classifyTracking(result(near, 40000, 45000, 4e-6f, 300000), prev) == TrackQuality::Good,
           "fit ratio counts model correspondences, not live points: Good");
    EXPECT(classifyTracking(result(near, 1500, 1600, 4e-6f), prev) == TrackQuality::Poor,
           "under 2000 inliers: Poor");
    EXPECT(classifyTracking(result(near, 150000, 200000, 9e-4f), prev) == TrackQuality::Poor,
           "RMS 3 cm > 1.5 cm: Poor");
    EXPECT(classifyTracking(result(prev * yawMove(0.2f, 0.0f), 150000, 200000, 4e-6f), prev) ==
               TrackQuality::Failed,
           "20 cm in one frame: Failed");
    EXPECT(classifyTracking(result(prev * yawMove(0.0f, 0.7f), 150000, 200000, 4e-6f), prev) ==
               TrackQuality::Failed,
           "40 degrees in one frame: Failed");
    EXPECT(classifyTracking(result(near, 99, 200000, 4e-6f), prev) == TrackQuality::Failed,
           "under 100 inliers: Failed");
    Mat4 bad = near;
    bad.at(0, 3) = std::numeric_limits<float>::quiet_NaN();
    EXPECT(classifyTracking(result(bad, 150000, 200000, 4e-6f), prev) == TrackQuality::Failed,
           "non-finite pose: Failed");
    EXPECT(classifyTracking(result(near, 150000, 200000, std::numeric_limits<float>::infinity()), prev) ==
               TrackQuality::Failed,
           "non-finite error: Failed");
    EXPECT(TrackingPolicy{}.failures_before_lost == 3, "three consecutive failures before Lost");

    // classifyFit grades without the motion gate.
    const Mat4 far = prev * yawMove(0.45f, 0.7f);
    EXPECT(classifyFit(result(far, 150000, 200000, 4e-6f)) == TrackQuality::Good, "classifyFit: far but tight: Good");
    EXPECT(classifyTracking(result(far, 150000, 200000, 4e-6f), prev) == TrackQuality::Failed,
           "classifyTracking: far fails the motion gate");
    EXPECT(classifyFit(result(far, 30000, 200000, 4e-6f)) == TrackQuality::Poor, "classifyFit: weak fit: Poor");
    EXPECT(classifyFit(result(far, 99, 200000, 4e-6f)) == TrackQuality::Failed, "classifyFit: under 100 inliers");
    EXPECT(withinFrameMotion(near, prev) && !withinFrameMotion(far, prev), "withinFrameMotion: near in, far out");

    // Gate edges: 0.14 m and 0.16 m of pure translation; 0.50 and 0.54 rad of pure yaw.
    EXPECT(withinFrameMotion(prev * yawMove(0.14f, 0.0f), prev), "0.14 m inside the 0.15 m gate");
    EXPECT(!withinFrameMotion(prev * yawMove(0.16f, 0.0f), prev), "0.16 m outside the 0.15 m gate");
    EXPECT(withinFrameMotion(prev * yawMove(0.0f, 0.50f), prev), "0.50 rad inside the 0.52 rad gate");
    EXPECT(!withinFrameMotion(prev * yawMove(0.0f, 0.54f), prev), "0.54 rad outside the 0.52 rad gate");

    // icpFitRatio / icpRmsMeters.
    EXPECT(icpFitRatio(result(near, 10, 0, 0.0f)) == 0.0f, "fit ratio of no model correspondences is 0");
    EXPECT(std::fabs(icpRmsMeters(result(near, 10, 10, 4e-6f)) - 2e-3f) < 1e-9f, "RMS of 4e-6 is 2 mm");
    EXPECT(icpRmsMeters(result(near, 10, 10, -1.0f)) == 0.0f, "negative error clamps to 0");

    // weakestDirectionRatio.
    Mat6 info = diag6(100, 100, 100, 100, 100, 1);
    EXPECT(std::fabs(weakestDirectionRatio(info) - 0.01f) < 1e-5f, "weakestDirectionRatio: 1/100");
    info.m[35] = 0.0f;
    EXPECT(weakestDirectionRatio(info) == 0.0f, "weakestDirectionRatio: an unobserved direction is 0");
    EXPECT(weakestDirectionRatio(Mat6{}) == 0.0f, "weakestDirectionRatio: an empty system is 0");
    info.m[7] = std::numeric_limits<float>::quiet_NaN();
    EXPECT(weakestDirectionRatio(info) == 0.0f, "weakestDirectionRatio: a non-finite system is 0");

    // keepObservableMotion: full rank keeps the estimate as it is.
    const Mat4 est = prev * rollMove(0.03f, -0.02f, 0.01f, 0.05f);
    ObservableMotion om = keepObservableMotion(prev, est, diag6(50, 60, 70, 80, 90, 100));
    EXPECT(om.degenerate_dofs == 0 && maxAbsDiff(om.pose.data(), est.m) == 0.0f, "full rank: estimate unchanged");
    // Roll about the optical axis unobserved (w_z): the kept motion is the translation.
    om = keepObservableMotion(prev, est, diag6(100, 100, 100, 100, 100, 0.1f));
    EXPECT(om.degenerate_dofs == 1, "one unobserved direction");
    const Mat4 want_t = prev * rollMove(0.03f, -0.02f, 0.01f, 0.0f);
    EXPECT(maxAbsDiff(om.pose.data(), want_t.m) < 2e-6f, "unobserved roll: translation kept, roll dropped");
    // Sliding along x unobserved (t_x): the kept motion is the roll and the other translation.
    om = keepObservableMotion(prev, est, diag6(0.2f, 100, 100, 100, 100, 100));
    const Mat4 want_x = prev * rollMove(0.0f, -0.02f, 0.01f, 0.05f);
    EXPECT(om.degenerate_dofs == 1 && maxAbsDiff(om.pose.data(), want_x.m) < 2e-6f, "unobserved slide: x dropped");
    // Every direction below the threshold: the previous pose.
    om = keepObservableMotion(prev, est, diag6(1e-4f, 1e-4f, 1e-4f, 1e-4f, 1e-4f, 1));
    const Mat4 want_roll = prev * rollMove(0.0f, 0.0f, 0.0f, 0.05f);
    EXPECT(om.degenerate_dofs == 5 && maxAbsDiff(om.pose.data(), want_roll.m) < 2e-6f,
           "only the roll observed: roll kept, translation dropped");
    // A non-finite or empty system leaves the estimate as it is.
    om = keepObservableMotion(prev, est, Mat6{});
    EXPECT(om.degenerate_dofs == 0 && maxAbsDiff(om.pose.data(), est.m) == 0.0f, "empty system: estimate unchanged");

    if (g_failures == 0) {
        std::printf("tracking_policy_contract: PASS (%d checks)\n", g_checks);
        return 0;
    }
    std::printf("tracking_policy_contract: FAIL (%d of %d checks)\n", g_failures, g_checks);
    return 1;
}
