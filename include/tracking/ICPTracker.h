// kfusion::tracking::ICPTracker for the MI355X backend: a header-only C++ host class over
// the C ABI of include/azu_hip.h, shaped like the reference's tracker
// (include/tracking/ICPTracker.h:19-134 of iphands/azu). The live pyramid, every ICP
// iteration and the 6x6 solve run on the GPU; the result is read back once per track.
// Poses: any type with data() returning a column-major 4x4 float array (Eigen::Matrix4f),
// or a const float[16] in that layout. Link with -lazu_hip.
#pragma once

#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "../azu_hip.h"
#include "../tsdf/TSDFVolume.h"

namespace kfusion {
namespace tracking {

inline constexpr int PYRAMID_LEVELS = 3;

// CameraIntrinsics (reference include/sensor/CameraIntrinsics.h:17-22), same defaults.
struct CameraIntrinsics {
    float fx = 525.0f;
    float fy = 525.0f;
    float cx = 319.5f;
    float cy = 239.5f;
};

// ICPParams (ICPTracker.h:19-25), same defaults.
struct ICPParams {
    int max_iterations[PYRAMID_LEVELS] = {10, 5, 4};
    float dist_threshold = 0.1f;   // metres
    float angle_threshold = 30.0f; // degrees
    float min_depth = 0.3f;
    float max_depth = 5.0f;
};

// ICPResult (ICPTracker.h:27-47). pose is column-major (Eigen::Map<Eigen::Matrix4f>
// views it); information is the symmetric undamped J^T W J of the last solve.
struct ICPResult {
    float pose[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float error = 0.0f;
    int inliers = 0;
    bool converged = false;
    bool tracking_ok = false;
    float final_step = std::numeric_limits<float>::infinity();
    float information[36] = {};
    int valid_live_points = 0;
    int valid_model_points = 0;
    int projected_points = 0;
    int dist_filtered = 0;
    int angle_filtered = 0;
    int iterations[PYRAMID_LEVELS] = {}; // accepted steps per level (0 = finest)
};

// ModelFrame (ICPTracker.h:49-84) on the device only: width*height packed world-space
// xyz vertices and normals (a miss is all zeros), raycast at `pose`. The buffers are
// the ones TSDFVolume::raycastGPU fills; the frame owns them.
struct ModelFrame {
    int width = 640;
    int height = 480;
    float pose[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    uint64_t source_frame_id = 0;

    explicit ModelFrame(int w = 640, int h = 480) : width(w), height(h) { allocate(); }
    float* d_vertices() const { return static_cast<float*>(buf_.p); }
    float* d_normals() const { return d_vertices() + 3 * static_cast<size_t>(width) * height; }

    // Raycasts the volume into this frame at `pose` (same camera as the live frames,
    // scaled to width x height by the tracker).
    template <class Pose>
    void render(tsdf::TSDFVolume& vol, const Pose& p, const CameraIntrinsics& k) {
        std::memcpy(pose, detail::poseData(p), sizeof(pose));
        vol.raycastGPU(pose, k.fx, k.fy, k.cx, k.cy, width, height, d_vertices(), d_normals());
    }

private:
    void allocate() { buf_.resize(2 * static_cast<size_t>(width) * height * 3 * sizeof(float)); }
    detail::DeviceBuffer buf_;
};

class ICPTracker {
public:
    explicit ICPTracker(const ICPParams& params = ICPParams{}) : params_(params) {}
    ~ICPTracker() { freeGPU(); }
    ICPTracker(const ICPTracker&) = delete;
    ICPTracker& operator=(const ICPTracker&) = delete;

    const ICPParams& params() const { return params_; }
    // Sanitised by the library (ICPTracker::sanitizeParams): NaN or out-of-range angle
    // thresholds clamp to [0, 85] degrees, iteration counts to >= 0.
    void setParams(const ICPParams& p) {
        params_ = p;
        if (h_) {
            const azu_icp_params c = toC(params_);
            detail::check(azu_tracker_set_params(h_, &c), "ICPTracker::setParams");
        }
    }
    void setIntrinsics(const CameraIntrinsics& k) {
        if (k.fx != k_.fx || k.fy != k_.fy || k.cx != k_.cx || k.cy != k_.cy) freeGPU();
        k_ = k;
    }
    const CameraIntrinsics& intrinsics() const { return k_; }
    void setStream(azu_stream s) { stream_ = s; }

    // prepareLiveGPU (ICPTracker.h:108): live depth (metres, device) -> vertex/normal
    // pyramid, with the CPU buildFrameData / computeNormals / buildFramePyramid rules.
    void prepareLiveGPU(const float* d_depth_m, int width, int height) {
        ensure(width, height);
        detail::check(azu_tracker_prepare_m(h_, d_depth_m, stream_), "ICPTracker::prepareLiveGPU");
    }
    // The sensor's raw uint16 millimetre frame instead of metres.
    void prepareLiveGPU(const uint16_t* d_depth_mm, int width, int height) {
        ensure(width, height);
        detail::check(azu_tracker_prepare_mm(h_, d_depth_mm, stream_), "ICPTracker::prepareLiveGPU");
    }
    // The same from device raw Kinect v1 11-bit codes (the reference's sensor
    // boundary, DepthValidity.h:55-93: invalid codes and out-of-band metres -> 0).
    void prepareLiveRawGPU(const uint16_t* d_raw, int width, int height) {
        ensure(width, height);
        detail::check(azu_tracker_prepare_raw(h_, d_raw, stream_), "ICPTracker::prepareLiveRawGPU");
    }
    // Host depth in metres (uploaded, then prepareLiveGPU).
    void prepareLive(const float* depth_m, int width, int height) {
        const size_t bytes = static_cast<size_t>(width) * height * sizeof(float);
        depth_.resize(bytes);
        detail::check(azu_memcpy_h2d(depth_.p, depth_m, bytes, stream_), "ICPTracker::prepareLive upload");
        prepareLiveGPU(static_cast<const float*>(depth_.p), width, height);
    }

    // trackPreparedGPU (ICPTracker.h:109-113): 3-level point-to-plane ICP of the prepared
    // live frame against the model image, from pose_estimate, with model.pose as the
    // reference pose. Synchronises the stream to return the result.
    template <class Pose>
    ICPResult trackPreparedGPU(int width, int height, const ModelFrame& model, const Pose& pose_estimate) {
        ensure(width, height);
        detail::check(azu_tracker_track(h_, model.d_vertices(), model.d_normals(), model.width, model.height,
                                        detail::poseData(pose_estimate), model.pose, stream_),
                      "ICPTracker::trackPreparedGPU");
        return result();
    }
    // Same with an explicit reference pose (the reference's four-argument form).
    template <class Pose, class Ref>
    ICPResult trackPreparedGPU(int width, int height, const ModelFrame& model, const Pose& pose_estimate,
                               const Ref& ref_pose) {
        ensure(width, height);
        detail::check(azu_tracker_track(h_, model.d_vertices(), model.d_normals(), model.width, model.height,
                                        detail::poseData(pose_estimate), detail::poseData(ref_pose), stream_),
                      "ICPTracker::trackPreparedGPU");
        return result();
    }

    // trackGPU (ICPTracker.h:95-101) == prepareLiveGPU + trackPreparedGPU.
    template <class Pose, class Ref>
    ICPResult trackGPU(const float* d_depth_m, const uint8_t* /*d_rgb*/, int width, int height,
                       const ModelFrame& model, const Pose& pose_estimate, const Ref& ref_pose) {
        prepareLiveGPU(d_depth_m, width, height);
        return trackPreparedGPU(width, height, model, pose_estimate, ref_pose);
    }

    // Device pointer to the tracked pose (16 floats, column-major), for callers that
    // keep the per-frame loop on the device (raycast / integrate take it directly).
    const float* devicePose() const {
        const float* p = nullptr;
        detail::check(azu_tracker_pose_device(h_, &p), "ICPTracker::devicePose");
        return p;
    }

    azu_tracker* handle() const { return h_; }
    void initGPU() {}
    void freeGPU() {
        if (h_) azu_tracker_destroy(h_);
        h_ = nullptr;
        w_ = h_h_ = 0;
    }

private:
    static azu_icp_params toC(const ICPParams& p) {
        azu_icp_params c;
        for (int l = 0; l < PYRAMID_LEVELS; ++l) c.max_iterations[l] = p.max_iterations[l];
        c.dist_threshold = p.dist_threshold;
        c.angle_threshold = p.angle_threshold;
        c.min_depth = p.min_depth;
        c.max_depth = p.max_depth;
        return c;
    }
    // The live frame size and intrinsics are fixed per device tracker: (re)create on change.
    void ensure(int width, int height) {
        if (h_ && width == w_ && height == h_h_) return;
        freeGPU();
        const azu_icp_params c = toC(params_);
        const float K[4] = {k_.fx, k_.fy, k_.cx, k_.cy};
        detail::check(azu_tracker_create(width, height, K, &c, &h_), "ICPTracker");
        w_ = width;
        h_h_ = height;
    }
    ICPResult result() {
        azu_icp_result r;
        detail::check(azu_tracker_result(h_, &r, stream_), "ICPTracker::result");
        ICPResult out;
        std::memcpy(out.pose, r.pose, sizeof(out.pose));
        out.error = r.error;
        out.inliers = r.inliers;
        out.converged = r.converged != 0;
        out.tracking_ok = r.tracking_ok != 0;
        out.final_step = r.final_step;
        std::memcpy(out.information, r.information, sizeof(out.information));
        out.valid_live_points = r.valid_live_points;
        out.valid_model_points = r.valid_model_points;
        out.projected_points = r.projected_points;
        out.dist_filtered = r.dist_filtered;
        out.angle_filtered = r.angle_filtered;
        for (int l = 0; l < PYRAMID_LEVELS; ++l) out.iterations[l] = (r.iterations >> (8 * l)) & 0xff;
        return out;
    }

    ICPParams params_;
    CameraIntrinsics k_{};
    azu_tracker* h_ = nullptr;
    int w_ = 0, h_h_ = 0;
    azu_stream stream_ = nullptr;
    detail::DeviceBuffer depth_;
};

}  // namespace tracking
}  // namespace kfusion
