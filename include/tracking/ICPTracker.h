// kfusion::tracking::ICPTracker for the MI355X backend: a header-only C++ host class over
// the C ABI of include/azu_hip.h, shaped like the reference's tracker
// (include/tracking/ICPTracker.h:19-134 of iphands/azu). The live pyramid, every ICP
// iteration and the 6x6 solve run on the GPU; the result is read back once per track.
// Poses: Eigen::Matrix4f (kfusion::Mat4f), any type whose data() returns a column-major
// 4x4 float array, or a const float[16] in that layout. Link with -lazu_hip.
#pragma once

#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../azu_hip.h"
#include "../azu_types.h"
#include "../sensor/FrameData.h"
#include "../tsdf/TSDFVolume.h"

namespace kfusion {
namespace tracking {

inline constexpr int PYRAMID_LEVELS = sensor::FramePyramid::LEVELS;
using sensor::CameraIntrinsics;

// ICPParams (ICPTracker.h:19-25), same defaults.
struct ICPParams {
    int max_iterations[PYRAMID_LEVELS] = {10, 5, 4};
    float dist_threshold = 0.1f;   // metres
    float angle_threshold = 30.0f; // degrees
    float min_depth = 0.3f;
    float max_depth = 5.0f;
};

// ICPResult (ICPTracker.h:27-47): pose defaults to the identity, final_step to +inf;
// information is the symmetric undamped J^T W J of the last solve ([t; w] order).
struct ICPResult {
    Mat4f pose = Mat4f::Identity();
    float error = 0.0f;
    int inliers = 0;
    bool converged = false;
    bool tracking_ok = false;
    float final_step = std::numeric_limits<float>::infinity();
    Mat6f information = Mat6f::Zero();
    int valid_live_points = 0;
    int valid_model_points = 0;
    int projected_points = 0;
    int dist_filtered = 0;
    int angle_filtered = 0;
    int iterations[PYRAMID_LEVELS] = {}; // accepted steps per level (0 = finest); not in the reference
};

// ModelFrame (ICPTracker.h:49-85): width*height world-space vertices and normals (a miss
// is all zeros) raycast at `pose`, as host vectors and their device mirrors d_vertices /
// d_normals (packed xyz floats). ICPTracker::track reads the host vectors (uploading
// them), trackGPU / trackPreparedGPU the device mirrors; render() fills the device
// mirrors and download() copies them into the host vectors.
struct ModelFrame {
    std::vector<Vec3f> vertices;
    std::vector<Vec3f> normals;
    std::vector<uint8_t> colors;
    int width = sensor::FRAME_W;
    int height = sensor::FRAME_H;
    Mat4f pose = Mat4f::Identity();
    uint64_t source_frame_id = 0;

    explicit ModelFrame(int w = sensor::FRAME_W, int h = sensor::FRAME_H) : width(w), height(h) {
        vertices.assign(static_cast<size_t>(width) * height, Vec3f::Zero());
        normals.assign(static_cast<size_t>(width) * height, Vec3f::Zero());
        colors.assign(static_cast<size_t>(width) * height * 3, 0);
        buf_.resize(2 * static_cast<size_t>(width) * height * 3 * sizeof(float));
    }
    float* d_vertices() const { return static_cast<float*>(buf_.p); }
    float* d_normals() const { return d_vertices() + 3 * static_cast<size_t>(width) * height; }

    // Raycasts the volume into the device mirrors at `p` (intrinsics k, the camera of the
    // live frames, scaled to width x height by the tracker).
    template <class Pose>
    void render(const tsdf::TSDFVolume& vol, const Pose& p, const CameraIntrinsics& k) {
        std::memcpy(pose.data(), detail::poseData(p), 16 * sizeof(float));
        vol.raycastGPU(pose, k.fx, k.fy, k.cx, k.cy, width, height, d_vertices(), d_normals());
    }
    // Host vectors -> device mirrors / device mirrors -> host vectors.
    void upload(azu_stream s = nullptr) {
        const size_t b = static_cast<size_t>(width) * height * 12;
        detail::check(azu_memcpy_h2d(d_vertices(), vertices.data(), b, s), "ModelFrame::upload");
        detail::check(azu_memcpy_h2d(d_normals(), normals.data(), b, s), "ModelFrame::upload");
        detail::check(azu_stream_sync(s), "ModelFrame::upload");
    }
    void download(azu_stream s = nullptr) {
        const size_t b = static_cast<size_t>(width) * height * 12;
        detail::check(azu_stream_sync(s), "ModelFrame::download");
        detail::check(azu_memcpy_d2h(vertices.data(), d_vertices(), b, s), "ModelFrame::download");
        detail::check(azu_memcpy_d2h(normals.data(), d_normals(), b, s), "ModelFrame::download");
        detail::check(azu_stream_sync(s), "ModelFrame::download");
    }

private:
    detail::DeviceBuffer buf_;
};

class ICPTracker {
public:
    explicit ICPTracker(const ICPParams& params = ICPParams{}) : params_(sanitizeParams(params)) {}
    ~ICPTracker() { freeGPU(); }
    ICPTracker(const ICPTracker&) = delete;
    ICPTracker& operator=(const ICPTracker&) = delete;

    // params() / setParams (ICPTracker.h:128-134): sanitised as ICPTracker::sanitizeParams
    // does (src/tracking/ICPTracker.cpp:24-34): the angle threshold clamps to [0, 85]
    // degrees, a non-finite one becomes 30.
    const ICPParams& params() const { return params_; }
    void setParams(const ICPParams& p) {
        params_ = sanitizeParams(p);
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
    void setNumThreads(int) {}  // the reductions are fixed-order on the GPU

    // track (ICPTracker.h:91-94, ICPTracker.cpp:40-134): 3-level point-to-plane ICP of a
    // host live pyramid against a host model image (ModelFrame's vectors), from
    // pose_estimate, with ref_pose as the reference pose. The pyramid levels and the
    // model are uploaded; the iterations run on the GPU.
    template <class Pose, class Ref>
    ICPResult track(const sensor::FramePyramid& live, ModelFrame& model, const Pose& pose_estimate,
                    const Ref& ref_pose) {
        const sensor::FrameData& l0 = live.levels[0];
        ensure(l0.width, l0.height);
        for (int l = 0; l < PYRAMID_LEVELS; ++l) {
            const sensor::FrameData& f = live.levels[l];
            const size_t n = static_cast<size_t>(f.width) * f.height;
            stage_.resize(n * 28);
            float* dd = static_cast<float*>(stage_.p);
            float* dv = dd + n;
            float* dn = dv + 3 * n;
            detail::check(azu_memcpy_h2d(dd, f.depth_meters.data(), n * 4, stream_), "ICPTracker::track upload");
            detail::check(azu_memcpy_h2d(dv, f.vertices.data(), n * 12, stream_), "ICPTracker::track upload");
            detail::check(azu_memcpy_h2d(dn, f.normals.data(), n * 12, stream_), "ICPTracker::track upload");
            detail::check(azu_tracker_load_level(h_, l, dd, dv, dn, stream_), "ICPTracker::track");
            detail::check(azu_stream_sync(stream_), "ICPTracker::track");
        }
        model.upload(stream_);
        detail::check(azu_tracker_track(h_, model.d_vertices(), model.d_normals(), model.width, model.height,
                                        detail::poseData(pose_estimate), detail::poseData(ref_pose), stream_),
                      "ICPTracker::track");
        return result();
    }

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
    // live frame against the model image's device mirrors, from pose_estimate, with
    // ref_pose as the reference pose (the model's own pose, ICPTracker.h:68-73, when
    // omitted). Synchronises the stream to return the result.
    template <class Pose, class Ref>
    ICPResult trackPreparedGPU(int width, int height, const ModelFrame& model, const Pose& pose_estimate,
                               const Ref& ref_pose) {
        ensure(width, height);
        detail::check(azu_tracker_track(h_, model.d_vertices(), model.d_normals(), model.width, model.height,
                                        detail::poseData(pose_estimate), detail::poseData(ref_pose), stream_),
                      "ICPTracker::trackPreparedGPU");
        return result();
    }
    template <class Pose>
    ICPResult trackPreparedGPU(int width, int height, const ModelFrame& model, const Pose& pose_estimate) {
        return trackPreparedGPU(width, height, model, pose_estimate, model.pose);
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

    static ICPParams sanitizeParams(const ICPParams& p) {
        const azu_icp_params in = toC(p);
        azu_icp_params out;
        detail::check(azu_icp_params_sanitize(&in, &out), "ICPTracker::sanitizeParams");
        ICPParams o = p;
        o.angle_threshold = out.angle_threshold;
        return o;
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
        std::memcpy(out.pose.data(), r.pose, sizeof(r.pose));
        out.error = r.error;
        out.inliers = r.inliers;
        out.converged = r.converged != 0;
        out.tracking_ok = r.tracking_ok != 0;
        out.final_step = r.final_step;
        // row-major from the ABI; the matrix is symmetric, so the storage order is moot
        std::memcpy(out.information.data(), r.information, sizeof(r.information));
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
    detail::DeviceBuffer depth_, stage_;
};

}  // namespace tracking
}  // namespace kfusion
