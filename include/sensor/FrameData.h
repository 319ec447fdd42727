// kfusion::sensor::FrameData / FramePyramid for the MI355X backend: the host containers
// of a live frame and its pyramid, shaped like the reference's
// include/sensor/FrameData.h:24-78 (iphands/azu) so that code filling them (fixtures,
// recorded frames, tests) hands them to ICPTracker::track unchanged. The GPU builds its
// own pyramid from depth (ICPTracker::prepareLiveGPU); these are the CPU-side form.
#pragma once

#include <cstdint>
#include <vector>

#include "../azu_types.h"

namespace kfusion {
namespace sensor {

static constexpr int FRAME_W = 640;
static constexpr int FRAME_H = 480;

// CameraIntrinsics (include/sensor/CameraIntrinsics.h:17-22), same defaults.
struct CameraIntrinsics {
    float fx = 525.0f;
    float fy = 525.0f;
    float cx = 319.5f;
    float cy = 239.5f;
};
inline constexpr CameraIntrinsics kLegacyIntrinsics{};
static constexpr float FX = kLegacyIntrinsics.fx;
static constexpr float FY = kLegacyIntrinsics.fy;
static constexpr float CX = kLegacyIntrinsics.cx;
static constexpr float CY = kLegacyIntrinsics.cy;

// scaleIntrinsics (CameraIntrinsics.h:33-40): the same camera at w x h instead of
// ref_w x ref_h; pixel centres stay pixel centres.
inline CameraIntrinsics scaleIntrinsics(const CameraIntrinsics& k, int ref_w, int ref_h, int w, int h) {
    if (w == ref_w && h == ref_h) return k;
    const float sx = static_cast<float>(w) / static_cast<float>(ref_w);
    const float sy = static_cast<float>(h) / static_cast<float>(ref_h);
    return CameraIntrinsics{k.fx * sx, k.fy * sy, (k.cx + 0.5f) * sx - 0.5f, (k.cy + 0.5f) * sy - 0.5f};
}

// FrameData (FrameData.h:24-59): per-pixel camera-space vertex and normal (zero =
// invalid), rgb bytes and depth in metres.
struct FrameData {
    std::vector<Vec3f> vertices;
    std::vector<Vec3f> normals;
    std::vector<uint8_t> rgb;
    std::vector<float> depth_meters;
    int width = FRAME_W;
    int height = FRAME_H;
    uint64_t frame_id = 0;
    bool rgb_valid = true;
    double timestamp_ms = 0.0;
    CameraIntrinsics intrinsics{};
    Mat4f pose = Mat4f::Identity();

    FrameData() { resize(FRAME_W, FRAME_H); }
    void resize(int w, int h) {
        width = w;
        height = h;
        const size_t n = static_cast<size_t>(w) * h;
        vertices.assign(n, Vec3f::Zero());
        normals.assign(n, Vec3f::Zero());
        rgb.assign(3 * n, 0);
        depth_meters.assign(n, 0.0f);
    }
    bool isValid(int x, int y) const { return depth_meters[static_cast<size_t>(y) * width + x] > 0.0f; }
};

// FramePyramid (FrameData.h:62-78): levels[0] full resolution, levels[2] coarsest.
struct FramePyramid {
    static constexpr int LEVELS = 3;
    FrameData levels[LEVELS];
    FramePyramid() {
        for (int l = 1; l < LEVELS; ++l) levels[l].resize(FRAME_W >> l, FRAME_H >> l);
    }
};

}  // namespace sensor
}  // namespace kfusion
