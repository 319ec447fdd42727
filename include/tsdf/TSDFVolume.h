// kfusion::tsdf::TSDFVolume for the MI355X backend: a header-only C++ host class over the
// C ABI of include/azu_hip.h, shaped like the reference's class
// (include/tsdf/TSDFVolume.h:39-293 of iphands/azu) so a `--backend hip` build can use it
// where the CPU/OpenMP volume was used. Differences from the reference, by design:
//   - the voxels live on the GPU; voxelAt/voxelData download (tests, meshing on the host);
//   - colour is not stored (rgb arguments are accepted and ignored; raycast colours are 0);
//   - a device failure throws std::runtime_error (the reference's HipUniquePtr throws on
//     allocation failure, HipUniquePtr.h:32-42); numeric outcomes never throw.
// Poses: any type with data() returning a column-major 4x4 float array (Eigen::Matrix4f),
// or a const float[16] in that layout. Link with -lazu_hip.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../azu_hip.h"

namespace kfusion {

namespace detail {
inline void check(int rc, const char* what) {
    if (rc != AZU_OK) throw std::runtime_error(std::string(what) + ": " + azu_last_error());
}
inline const float* poseData(const float* p) { return p; }
template <class M>
inline auto poseData(const M& m) -> decltype(m.data()) {
    return m.data();
}
// RAII device buffer (the reference's HipUniquePtr role).
struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t b) { resize(b); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p) azu_free(p);
    }
    void resize(size_t b) {
        if (b <= bytes) return;
        if (p) azu_free(p);
        p = nullptr;
        bytes = 0;
        check(azu_malloc(&p, b), "device allocation");
        bytes = b;
    }
};
}  // namespace detail

namespace tsdf {

inline constexpr float EMPTY_TSDF = 1.0f;
inline constexpr float EMPTY_WEIGHT = 0.0f;

// TSDFParams (reference include/tsdf/TSDFVolume.h:39-53), same defaults.
struct TSDFParams {
    int resolution = 256;
    float voxel_size = 0.010f;
    float truncation = 0.030f;
    float max_weight = 128.0f;
    float origin[3] = {-1.28f, -1.28f, 0.0f};
    float min_depth = 0.30f;
    float max_depth = 5.00f;
};

// Geometry part of the reference's Voxel (TSDFVolume.h:55-62).
struct Voxel {
    float tsdf = EMPTY_TSDF;
    float weight = EMPTY_WEIGHT;
};

// Voxel storage on the device: FP32 tsdf + weight, or binary16 tsdf + uint16 weight.
enum class Storage { F32 = AZU_FMT_F32, F16_U16 = AZU_FMT_F16_U16 };

class TSDFVolume {
public:
    explicit TSDFVolume(const TSDFParams& params = TSDFParams{}, Storage storage = Storage::F32)
        : params_(params), storage_(storage) {
        const azu_tsdf_params p = toC(params_);
        detail::check(azu_vol_create(&p, static_cast<int>(storage_), &h_), "TSDFVolume");
    }
    ~TSDFVolume() { azu_vol_destroy(h_); }
    TSDFVolume(const TSDFVolume&) = delete;
    TSDFVolume& operator=(const TSDFVolume&) = delete;

    const TSDFParams& params() const { return params_; }
    Storage storage() const { return storage_; }

    // Any field difference clears the volume; a resolution change reallocates
    // (TSDFVolume.cpp:124-146).
    void setParams(const TSDFParams& p) {
        const azu_tsdf_params c = toC(p);
        detail::check(azu_vol_set_params(h_, &c), "TSDFVolume::setParams");
        params_ = p;
    }

    void reset() {
        detail::check(azu_vol_reset(h_, stream_), "TSDFVolume::reset");
        detail::check(azu_stream_sync(stream_), "TSDFVolume::reset");
    }

    // TSDFVolume::integrate (TSDFVolume.h:89-96): host depth in metres, uploaded, then the
    // voxel-projective update on the GPU.
    template <class Pose>
    void integrate(const float* depth_meters, const uint8_t* /*rgb*/, const Pose& pose, float fx, float fy,
                   float cx, float cy, int width, int height, float min_depth = 0.3f, float max_depth = 5.0f) {
        const size_t bytes = static_cast<size_t>(width) * height * sizeof(float);
        depth_.resize(bytes);
        detail::check(azu_memcpy_h2d(depth_.p, depth_meters, bytes, stream_), "TSDFVolume::integrate upload");
        integrateGPU(static_cast<const float*>(depth_.p), nullptr, pose, fx, fy, cx, cy, width, height, min_depth,
                     max_depth);
        detail::check(azu_stream_sync(stream_), "TSDFVolume::integrate");
    }

    // TSDFVolume::integrateGPU (TSDFVolume.h:169-176): device depth, asynchronous.
    template <class Pose>
    void integrateGPU(const float* d_depth, const uint8_t* /*d_rgb*/, const Pose& pose, float fx, float fy, float cx,
                      float cy, int width, int height, float min_depth = 0.3f, float max_depth = 5.0f) {
        const float K[4] = {fx, fy, cx, cy};
        detail::check(azu_integrate(h_, d_depth, width, height, K, detail::poseData(pose), min_depth, max_depth,
                                    stream_),
                      "TSDFVolume::integrateGPU");
    }

    // TSDFVolume::raycast (TSDFVolume.h:99-104) into host arrays of width*height xyz
    // triples (Eigen::Vector3f or float[3]); a miss is all zeros.
    template <class Pose, class Vec3>
    void raycast(const Pose& pose, float fx, float fy, float cx, float cy, int width, int height,
                 Vec3* vertices_out, Vec3* normals_out, uint8_t* colors_out = nullptr) {
        static_assert(sizeof(Vec3) == 3 * sizeof(float), "vertices are packed xyz floats");
        const size_t n = static_cast<size_t>(width) * height;
        model_.resize(2 * n * 12);
        float* dv = static_cast<float*>(model_.p);
        float* dn = dv + 3 * n;
        raycastGPU(pose, fx, fy, cx, cy, width, height, dv, dn);
        detail::check(azu_memcpy_d2h(vertices_out, dv, n * 12, stream_), "TSDFVolume::raycast download");
        detail::check(azu_memcpy_d2h(normals_out, dn, n * 12, stream_), "TSDFVolume::raycast download");
        detail::check(azu_stream_sync(stream_), "TSDFVolume::raycast");
        if (colors_out)
            for (size_t i = 0; i < 3 * n; ++i) colors_out[i] = 0;
    }

    // TSDFVolume::raycastGPU (TSDFVolume.h:177-181): device outputs, asynchronous.
    template <class Pose>
    void raycastGPU(const Pose& pose, float fx, float fy, float cx, float cy, int width, int height,
                    float* d_vertices, float* d_normals) {
        const float K[4] = {fx, fy, cx, cy};
        detail::check(azu_raycast(h_, detail::poseData(pose), K, width, height, d_vertices, d_normals, stream_),
                      "TSDFVolume::raycastGPU");
    }

    // Whole-volume download (voxelData, TSDFVolume.h:134), index z*res*res + y*res + x.
    std::vector<Voxel> voxelData() const {
        const size_t n = static_cast<size_t>(params_.resolution) * params_.resolution * params_.resolution;
        std::vector<float> t(n), w(n);
        detail::check(azu_vol_download(h_, t.data(), w.data()), "TSDFVolume::voxelData");
        std::vector<Voxel> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = Voxel{t[i], w[i]};
        return out;
    }
    // TSDFVolume::extractGlobalPointCloud (TSDFVolume.cpp:571-632): the voxels with
    // weight > 1 and |tsdf| < 0.2 at origin + (x, y, z) * voxel_size, in linear voxel
    // order, computed on the GPU. Vec3 is any type built from three floats
    // (Eigen::Vector3f, std::array<float, 3>); colours are zeros (colour is not stored).
    template <class Vec3>
    void extractGlobalPointCloud(std::vector<Vec3>& points_out, std::vector<uint8_t>& colors_out) const {
        unsigned long long n = 0;
        detail::check(azu_vol_point_cloud(h_, nullptr, 0, &n, stream_), "TSDFVolume::extractGlobalPointCloud");
        std::vector<float> xyz(3 * static_cast<size_t>(n));
        if (n > 0) {
            detail::DeviceBuffer d(static_cast<size_t>(n) * 12);
            unsigned long long m = 0;
            detail::check(azu_vol_point_cloud(h_, static_cast<float*>(d.p), n, &m, stream_),
                          "TSDFVolume::extractGlobalPointCloud");
            detail::check(azu_memcpy_d2h(xyz.data(), d.p, static_cast<size_t>(n) * 12, stream_),
                          "TSDFVolume::extractGlobalPointCloud download");
            detail::check(azu_stream_sync(stream_), "TSDFVolume::extractGlobalPointCloud");
        }
        points_out.clear();
        points_out.reserve(static_cast<size_t>(n));
        for (size_t i = 0; i < static_cast<size_t>(n); ++i)
            points_out.push_back(Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
        colors_out.assign(3 * static_cast<size_t>(n), 0);
    }
    Voxel voxelAt(int x, int y, int z) const {
        const size_t r = static_cast<size_t>(params_.resolution);
        return voxelData()[(static_cast<size_t>(z) * r + y) * r + x];
    }
    int integratedFrames() const {
        int f = 0;
        detail::check(azu_vol_integrated_frames(h_, &f), "TSDFVolume::integratedFrames");
        return f;
    }

    azu_volume* handle() const { return h_; }
    void setStream(azu_stream s) { stream_ = s; }
    // Empty-space skipping in the raycast (on by default; the same results either way).
    void setEmptySpaceSkipping(bool on) {
        detail::check(azu_vol_set_skip(h_, on ? 1 : 0), "TSDFVolume::setEmptySpaceSkipping");
    }
    void setGPUEnabled(bool) {}  // always on the GPU
    bool isGPUEnabled() const { return true; }

private:
    static azu_tsdf_params toC(const TSDFParams& p) {
        azu_tsdf_params c;
        c.resolution = p.resolution;
        c.voxel_size = p.voxel_size;
        c.truncation = p.truncation;
        c.max_weight = p.max_weight;
        for (int i = 0; i < 3; ++i) c.origin[i] = p.origin[i];
        c.min_depth = p.min_depth;
        c.max_depth = p.max_depth;
        return c;
    }

    TSDFParams params_;
    Storage storage_;
    azu_volume* h_ = nullptr;
    azu_stream stream_ = nullptr;
    detail::DeviceBuffer depth_, model_;
};

}  // namespace tsdf
}  // namespace kfusion
