// kfusion::tsdf::TSDFVolume for the MI355X backend: a header-only C++ host class over the
// C ABI of include/azu_hip.h, shaped like the reference's class
// (include/tsdf/TSDFVolume.h:39-293 of iphands/azu) so a `--backend hip` build can use it
// where the CPU/OpenMP volume was used. The voxels live on the GPU.
//
// Host access (voxelData, voxelAt) goes through a host mirror, the reference's own model
// for its GPU volume (syncFromGPU / syncToGPU, TSDFVolume.h:161-162): the first host
// access after a GPU write downloads the volume once; later reads are plain memory reads
// until the GPU writes again. A non-const voxelData() / voxelAt() may write the mirror,
// so the mirror is uploaded before the next GPU operation (integrate, raycast, meshing,
// point cloud) or by syncToGPU(). readVoxel() / writeVoxel() touch one voxel on the
// device without a mirror.
//
// Differences from the reference, by design:
//   - colour is not stored on the device (rgb arguments are accepted and ignored; raycast
//     colours are 0; the mirror's r, g, b read EMPTY_COLOR after a download and colour
//     written into the mirror is not uploaded);
//   - a device failure throws std::runtime_error (the reference's HipUniquePtr throws on
//     allocation failure, HipUniquePtr.h:32-42); numeric outcomes never throw.
// Poses: Eigen::Matrix4f (kfusion::Mat4f), any type whose data() returns a column-major
// 4x4 float array, or a const float[16] in that layout. Link with -lazu_hip.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../azu_hip.h"
#include "../azu_types.h"

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
inline constexpr float EMPTY_COLOR = 128.0f / 255.0f;

// TSDFParams (reference include/tsdf/TSDFVolume.h:39-53), same fields and defaults.
struct TSDFParams {
    int resolution = 256;
    float voxel_size = 0.010f;
    float truncation = 0.030f;
    float max_weight = 128.0f;
    Vec3f origin = Vec3f(-1.28f, -1.28f, 0.0f);
    float min_depth = 0.30f;
    float max_depth = 5.00f;
};

// Voxel (TSDFVolume.h:55-62): the empty state is tsdf 1, weight 0, neutral colour.
struct Voxel {
    float tsdf = EMPTY_TSDF;
    float weight = EMPTY_WEIGHT;
    float r = EMPTY_COLOR, g = EMPTY_COLOR, b = EMPTY_COLOR;
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
    // (TSDFVolume.cpp:124-146). Equal parameters keep every voxel, the mirror included.
    void setParams(const TSDFParams& p) {
        const azu_tsdf_params c = toC(p);
        flush();
        detail::check(azu_vol_set_params(h_, &c), "TSDFVolume::setParams");
        const azu_tsdf_params old = toC(params_);
        params_ = p;
        if (!sameParams(old, c)) dropMirror();
    }

    void reset() {
        detail::check(azu_vol_reset(h_, stream_), "TSDFVolume::reset");
        detail::check(azu_stream_sync(stream_), "TSDFVolume::reset");
        dropMirror();
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
        flush();
        detail::check(azu_integrate(h_, d_depth, width, height, K, detail::poseData(pose), min_depth, max_depth,
                                    stream_),
                      "TSDFVolume::integrateGPU");
        dropMirror();
    }

    // TSDFVolume::raycast (TSDFVolume.h:99-104) into host arrays of width*height xyz
    // triples (Eigen::Vector3f, kfusion::Vec3f or float[3]); a miss is all zeros. Colours,
    // when asked for, are zeros (colour is not stored).
    template <class Pose, class Vec3>
    void raycast(const Pose& pose, float fx, float fy, float cx, float cy, int width, int height,
                 Vec3* vertices_out, Vec3* normals_out, uint8_t* colors_out = nullptr) const {
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

    // TSDFVolume::raycastGPU (TSDFVolume.h:177-181): device outputs (float3 / packed xyz
    // floats), asynchronous. d_colors (uchar3, 3 bytes per pixel), when given, is zeroed.
    template <class Pose, class F3, class C3 = unsigned char>
    void raycastGPU(const Pose& pose, float fx, float fy, float cx, float cy, int width, int height, F3* d_vertices,
                    F3* d_normals, C3* d_colors = nullptr) const {
        static_assert(sizeof(F3) == 12 || sizeof(F3) == 4, "float3 or float");
        const float K[4] = {fx, fy, cx, cy};
        const_cast<TSDFVolume*>(this)->flush();
        detail::check(azu_raycast(h_, detail::poseData(pose), K, width, height, reinterpret_cast<float*>(d_vertices),
                                  reinterpret_cast<float*>(d_normals), stream_),
                      "TSDFVolume::raycastGPU");
        if (d_colors) {
            const size_t bytes = static_cast<size_t>(width) * height * 3;
            if (zeros_.size() < bytes) zeros_.assign(bytes, 0);
            detail::check(azu_memcpy_h2d(d_colors, zeros_.data(), bytes, stream_), "TSDFVolume::raycastGPU colours");
            detail::check(azu_stream_sync(stream_), "TSDFVolume::raycastGPU colours");
        }
    }

    // voxelData (TSDFVolume.h:134): the host mirror, index z*res*res + y*res + x,
    // downloaded when the GPU has written since the last download.
    const std::vector<Voxel>& voxelData() const {
        const_cast<TSDFVolume*>(this)->refresh();
        return mirror_;
    }
    // Mutable form: writes into the returned voxels reach the GPU before its next
    // operation on this volume (or at syncToGPU()).
    std::vector<Voxel>& voxelData() {
        refresh();
        host_dirty_ = true;
        return mirror_;
    }
    // voxelAt (TSDFVolume.h:117-118). PRECONDITION: 0 <= x, y, z < resolution.
    const Voxel& voxelAt(int x, int y, int z) const { return voxelData()[index(x, y, z)]; }
    Voxel& voxelAt(int x, int y, int z) { return voxelData()[index(x, y, z)]; }

    // One voxel read from / written to the device storage, without the mirror (pending
    // host writes of the mirror are uploaded first; a write updates a current mirror).
    Voxel readVoxel(int x, int y, int z) const {
        const_cast<TSDFVolume*>(this)->flush();
        Voxel v;
        detail::check(azu_vol_read_voxels(h_, index(x, y, z), 1, &v.tsdf, &v.weight), "TSDFVolume::readVoxel");
        return v;
    }
    void writeVoxel(int x, int y, int z, float tsdf, float weight) {
        flush();
        detail::check(azu_vol_write_voxels(h_, index(x, y, z), 1, &tsdf, &weight), "TSDFVolume::writeVoxel");
        if (mirror_valid_) {
            mirror_[index(x, y, z)].tsdf = tsdf;
            mirror_[index(x, y, z)].weight = weight;
        }
    }

    // syncToGPU / syncFromGPU (TSDFVolume.h:161-162): upload the mirror now / download it now.
    void syncToGPU() {
        if (!mirror_valid_) return;
        upload(mirror_);
        host_dirty_ = false;
    }
    void syncFromGPU() {
        flush();
        mirror_valid_ = false;
        refresh();
    }

    // usageFraction (TSDFVolume.cpp:563-569): fraction of voxels with weight > 0, counted
    // on the GPU. observedFraction (TSDFVolume.h:121-128): voxels first observed by
    // integrate since the last reset, over the voxel count.
    float usageFraction() const {
        const_cast<TSDFVolume*>(this)->flush();
        unsigned long long n = 0;
        detail::check(azu_vol_usage(h_, &n), "TSDFVolume::usageFraction");
        return static_cast<float>(n) / static_cast<float>(voxelCount());
    }
    float observedFraction() const {
        unsigned long long upd = 0, newly = 0;
        detail::check(azu_vol_counters(h_, &upd, &newly, 0), "TSDFVolume::observedFraction");
        return static_cast<float>(newly) / static_cast<float>(voxelCount());
    }

    // worldToVoxel / voxelToWorld (TSDFVolume.cpp:537-561): floor per axis; a non-finite
    // or out-of-int-range axis gives the out-of-bounds sentinel INT32_MIN.
    Vec3i worldToVoxel(const Vec3f& world) const {
        const float q[3] = {(world(0) - params_.origin(0)) / params_.voxel_size,
                            (world(1) - params_.origin(1)) / params_.voxel_size,
                            (world(2) - params_.origin(2)) / params_.voxel_size};
        int o[3];
        for (int i = 0; i < 3; ++i) {
            const float f = std::floor(q[i]);
            const bool ok = std::isfinite(f) && f >= -2147483648.0f && f < 2147483648.0f;
            o[i] = ok ? static_cast<int>(f) : INT32_MIN;
        }
        return Vec3i(o[0], o[1], o[2]);
    }
    Vec3f voxelToWorld(int x, int y, int z) const {
        return Vec3f(params_.origin(0) + static_cast<float>(x) * params_.voxel_size,
                     params_.origin(1) + static_cast<float>(y) * params_.voxel_size,
                     params_.origin(2) + static_cast<float>(z) * params_.voxel_size);
    }
    Vec3f voxelToWorld(const Vec3i& v) const { return voxelToWorld(v(0), v(1), v(2)); }
    bool inBounds(int x, int y, int z) const {
        const int r = params_.resolution;
        return x >= 0 && y >= 0 && z >= 0 && x < r && y < r && z < r;
    }

    // TSDFVolume::extractGlobalPointCloud (TSDFVolume.cpp:571-632): the voxels with
    // weight > 1 and |tsdf| < 0.2 at origin + (x, y, z) * voxel_size, in linear voxel
    // order, computed on the GPU. Vec3 is any type built from three floats
    // (Eigen::Vector3f, std::array<float, 3>); colours are zeros (colour is not stored).
    template <class Vec3>
    void extractGlobalPointCloud(std::vector<Vec3>& points_out, std::vector<uint8_t>& colors_out) const {
        const_cast<TSDFVolume*>(this)->flush();
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
    template <class Vec3>
    void extractGlobalPointCloudGPU(std::vector<Vec3>& points_out, std::vector<uint8_t>& colors_out) const {
        extractGlobalPointCloud(points_out, colors_out);
    }

    int integratedFrames() const {
        int f = 0;
        detail::check(azu_vol_integrated_frames(h_, &f), "TSDFVolume::integratedFrames");
        return f;
    }

    // The device handle, with any pending host writes of the mirror uploaded first.
    azu_volume* handle() const {
        const_cast<TSDFVolume*>(this)->flush();
        return h_;
    }
    void* getGPUVoxels() const {
        void* p = nullptr;
        size_t b = 0;
        detail::check(azu_vol_storage(handle(), &p, &b), "TSDFVolume::getGPUVoxels");
        const_cast<TSDFVolume*>(this)->dropMirror();  // the caller may write through it
        return p;
    }
    void setStream(azu_stream s) { stream_ = s; }
    azu_stream stream() const { return stream_; }
    // Empty-space skipping in the raycast (on by default; the same results either way).
    void setEmptySpaceSkipping(bool on) {
        detail::check(azu_vol_set_skip(h_, on ? 1 : 0), "TSDFVolume::setEmptySpaceSkipping");
    }
    void initGPU() {}
    void freeGPU() {}
    void setGPUEnabled(bool) {}  // always on the GPU
    bool isGPUEnabled() const { return true; }

private:
    static azu_tsdf_params toC(const TSDFParams& p) {
        azu_tsdf_params c;
        c.resolution = p.resolution;
        c.voxel_size = p.voxel_size;
        c.truncation = p.truncation;
        c.max_weight = p.max_weight;
        for (int i = 0; i < 3; ++i) c.origin[i] = p.origin(i);
        c.min_depth = p.min_depth;
        c.max_depth = p.max_depth;
        return c;
    }
    static bool sameParams(const azu_tsdf_params& a, const azu_tsdf_params& b) {
        return a.resolution == b.resolution && a.voxel_size == b.voxel_size && a.truncation == b.truncation &&
               a.max_weight == b.max_weight && a.origin[0] == b.origin[0] && a.origin[1] == b.origin[1] &&
               a.origin[2] == b.origin[2] && a.min_depth == b.min_depth && a.max_depth == b.max_depth;
    }
    size_t voxelCount() const {
        const size_t r = static_cast<size_t>(params_.resolution);
        return r * r * r;
    }
    size_t index(int x, int y, int z) const {
        const size_t r = static_cast<size_t>(params_.resolution);
        return (static_cast<size_t>(z) * r + static_cast<size_t>(y)) * r + static_cast<size_t>(x);
    }
    // Download the mirror if the GPU has written since.
    void refresh() {
        if (mirror_valid_) return;
        const size_t n = voxelCount();
        std::vector<float> t(n), w(n);
        detail::check(azu_vol_download(h_, t.data(), w.data()), "TSDFVolume::voxelData");
        mirror_.assign(n, Voxel{});
        for (size_t i = 0; i < n; ++i) {
            mirror_[i].tsdf = t[i];
            mirror_[i].weight = w[i];
        }
        mirror_valid_ = true;
        host_dirty_ = false;
    }
    // Upload host writes of the mirror before the GPU uses the volume.
    void flush() {
        if (host_dirty_ && mirror_valid_) upload(mirror_);
        host_dirty_ = false;
    }
    void upload(const std::vector<Voxel>& v) {
        const size_t n = voxelCount();
        std::vector<float> t(n), w(n);
        for (size_t i = 0; i < n; ++i) {
            t[i] = v[i].tsdf;
            w[i] = v[i].weight;
        }
        detail::check(azu_vol_upload(h_, t.data(), w.data()), "TSDFVolume::syncToGPU");
    }
    // The GPU has written (or may write): the mirror is stale.
    void dropMirror() {
        mirror_valid_ = false;
        host_dirty_ = false;
    }

    TSDFParams params_;
    Storage storage_;
    azu_volume* h_ = nullptr;
    azu_stream stream_ = nullptr;
    detail::DeviceBuffer depth_;
    mutable detail::DeviceBuffer model_;
    mutable std::vector<unsigned char> zeros_;
    std::vector<Voxel> mirror_;
    bool mirror_valid_ = false;
    bool host_dirty_ = false;
};

}  // namespace tsdf
}  // namespace kfusion
