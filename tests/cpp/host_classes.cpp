// Runs the C++ host classes (include/tsdf, include/tracking, include/meshing) on the GPU,
// the way the reference pipeline drives TSDFVolume / ICPTracker / MarchingCubes. Driven
// by tests/test_gpu_host_classes.py, which compares what this program writes with the
// C ABI path through ctypes.
//
//   host_classes loop IN OUT   frame loop over the frames in IN; writes poses, the volume
//                              (through the host mirror, voxelData) and mesh sizes to OUT
//   host_classes reset         the reference's tsdf_reset_contract restated (8^3 volume)
//   host_classes track         the reference's weighting fixture through
//                              ICPTracker::track(FramePyramid, ModelFrame, est, ref)
//   host_classes voxels        voxelAt writes reach the raycast; readVoxel / writeVoxel
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "meshing/MarchingCubes.h"
#include "tracking/ICPTracker.h"
#include "tsdf/TSDFVolume.h"

using namespace kfusion;

static int g_fail = 0;
#define EXPECT(c, what)                                                  \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL: %s [%s:%d]\n", what, __FILE__, __LINE__); \
            ++g_fail;                                                    \
        }                                                                \
    } while (0)

struct LoopHeader {
    int W, H, frames, res, storage;
    float vs, trunc, origin[3], K[4], pose0[16];
};

static int runLoop(const char* in_path, const char* out_path) {
    std::ifstream in(in_path, std::ios::binary);
    LoopHeader h;
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    const size_t px = static_cast<size_t>(h.W) * h.H;
    std::vector<uint16_t> mm(px * h.frames);
    in.read(reinterpret_cast<char*>(mm.data()), mm.size() * 2);
    if (!in) return 3;

    tsdf::TSDFParams p;
    p.resolution = h.res;
    p.voxel_size = h.vs;
    p.truncation = h.trunc;
    p.origin = Vec3f(h.origin[0], h.origin[1], h.origin[2]);
    tsdf::TSDFVolume vol(p, h.storage ? tsdf::Storage::F16_U16 : tsdf::Storage::F32);
    tracking::ICPTracker icp;
    tracking::CameraIntrinsics k{h.K[0], h.K[1], h.K[2], h.K[3]};
    icp.setIntrinsics(k);
    tracking::ModelFrame model(h.W, h.H);
    Mat4f pose;
    std::memcpy(pose.data(), h.pose0, sizeof(h.pose0));
    detail::DeviceBuffer d_mm(px * 2);
    std::vector<float> poses;
    for (int f = 0; f < h.frames; ++f) {
        detail::check(azu_memcpy_h2d(d_mm.p, &mm[f * px], px * 2, nullptr), "frame upload");
        icp.prepareLiveGPU(static_cast<const uint16_t*>(d_mm.p), h.W, h.H);
        if (f > 0) {
            // the reference pipeline's order: ICP against the model raycast at the previous
            // pose, with that pose as estimate and reference (ModelFrame::pose)
            const tracking::ICPResult r = icp.trackPreparedGPU(h.W, h.H, model, pose);
            pose = r.pose;
        }
        const float* d_depth = nullptr;
        detail::check(azu_tracker_level(icp.handle(), 0, &d_depth, nullptr, nullptr, nullptr, nullptr), "level 0");
        vol.integrateGPU(d_depth, nullptr, pose, k.fx, k.fy, k.cx, k.cy, h.W, h.H, p.min_depth, p.max_depth);
        model.render(vol, pose, k);
        poses.insert(poses.end(), pose.data(), pose.data() + 16);
    }
    const std::vector<tsdf::Voxel>& vox = vol.voxelData();
    meshing::MarchingCubes mc;
    float last_progress = -1.0f;
    auto mesh = mc.extract(vol, [&](float x) { last_progress = x; });
    std::string why;
    const bool valid = mesh->validate(&why);
    std::ofstream out(out_path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(poses.data()), poses.size() * 4);
    std::vector<float> t(vox.size()), w(vox.size());
    for (size_t i = 0; i < vox.size(); ++i) {
        t[i] = vox[i].tsdf;
        w[i] = vox[i].weight;
    }
    out.write(reinterpret_cast<const char*>(t.data()), t.size() * 4);
    out.write(reinterpret_cast<const char*>(w.data()), w.size() * 4);
    const long long counts[4] = {static_cast<long long>(mesh->positions.size()),
                                 static_cast<long long>(mesh->triangleCount()), valid ? 1 : 0,
                                 last_progress == 1.0f ? 1 : 0};
    out.write(reinterpret_cast<const char*>(counts), sizeof(counts));
    out.write(reinterpret_cast<const char*>(mesh->positions.data()), mesh->positions.size() * 12);
    out.write(reinterpret_cast<const char*>(mesh->indices.data()), mesh->indices.size() * 4);
    return out ? 0 : 4;
}

// ---- tests/tsdf_reset_contract.cpp (reference), restated ---------------------------
static tsdf::TSDFParams baseParams() {
    tsdf::TSDFParams p;
    p.resolution = 8;
    p.voxel_size = 0.05f;
    p.truncation = 0.05f;
    p.max_weight = 128.0f;
    p.origin = Vec3f(-0.2f, -0.2f, 0.3f);
    return p;
}
static const int kDirty[4][3] = {{0, 0, 0}, {1, 2, 3}, {3, 4, 5}, {7, 7, 7}};
static bool clean(const tsdf::Voxel& v) { return v.tsdf == tsdf::EMPTY_TSDF && v.weight == tsdf::EMPTY_WEIGHT; }
static size_t countDirty(const tsdf::TSDFVolume& vol) {
    size_t n = 0;
    for (const tsdf::Voxel& v : vol.voxelData()) n += clean(v) ? 0 : 1;
    return n;
}
static void dirty(tsdf::TSDFVolume& vol) {
    for (int i = 0; i < 4; ++i) {
        tsdf::Voxel& v = vol.voxelAt(kDirty[i][0], kDirty[i][1], kDirty[i][2]);
        v.tsdf = 0.25f + 0.1f * i;
        v.weight = 7.0f + i;
    }
}
static bool dirtyStuck(tsdf::TSDFVolume& vol) {
    bool ok = countDirty(vol) == 4 && vol.usageFraction() > 0.0f;
    for (int i = 0; i < 4; ++i) {
        const tsdf::Voxel& v = static_cast<const tsdf::TSDFVolume&>(vol).voxelAt(kDirty[i][0], kDirty[i][1], kDirty[i][2]);
        ok = ok && v.tsdf == 0.25f + 0.1f * i && v.weight == 7.0f + i;
        // the device holds them too (a single-voxel read of the device storage)
        const tsdf::Voxel d = vol.readVoxel(kDirty[i][0], kDirty[i][1], kDirty[i][2]);
        ok = ok && d.tsdf == 0.25f + 0.1f * i && d.weight == 7.0f + i;
    }
    return ok;
}
static bool allEmpty(const tsdf::TSDFVolume& vol) {
    const size_t r = vol.params().resolution;
    return vol.voxelData().size() == r * r * r && countDirty(vol) == 0 && vol.usageFraction() == 0.0f &&
           vol.integratedFrames() == 0;
}
static void integrateOne(tsdf::TSDFVolume& vol) {
    std::vector<float> depth(64, 0.5f);
    std::vector<uint8_t> rgb(64 * 3, 200);
    vol.integrate(depth.data(), rgb.data(), Mat4f::Identity(), 500.0f, 500.0f, 3.5f, 3.5f, 8, 8);
}

static int runReset() {
    {
        const tsdf::TSDFVolume vol(baseParams());
        EXPECT(allEmpty(vol), "fresh volume is empty");
        const tsdf::Voxel d;
        EXPECT(clean(d) && d.r == tsdf::EMPTY_COLOR, "default Voxel is the empty state");
    }
    {
        tsdf::TSDFVolume vol(baseParams());
        dirty(vol);
        EXPECT(dirtyStuck(vol), "dirty voxels are visible");
        vol.reset();
        EXPECT(allEmpty(vol), "reset clears every voxel");
    }
    {
        tsdf::TSDFVolume vol(baseParams());
        integrateOne(vol);
        EXPECT(vol.integratedFrames() == 1 && vol.usageFraction() > 0.0f, "one frame observed voxels");
        vol.reset();
        EXPECT(allEmpty(vol), "reset after integrate clears");
    }
    {
        tsdf::TSDFVolume vol(baseParams());
        dirty(vol);
        vol.setParams(baseParams());
        EXPECT(dirtyStuck(vol), "equal params keep the voxels");
    }
    const std::function<void(tsdf::TSDFParams&)> mutate[] = {
        [](tsdf::TSDFParams& p) { p.voxel_size = 0.025f; },
        [](tsdf::TSDFParams& p) { p.origin = Vec3f(0.1f, -0.3f, 0.4f); },
        [](tsdf::TSDFParams& p) { p.truncation = 0.02f; },
        [](tsdf::TSDFParams& p) { p.max_weight = 32.0f; },
        [](tsdf::TSDFParams& p) { p.resolution = 5; },
        [](tsdf::TSDFParams& p) { p.min_depth = 0.4f; },
        [](tsdf::TSDFParams& p) { p.max_depth = 4.0f; },
    };
    for (const auto& m : mutate) {
        tsdf::TSDFVolume vol(baseParams());
        dirty(vol);
        EXPECT(dirtyStuck(vol), "dirty before the change");
        tsdf::TSDFParams next = baseParams();
        m(next);
        vol.setParams(next);
        const tsdf::TSDFParams& got = vol.params();
        EXPECT(got.resolution == next.resolution && got.voxel_size == next.voxel_size &&
                   got.truncation == next.truncation && got.max_weight == next.max_weight &&
                   got.origin(0) == next.origin(0) && got.origin(1) == next.origin(1) &&
                   got.origin(2) == next.origin(2),
               "params() round-trips the new values");
        EXPECT(allEmpty(vol), "a changed field clears the volume");
        const Vec3f w = vol.voxelToWorld(1, 0, 0);
        EXPECT(std::fabs(w(0) - (next.origin(0) + next.voxel_size)) < 1e-6f, "voxelToWorld uses the new params");
        const Vec3i vi = vol.worldToVoxel(Vec3f(next.origin(0) + next.voxel_size * 4.5f,
                                                next.origin(1) + next.voxel_size * 0.5f,
                                                next.origin(2) + next.voxel_size * 2.5f));
        EXPECT(vi(0) == 4 && vi(1) == 0 && vi(2) == 2, "worldToVoxel uses the new params");
    }
    std::printf("reset: %s\n", g_fail ? "FAIL" : "PASS");
    return g_fail ? 1 : 0;
}

// ---- tests/icp_weighting_contract.cpp:62-216 (reference) through the host entry -------
static int runTrack() {
    sensor::FramePyramid live;
    tracking::ModelFrame model;
    auto put = [&](int px, int py, float err) {
        const float z = 1.0f;
        const Vec3f v((static_cast<float>(px) - 319.5f) / 525.0f * z, (static_cast<float>(py) - 239.5f) / 525.0f * z,
                      z);
        const size_t i = static_cast<size_t>(py) * 640 + px;
        live.levels[0].vertices[i] = v;
        live.levels[0].normals[i] = Vec3f(0, 0, 1);
        live.levels[0].depth_meters[i] = z;
        model.vertices[i] = Vec3f(v(0), v(1), z - err);
        model.normals[i] = Vec3f(0, 0, 1);
    };
    for (int k = 0; k < 5; ++k) {
        put(319 - k, 239 - k, 0.01f);
        put(320 + k, 240 + k, 0.01f);
    }
    put(320, 239, 0.08f);
    tracking::ICPParams prm;
    prm.max_iterations[0] = 1;
    prm.max_iterations[1] = prm.max_iterations[2] = 0;
    tracking::ICPTracker icp(prm);
    const tracking::ICPResult r = icp.track(live, model, Mat4f::Identity(), Mat4f::Identity());
    const tracking::ICPResult r2 = icp.track(live, model, Mat4f::Identity(), Mat4f::Identity());
    EXPECT(r.inliers == 11 && r.valid_live_points == 11 && r.valid_model_points == 11, "11 correspondences");
    EXPECT(std::fabs(r.pose(2, 3) - (-0.12 / 11.25)) <= 5e-6, "t_z = -0.12 / 11.25");
    EXPECT(std::fabs(r.error - 0.0038 / 11.0) <= 5e-6, "error = 0.0038 / 11");
    EXPECT(std::memcmp(r.pose.data(), r2.pose.data(), 64) == 0, "repeat is bit-identical");
    const tracking::ICPResult fresh{};
    EXPECT(fresh.pose == Mat4f::Identity() && std::isinf(fresh.final_step), "ICPResult{} defaults");
    tracking::ICPParams wide;
    wide.angle_threshold = 120.0f;
    EXPECT(tracking::ICPTracker(wide).params().angle_threshold == 85.0f, "angle threshold clamps to 85");
    std::printf("track: %s t_z=%.9f\n", g_fail ? "FAIL" : "PASS", r.pose(2, 3));
    return g_fail ? 1 : 0;
}

// ---- voxelAt writes through the mirror reach the GPU raycast; single-voxel access -------
static int runVoxels() {
    tsdf::TSDFParams p;
    p.resolution = 32;
    p.voxel_size = 0.04f;
    p.truncation = 0.12f;
    p.origin = Vec3f(-0.64f, -0.64f, 0.0f);
    tsdf::TSDFVolume vol(p);
    // a wall at z = 0.38 m: voxel planes k = 9 (+) / 10 (-), all observed
    for (int z = 0; z < 32; ++z)
        for (int y = 0; y < 32; ++y)
            for (int x = 0; x < 32; ++x) {
                tsdf::Voxel& v = vol.voxelAt(x, y, z);
                v.tsdf = std::fmax(-1.0f, std::fmin(1.0f, (0.38f - z * 0.04f) / 0.12f));
                v.weight = 1.0f;
            }
    std::vector<Vec3f> verts(64 * 48), norms(64 * 48);
    vol.raycast(Mat4f::Identity(), 50.0f, 50.0f, 31.5f, 23.5f, 64, 48, verts.data(), norms.data());
    const Vec3f c = verts[24 * 64 + 32];
    EXPECT(std::fabs(c(2) - 0.38f) < 1e-4f, "the mirror's wall is what the GPU raycast hits");
    EXPECT(std::fabs(norms[24 * 64 + 32](2) + 1.0f) < 1e-4f, "normal (0, 0, -1)");
    const tsdf::Voxel a = vol.readVoxel(3, 4, 5);
    const tsdf::TSDFVolume& cv = vol;
    EXPECT(a.tsdf == cv.voxelAt(3, 4, 5).tsdf && a.weight == 1.0f, "readVoxel sees the uploaded mirror");
    vol.writeVoxel(3, 4, 5, -0.5f, 9.0f);
    EXPECT(cv.voxelAt(3, 4, 5).tsdf == -0.5f && cv.voxelAt(3, 4, 5).weight == 9.0f, "writeVoxel updates the mirror");
    vol.syncFromGPU();
    EXPECT(cv.voxelAt(3, 4, 5).tsdf == -0.5f && cv.voxelAt(3, 4, 5).weight == 9.0f, "and the device");
    std::printf("voxels: %s z=%.6f\n", g_fail ? "FAIL" : "PASS", c(2));
    return g_fail ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "loop" && argc == 4) return runLoop(argv[2], argv[3]);
        if (mode == "reset") return runReset();
        if (mode == "track") return runTrack();
        if (mode == "voxels") return runVoxels();
    } catch (const std::exception& e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 5;
    }
    return 2;
}
