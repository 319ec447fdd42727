// TEST INFRASTRUCTURE ONLY: C entry points over the CPU oracle (oracle.hpp) for the
// Python tests (ctypes) and the CPU-baseline timing in bench.py. Functions come in
// _f32 (FP32 parity oracle) and _f64 (FP64 oracle) pairs. Vectors are packed xyz
// triples, poses 4x4 column-major world-from-camera, intrinsics {fx, fy, cx, cy}.
#include <chrono>

#include "oracle.hpp"
#include "oracle_condition.hpp"
#include "oracle_mc.hpp"

using namespace orc;

namespace {
Params mkParams(int res, float vs, float trunc, float max_w, const float* origin, float min_d, float max_d) {
    Params p;
    p.res = res;
    p.vs = vs;
    p.trunc = trunc;
    p.max_w = max_w;
    p.origin[0] = origin[0];
    p.origin[1] = origin[1];
    p.origin[2] = origin[2];
    p.min_depth = min_d;
    p.max_depth = max_d;
    return p;
}
Intr mkK(const float* K) { return Intr{K[0], K[1], K[2], K[3]}; }
IcpParams mkIcp(const int* iters, float dist, float angle) {
    IcpParams p;
    p.iters[0] = iters[0];
    p.iters[1] = iters[1];
    p.iters[2] = iters[2];
    // sanitizeParams (ICPTracker.cpp:24-34)
    if (!std::isfinite(angle)) angle = 30.0f;
    else if (angle < 0.0f) angle = 0.0f;
    else if (angle > 85.0f) angle = 85.0f;
    p.dist_threshold = dist;
    p.angle_threshold = angle;
    return p;
}

// Full live pyramid from a metre depth frame (FrameData.cpp:137-345).
template <class T> struct Pyr {
    std::vector<T> d[3];
    std::vector<V3<T>> v[3], n[3];
    int W[3], H[3];
    void build(const T* depth, int W0, int H0, Intr K) {
        W[0] = W0; H[0] = H0;
        for (int l = 1; l < 3; ++l) { W[l] = W[l - 1] / 2; H[l] = H[l - 1] / 2; }
        for (int l = 0; l < 3; ++l) {
            d[l].assign(static_cast<size_t>(W[l]) * H[l], T(0));
            v[l].assign(d[l].size(), V3<T>{0, 0, 0});
            n[l].assign(d[l].size(), V3<T>{0, 0, 0});
        }
        std::copy(depth, depth + static_cast<size_t>(W0) * H0, d[0].begin());
        vertices(d[0].data(), W0, H0, K, v[0].data());
        normals(d[0].data(), v[0].data(), W0, H0, n[0].data());
        for (int l = 1; l < 3; ++l)
            downsample(d[l - 1].data(), v[l - 1].data(), n[l - 1].data(), W[l - 1], H[l - 1], d[l].data(),
                       v[l].data(), n[l].data());
    }
    void levels(Level<T> out[3]) const {
        for (int l = 0; l < 3; ++l) out[l] = Level<T>{v[l].data(), n[l].data(), W[l], H[l]};
    }
};

// The benchmark frame loop (the pipeline order): frame 0 integrates at pose0 and
// raycasts; frame k >= 1 builds the live pyramid, tracks against the model raycast at the
// previous tracked pose (ref = that pose), integrates and raycasts at the tracked pose
// (velocity != 0: the estimate is the constant-velocity prediction, else the previous
// pose). Frames are uint16 mm (frames_mm, band-gated here) or float metres (frames_m, as
// given). poses_out: 16 floats per frame; ms_out: 5 doubles per frame {pyramid, icp,
// integrate, raycast, total}; with poses_in, frames k >= 1 use poses_in[k] instead of ICP.
template <class T>
int runSequence(Volume<T>& V, const uint16_t* frames_mm, const float* frames_m, int nframes, int W, int H,
                const float* K, const float* pose0, const int* iters, float dist, float angle, const float* poses_in,
                float* poses_out, double* ms_out, long long* n_upd_out, int velocity, double* icp_out) {
    const IcpParams prm = mkIcp(iters, dist, angle);
    const Intr KK = mkK(K);
    const size_t npx = static_cast<size_t>(W) * H;
    std::vector<T> depth(npx);
    std::vector<V3<T>> mv(npx), mn(npx);
    Pose<T> pose = poseFrom<T>(pose0);
    Pose<T> last = pose;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const T mind = static_cast<T>(V.p.min_depth), maxd = static_cast<T>(V.p.max_depth);
    for (int f = 0; f < nframes; ++f) {
        const auto t0 = clk::now();
        if (frames_mm)
            depthMmToMeters<T>(frames_mm + npx * f, static_cast<int>(npx), mind, maxd, depth.data());
        else
            for (size_t i = 0; i < npx; ++i) depth[i] = static_cast<T>(frames_m[npx * f + i]);
        auto t1 = t0, t2 = t0;
        if (f > 0) {
            if (poses_in) {
                pose = poseFrom<T>(poses_in + 16 * f);
                t1 = t2 = clk::now();
            } else {
                Pyr<T> P;
                P.build(depth.data(), W, H, KK);
                Level<T> L[3];
                P.levels(L);
                t1 = clk::now();
                const Pose<T> est = velocity ? predictPose(pose, last) : pose;
                last = pose;
                const IcpResult<T> r = track(L, mv.data(), mn.data(), W, H, KK, est, pose, prm);
                if (icp_out) {
                    icp_out[4 * f + 0] = r.converged;
                    icp_out[4 * f + 1] = static_cast<double>(r.final_step);
                    icp_out[4 * f + 2] = r.inliers;
                    icp_out[4 * f + 3] = r.steps;
                }
                pose = poseFrom<T>(r.pose);
                t2 = clk::now();
            }
        } else {
            t1 = t2 = clk::now();
        }
        const long long nu = integrate(V, depth.data(), W, H, KK, pose, mind, maxd);
        const auto t3 = clk::now();
        raycast(V, pose, KK, W, H, mv.data(), mn.data());
        const auto t4 = clk::now();
        if (poses_out) poseTo(pose, poses_out + 16 * f);
        if (n_upd_out) n_upd_out[f] = nu;
        if (ms_out) {
            ms_out[5 * f + 0] = ms(t0, t1);
            ms_out[5 * f + 1] = ms(t1, t2);
            ms_out[5 * f + 2] = ms(t2, t3);
            ms_out[5 * f + 3] = ms(t3, t4);
            ms_out[5 * f + 4] = ms(t0, t4);
        }
    }
    return nframes;
}
}  // namespace

#define ORC_API extern "C" __attribute__((visibility("default")))

// The same set of functions for each scalar type.
#define ORC_DEFINE(SFX, T)                                                                                   \
    ORC_API void* orc_vol_create_##SFX(int res, float vs, float trunc, float max_w, const float* origin,       \
                                       float min_d, float max_d) {                                            \
        return new Volume<T>(mkParams(res, vs, trunc, max_w, origin, min_d, max_d));                          \
    }                                                                                                        \
    ORC_API void orc_vol_destroy_##SFX(void* h) { delete static_cast<Volume<T>*>(h); }                       \
    ORC_API void orc_vol_reset_##SFX(void* h) { static_cast<Volume<T>*>(h)->reset(); }                       \
    ORC_API void orc_vol_get_##SFX(void* h, T* tsdf, T* weight) {                                            \
        const auto& V = *static_cast<Volume<T>*>(h);                                                         \
        const long long n = static_cast<long long>(V.vox.size());                                            \
        _Pragma("omp parallel for") for (long long i = 0; i < n; ++i) {                                      \
            tsdf[i] = V.vox[i].tsdf;                                                                         \
            weight[i] = V.vox[i].weight;                                                                     \
        }                                                                                                    \
    }                                                                                                        \
    ORC_API void orc_vol_set_##SFX(void* h, const T* tsdf, const T* weight) {                                \
        auto& V = *static_cast<Volume<T>*>(h);                                                               \
        const long long n = static_cast<long long>(V.vox.size());                                            \
        _Pragma("omp parallel for") for (long long i = 0; i < n; ++i) {                                      \
            V.vox[i].tsdf = tsdf[i];                                                                         \
            V.vox[i].weight = weight[i];                                                                     \
        }                                                                                                    \
    }                                                                                                        \
    ORC_API void orc_depth_mm_##SFX(const uint16_t* mm, int n, T min_d, T max_d, T* out) {                   \
        depthMmToMeters<T>(mm, n, min_d, max_d, out);                                                        \
    }                                                                                                        \
    ORC_API long long orc_integrate_##SFX(void* h, const T* depth, int W, int H, const float* K,             \
                                          const float* pose, T min_d, T max_d, long long* newly) {           \
        return integrate(*static_cast<Volume<T>*>(h), depth, W, H, mkK(K), poseFrom<T>(pose), min_d, max_d,  \
                         newly);                                                                             \
    }                                                                                                        \
    ORC_API void orc_raycast_##SFX(void* h, const float* pose, const float* K, int W, int H, T* verts,       \
                                   T* norms, long long* nsamples) {                                          \
        raycast(*static_cast<Volume<T>*>(h), poseFrom<T>(pose), mkK(K), W, H,                                \
                reinterpret_cast<V3<T>*>(verts), reinterpret_cast<V3<T>*>(norms), nullptr, nsamples);        \
    }                                                                                                        \
    /* Distinct voxels read by all marches and normal stencils (serial pass). */                             \
    ORC_API long long orc_raycast_touched_##SFX(void* h, const float* pose, const float* K, int W, int H) {   \
        const auto& V = *static_cast<Volume<T>*>(h);                                                         \
        std::vector<uint8_t> touched(V.vox.size(), 0);                                                       \
        std::vector<V3<T>> v(static_cast<size_t>(W) * H), n(v.size());                                      \
        raycast(V, poseFrom<T>(pose), mkK(K), W, H, v.data(), n.data(), &touched, nullptr);                  \
        long long c = 0;                                                                                     \
        for (uint8_t t : touched) c += t;                                                                    \
        return c;                                                                                            \
    }                                                                                                        \
    ORC_API void orc_pyramid_##SFX(const T* depth, int W, int H, const float* K, T* d1, T* d2, T* v0, T* n0,  \
                                   T* v1, T* n1, T* v2, T* n2) {                                             \
        Pyr<T> P;                                                                                            \
        P.build(depth, W, H, mkK(K));                                                                        \
        T* dd[3] = {nullptr, d1, d2};                                                                        \
        T* vv[3] = {v0, v1, v2};                                                                             \
        T* nn[3] = {n0, n1, n2};                                                                             \
        for (int l = 0; l < 3; ++l) {                                                                        \
            if (dd[l]) std::copy(P.d[l].begin(), P.d[l].end(), dd[l]);                                       \
            std::memcpy(vv[l], P.v[l].data(), P.v[l].size() * sizeof(V3<T>));                                \
            std::memcpy(nn[l], P.n[l].data(), P.n[l].size() * sizeof(V3<T>));                                \
        }                                                                                                    \
    }                                                                                                        \
    /* buildLinearSystem on one live level. out_cnt = {valid_live, valid_model, projected, dist_f,           \
       angle_f, inliers}. Returns 1 when inliers > 0. */                                                     \
    ORC_API int orc_icp_system_##SFX(const T* lv, const T* ln, int W, int H, const T* mv, const T* mn,       \
                                     int MW, int MH, const float* K, const float* pose, const float* ref,    \
                                     float dist, float angle, T* A, T* b, T* residual, int* out_cnt) {       \
        const int it[3] = {1, 1, 1};                                                                         \
        const IcpParams prm = mkIcp(it, dist, angle);                                                        \
        Level<T> L{reinterpret_cast<const V3<T>*>(lv), reinterpret_cast<const V3<T>*>(ln), W, H};            \
        IcpSystemCounters c;                                                                                 \
        T res = 0;                                                                                           \
        const bool ok = buildLinearSystem(L, reinterpret_cast<const V3<T>*>(mv),                             \
                                          reinterpret_cast<const V3<T>*>(mn), MW, MH, mkK(K),                \
                                          poseFrom<T>(pose), poseFrom<T>(ref), prm, A, b, res, c);           \
        *residual = res;                                                                                     \
        out_cnt[0] = c.valid_live; out_cnt[1] = c.valid_model; out_cnt[2] = c.projected;                     \
        out_cnt[3] = c.dist_filtered; out_cnt[4] = c.angle_filtered; out_cnt[5] = c.inliers;                 \
        return ok ? 1 : 0;                                                                                   \
    }                                                                                                        \
    /* ICPTracker::track. live v/n per level are given explicitly (levels 0..2). res_pose[16],             \
       res_scalars = {error, final_step}, res_ints = {inliers, converged, tracking_ok, valid_live,           \
       valid_model, projected, dist_f, angle_f}, res_info[36]. */                                            \
    ORC_API void orc_icp_track_##SFX(const T* v0, const T* n0, const T* v1, const T* n1, const T* v2,        \
                                     const T* n2, int W, int H, const T* mv, const T* mn, int MW, int MH,    \
                                     const float* K, const float* est, const float* ref, const int* iters,   \
                                     float dist, float angle, T* res_pose, T* res_scalars, int* res_ints,    \
                                     T* res_info) {                                                          \
        const IcpParams prm = mkIcp(iters, dist, angle);                                                     \
        Level<T> L[3] = {{reinterpret_cast<const V3<T>*>(v0), reinterpret_cast<const V3<T>*>(n0), W, H},     \
                         {reinterpret_cast<const V3<T>*>(v1), reinterpret_cast<const V3<T>*>(n1), W / 2,     \
                          H / 2},                                                                            \
                         {reinterpret_cast<const V3<T>*>(v2), reinterpret_cast<const V3<T>*>(n2), W / 4,     \
                          H / 4}};                                                                           \
        const IcpResult<T> r = track(L, reinterpret_cast<const V3<T>*>(mv), reinterpret_cast<const V3<T>*>(mn), \
                                     MW, MH, mkK(K), poseFrom<T>(est), poseFrom<T>(ref), prm);               \
        for (int i = 0; i < 16; ++i) res_pose[i] = r.pose[i];                                                \
        res_scalars[0] = r.error;                                                                            \
        res_scalars[1] = r.final_step;                                                                       \
        res_ints[0] = r.inliers; res_ints[1] = r.converged; res_ints[2] = r.tracking_ok;                     \
        res_ints[3] = r.valid_live; res_ints[4] = r.valid_model; res_ints[5] = r.projected;                  \
        res_ints[6] = r.dist_filtered; res_ints[7] = r.angle_filtered; res_ints[8] = r.steps;                \
        for (int i = 0; i < 36; ++i) res_info[i] = r.information[i];                                         \
    }                                                                                                        \
    /* runSequence over uint16 mm frames, or over float metre frames (_m). */                            \
    ORC_API int orc_run_sequence_##SFX(void* h, const uint16_t* frames_mm, int nframes, int W, int H,        \
                                       const float* K, const float* pose0, const int* iters, float dist,     \
                                       float angle, const float* poses_in, float* poses_out, double* ms_out,  \
                                       long long* n_upd_out, int velocity, double* icp_out) {                \
        return runSequence(*static_cast<Volume<T>*>(h), frames_mm, nullptr, nframes, W, H, K, pose0, iters,   \
                           dist, angle, poses_in, poses_out, ms_out, n_upd_out, velocity, icp_out);          \
    }                                                                                                        \
    ORC_API int orc_run_sequence_m_##SFX(void* h, const float* frames_m, int nframes, int W, int H,          \
                                         const float* K, const float* pose0, const int* iters, float dist,   \
                                         float angle, float* poses_out, long long* n_upd_out, int velocity,  \
                                         double* icp_out) {                                                  \
        return runSequence(*static_cast<Volume<T>*>(h), nullptr, frames_m, nframes, W, H, K, pose0, iters,    \
                           dist, angle, nullptr, poses_out, nullptr, n_upd_out, velocity, icp_out);          \
    }                                                                                                        \
    ORC_API void orc_model_get_##SFX(void* h, const float* pose, const float* K, int W, int H, T* verts,     \
                                     T* norms) {                                                             \
        raycast(*static_cast<Volume<T>*>(h), poseFrom<T>(pose), mkK(K), W, H,                                \
                reinterpret_cast<V3<T>*>(verts), reinterpret_cast<V3<T>*>(norms));                           \
    }

ORC_DEFINE(f32, float)
ORC_DEFINE(f64, double)

ORC_API void orc_jacobi_eigenvalues(const double* A, int n, double* eig) { jacobiEigenvalues(A, n, eig); }
ORC_API int orc_ldlt_solve_f32(const float* A, const float* b, float* x) { return ldltSolve(A, b, x) ? 1 : 0; }
ORC_API int orc_ldlt_solve_f64(const double* A, const double* b, double* x) { return ldltSolve(A, b, x) ? 1 : 0; }
ORC_API void orc_project_so3(const double* M, double* R) { projectToSO3d(M, R); }
ORC_API float orc_damping_f32(const float* A, int level) { return dampingForHessian(A, level); }
ORC_API int orc_floor_to_int(float v, int* out) { return floorToInt(v, out) ? 1 : 0; }
ORC_API void orc_scale_intrinsics(const float* K, int rw, int rh, int w, int h, float* out) {
    const Intr k = scaleIntrinsics(mkK(K), rw, rh, w, h);
    out[0] = k.fx; out[1] = k.fy; out[2] = k.cx; out[3] = k.cy;
}
ORC_API void orc_set_icp_chunks(int n) { icpChunks() = n; }
// The constant-velocity estimate (PipelineController.cpp:1162-1187) of two
// column-major rigid poses: out = cur * (last^-1 * cur).
ORC_API void orc_predict_pose_f32(const float* cur, const float* last, float* out) {
    poseTo(predictPose(poseFrom<float>(cur), poseFrom<float>(last)), out);
}
ORC_API void orc_predict_pose_f64(const double* cur, const double* last, double* out) {
    poseTo(predictPose(poseFrom<double>(cur), poseFrom<double>(last)), out);
}
// One downsample step (FrameData.cpp:262-334) on explicit level arrays.
ORC_API void orc_downsample_f32(const float* d, const float* v, const float* n, int W, int H, float* d1, float* v1,
                                float* n1) {
    downsample(d, reinterpret_cast<const V3<float>*>(v), reinterpret_cast<const V3<float>*>(n), W, H, d1,
               reinterpret_cast<V3<float>*>(v1), reinterpret_cast<V3<float>*>(n1));
}
ORC_API int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Marching cubes over an f32 volume given as tsdf / weight arrays (res^3, x fastest).
// Two-call protocol: with null outputs it returns the vertex and index counts; then
// call again with buffers of those sizes.
ORC_API long long orc_mc_extract(const float* tsdf, const float* weight, int res, float vs, const float* origin,
                                 long long max_tris, float* pos, float* nrm, uint32_t* idx, long long* nvert,
                                 long long* nidx, int* truncated) {
    static thread_local MeshOut cache;
    static thread_local const float* cache_key = nullptr;
    if (!pos || !cache_key || cache_key != tsdf) {
        cache = mcExtract(tsdf, weight, res, vs, origin, static_cast<size_t>(max_tris));
        cache_key = tsdf;
    }
    *nvert = static_cast<long long>(cache.pos.size() / 3);
    *nidx = static_cast<long long>(cache.idx.size());
    *truncated = cache.truncated ? 1 : 0;
    if (pos) {
        std::memcpy(pos, cache.pos.data(), cache.pos.size() * sizeof(float));
        std::memcpy(nrm, cache.nrm.data(), cache.nrm.size() * sizeof(float));
        std::memcpy(idx, cache.idx.data(), cache.idx.size() * sizeof(uint32_t));
        cache_key = nullptr;
    }
    return *nvert;
}

// Depth conditioning stages over raw codes (W x H, row-major). luma may be null for the
// stages that do not read it; state is the per-pixel EMA history in metres.
ORC_API void orc_cond_median(const uint16_t* in, uint16_t* out, int W, int H, float lo, float hi) {
    orc::cond::median(in, out, W, H, lo, hi);
}
ORC_API void orc_cond_fill_holes(const uint16_t* in, uint16_t* out, int W, int H, float lo, float hi) {
    orc::cond::fillHoles(in, out, W, H, lo, hi);
}
ORC_API void orc_cond_guided(const uint16_t* in, const float* luma, uint16_t* out, int W, int H, float lo, float hi,
                             int exp_mode) {
    orc::cond::guided(in, luma, out, W, H, lo, hi, exp_mode);
}
ORC_API void orc_cond_ema(const uint16_t* in, uint16_t* out, float* state, int W, int H, float lo, float hi) {
    orc::cond::ema(in, out, state, W, H, lo, hi);
}
ORC_API void orc_cond_process(const uint16_t* in, const float* luma, uint16_t* out, float* state, int W, int H,
                              float lo, float hi, int stages, int exp_mode) {
    orc::cond::process(in, luma, out, state, W, H, lo, hi, stages, exp_mode);
}
