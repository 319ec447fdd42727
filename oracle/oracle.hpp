// TEST INFRASTRUCTURE ONLY. This is the CPU oracle of the KinectFusion frame loop: a
// from-scratch restatement of the reference CPU path, templated on the scalar type
// (float = the FP32 parity oracle, double = the FP64 oracle). Only tests/, the smoke
// check and bench.py's cpu_baseline leg may load it, and only as the checker or the
// timed CPU baseline. The product (libazu_hip.so) never links or calls it.
//
// Numeric policy: compiled with -ffp-contract=off and IEEE division / sqrt. Three-term
// sums follow Eigen's fixed-size reduction order a0 + (a1 + a2) (Eigen 3.4
// redux_novec_unroller splits a length-3 sum as 1 + 2), and 4x4 products follow
// Eigen's column-packet order ((p0 + p1) + p2) + p3. Bit-level agreement with an
// Eigen build of the reference is not claimed (the reference Release build contracts
// to FMA); agreement with the reference is pinned by its own test fixtures instead.
//
// Every function cites the reference file:line it restates.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Small vector / matrix helpers with a fixed evaluation order.
// ---------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> mul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> divs(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T sum3(T a0, T a1, T a2) { return a0 + (a1 + a2); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline T sqnorm(V3<T> a) { return dot(a, a); }
template <class T> inline T norm(V3<T> a) { return std::sqrt(sqnorm(a)); }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline bool finite3(V3<T> a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
// std::min / std::max semantics (first argument wins ties and NaN in b position).
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }

template <class T> struct M3 {
    T m[3][3];
};
template <class T> inline V3<T> mulv(const M3<T>& R, V3<T> v) {
    return {sum3(R.m[0][0] * v.x, R.m[0][1] * v.y, R.m[0][2] * v.z),
            sum3(R.m[1][0] * v.x, R.m[1][1] * v.y, R.m[1][2] * v.z),
            sum3(R.m[2][0] * v.x, R.m[2][1] * v.y, R.m[2][2] * v.z)};
}
template <class T> inline V3<T> mulvT(const M3<T>& R, V3<T> v) {  // R^T v
    return {sum3(R.m[0][0] * v.x, R.m[1][0] * v.y, R.m[2][0] * v.z),
            sum3(R.m[0][1] * v.x, R.m[1][1] * v.y, R.m[2][1] * v.z),
            sum3(R.m[0][2] * v.x, R.m[1][2] * v.y, R.m[2][2] * v.z)};
}
template <class T> inline M3<T> mulm(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C.m[i][j] = sum3(A.m[i][0] * B.m[0][j], A.m[i][1] * B.m[1][j], A.m[i][2] * B.m[2][j]);
    return C;
}

// Rigid pose (world-from-camera when read from the API). The API matrix is a 4x4
// column-major float[16] (Eigen's default storage), element (r,c) at p[c*4+r].
template <class T> struct Pose {
    M3<T> R;
    V3<T> t;
};
template <class T, class S> inline Pose<T> poseFrom(const S* p) {
    Pose<T> P;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P.R.m[r][c] = static_cast<T>(p[c * 4 + r]);
    P.t = {static_cast<T>(p[12]), static_cast<T>(p[13]), static_cast<T>(p[14])};
    return P;
}
template <class T, class S> inline void poseTo(const Pose<T>& P, S* p) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) p[c * 4 + r] = static_cast<S>(P.R.m[r][c]);
    p[12] = static_cast<S>(P.t.x);
    p[13] = static_cast<S>(P.t.y);
    p[14] = static_cast<S>(P.t.z);
    p[3] = p[7] = p[11] = 0;
    p[15] = 1;
}
// Rigid inverse: R^T, -(R^T t). The reference calls Eigen's general 4x4 inverse
// (TSDFVolume.cpp:238, ICPTracker.cpp:167); for a rigid pose both are the same
// matrix up to rounding, and the rigid form is the one fixed here.
template <class T> inline Pose<T> inverse(const Pose<T>& P) {
    Pose<T> I;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) I.R.m[r][c] = P.R.m[c][r];
    const V3<T> rt = mulv(I.R, P.t);
    I.t = {-rt.x, -rt.y, -rt.z};
    return I;
}

// A * B for rigid poses, each entry summed left to right ((a0 b0 + a1 b1) + a2 b2) + a3 b3
// (Eigen's 4x4 column-packet order) with the implicit bottom row (0, 0, 0, 1).
template <class T> inline Pose<T> compose(const Pose<T>& A, const Pose<T>& B) {
    Pose<T> C;
    const T at[3] = {A.t.x, A.t.y, A.t.z}, bt[3] = {B.t.x, B.t.y, B.t.z};
    T ct[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            C.R.m[r][c] = ((A.R.m[r][0] * B.R.m[0][c] + A.R.m[r][1] * B.R.m[1][c]) + A.R.m[r][2] * B.R.m[2][c]) +
                          at[r] * T(0);
        ct[r] = ((A.R.m[r][0] * bt[0] + A.R.m[r][1] * bt[1]) + A.R.m[r][2] * bt[2]) + at[r] * T(1);
    }
    C.t = {ct[0], ct[1], ct[2]};
    return C;
}

// Constant-velocity estimate of the reference pipeline (src/app/PipelineController.cpp:
// 1162-1187): cur * (last^-1 * cur).
template <class T> inline Pose<T> predictPose(const Pose<T>& cur, const Pose<T>& last) {
    return compose(cur, compose(inverse(last), cur));
}

struct Intr {
    float fx, fy, cx, cy;
};

// scaleIntrinsics (include/sensor/CameraIntrinsics.h:33-40)
inline Intr scaleIntrinsics(Intr k, int ref_w, int ref_h, int w, int h) {
    if (w == ref_w && h == ref_h) return k;
    const float sx = static_cast<float>(w) / static_cast<float>(ref_w);
    const float sy = static_cast<float>(h) / static_cast<float>(ref_h);
    return Intr{k.fx * sx, k.fy * sy, (k.cx + 0.5f) * sx - 0.5f, (k.cy + 0.5f) * sy - 0.5f};
}

// floorToInt (include/utils/CoordinateMath.h:27-49)
inline bool floorToInt(float value, int* out) {
    if (!std::isfinite(value)) return false;
    const double f = std::floor(static_cast<double>(value));
    if (f < static_cast<double>(std::numeric_limits<int>::min()) ||
        f > static_cast<double>(std::numeric_limits<int>::max()))
        return false;
    *out = static_cast<int>(f);
    return true;
}
inline bool floorToInt(double value, int* out) {
    if (!std::isfinite(value)) return false;
    const double f = std::floor(value);
    if (f < static_cast<double>(std::numeric_limits<int>::min()) ||
        f > static_cast<double>(std::numeric_limits<int>::max()))
        return false;
    *out = static_cast<int>(f);
    return true;
}

// ---------------------------------------------------------------------------
// Depth front end for uint16 millimetre frames. The reference's band rule
// (include/sensor/DepthValidity.h:81-93): accepted meters are finite and inside
// [min_depth, max_depth], anything else is exactly 0. Millimetres convert as mm*0.001.
// ---------------------------------------------------------------------------
template <class T>
void depthMmToMeters(const uint16_t* mm, int n, T min_d, T max_d, T* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const T m = static_cast<T>(mm[i]) * static_cast<T>(0.001);
        out[i] = (mm[i] != 0 && std::isfinite(m) && m >= min_d && m <= max_d) ? m : T(0);
    }
}

// updateVerticesFromDepth (src/sensor/FrameData.cpp:113-133)
template <class T>
void vertices(const T* depth, int W, int H, Intr K, V3<T>* v) {
    const T fx_inv = T(1) / static_cast<T>(K.fx);
    const T fy_inv = T(1) / static_cast<T>(K.fy);
    const T cx = K.cx, cy = K.cy;
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < W * H; ++idx) {
        const T d = depth[idx];
        if (d > T(0)) {
            const int x = idx % W, y = idx / W;
            v[idx] = {(static_cast<T>(x) - cx) * fx_inv * d, (static_cast<T>(y) - cy) * fy_inv * d, d};
        } else {
            v[idx] = {0, 0, 0};
        }
    }
}

// depthJumpThreshold (include/tracking/ICPShared.h:53-55)
template <class T> inline T depthJump(T d) { return smax(static_cast<T>(0.03f), d * static_cast<T>(0.05f)); }

// computeNormals (src/sensor/FrameData.cpp:196-251). Border pixels stay zero.
template <class T>
void normals(const T* depth, const V3<T>* v, int W, int H, V3<T>* n) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const int c = y * W + x;
            n[c] = {0, 0, 0};
            if (x < 1 || x >= W - 1 || y < 1 || y >= H - 1) continue;
            const int r = c + 1, l = c - 1, u = c - W, d = c + W;
            if (depth[c] <= 0 || depth[r] <= 0 || depth[l] <= 0 || depth[u] <= 0 || depth[d] <= 0) continue;
            const T dc = depth[c];
            const T jump = depthJump(dc);
            if (std::abs(dc - depth[r]) > jump || std::abs(dc - depth[l]) > jump ||
                std::abs(dc - depth[u]) > jump || std::abs(dc - depth[d]) > jump)
                continue;
            const V3<T> dx = sub(v[r], v[l]);
            const V3<T> dy = sub(v[d], v[u]);
            const V3<T> nn = cross(dx, dy);
            const T len = norm(nn);
            if (len > static_cast<T>(1e-6f)) n[c] = divs(nn, len);
        }
    }
}

// downsample (src/sensor/FrameData.cpp:262-334), 2x2 block with the depth-jump guard.
template <class T>
void downsample(const T* sd, const V3<T>* sv, const V3<T>* sn, int SW, int SH, T* dd, V3<T>* dv,
                V3<T>* dn) {
    const int DW = SW / 2, DH = SH / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < DH; ++y) {
        for (int x = 0; x < DW; ++x) {
            const int sx = x * 2, sy = y * 2, o = y * DW + x;
            V3<T> sum_v{0, 0, 0}, sum_n{0, 0, 0};
            T sum_d = 0, d_min = 0, d_max = 0;
            int count = 0, min_dx = 0, min_dy = 0;
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int i = (sy + dy) * SW + (sx + dx);
                    const T d = sd[i];
                    if (d > T(0)) {
                        sum_v = add(sum_v, sv[i]);
                        sum_n = add(sum_n, sn[i]);
                        sum_d = sum_d + d;
                        if (count == 0 || d < d_min) { d_min = d; min_dx = dx; min_dy = dy; }
                        if (count == 0 || d > d_max) d_max = d;
                        ++count;
                    }
                }
            dd[o] = 0;
            dv[o] = {0, 0, 0};
            dn[o] = {0, 0, 0};
            if (count == 0) continue;
            if (d_max - d_min > depthJump(d_min)) {
                const int i = (sy + min_dy) * SW + (sx + min_dx);
                dv[o] = sv[i];
                dd[o] = sd[i];
                const T nl = norm(sn[i]);
                dn[o] = (nl > static_cast<T>(1e-6f)) ? divs(sn[i], nl) : V3<T>{0, 0, 0};
            } else {
                const T inv = T(1) / static_cast<T>(count);
                dv[o] = mul(sum_v, inv);
                dd[o] = sum_d * inv;
                const T nl = norm(sum_n);
                dn[o] = (nl > static_cast<T>(1e-6f)) ? divs(sum_n, nl) : V3<T>{0, 0, 0};
            }
        }
    }
}

// ---------------------------------------------------------------------------
// TSDF volume (include/tsdf/TSDFVolume.h:39-62): the reference's 20-byte AoS voxel
// {tsdf, weight, r, g, b}, initialised to the empty state (tsdf 1, weight 0,
// colour 128/255). Colour is carried for layout fidelity and never written (the
// synthetic frames carry no RGB).
// ---------------------------------------------------------------------------
struct Params {
    int res;
    float vs, trunc, max_w;
    float origin[3];
    float min_depth, max_depth;
};

template <class T> struct Voxel {
    T tsdf, weight, r, g, b;
};

template <class T> struct Volume {
    Params p;
    std::vector<Voxel<T>> vox;
    explicit Volume(const Params& pp) : p(pp) {
        vox.resize(static_cast<size_t>(p.res) * p.res * p.res);
        reset();
    }
    void reset() {
        const T ec = static_cast<T>(128.0f / 255.0f);
        std::fill(vox.begin(), vox.end(), Voxel<T>{T(1), T(0), ec, ec, ec});
    }
    size_t idx(int x, int y, int z) const {
        return (static_cast<size_t>(z) * p.res + y) * p.res + x;
    }
};

// integrateCPU (src/tsdf/TSDFVolume.cpp:214-346): voxel-projective update over the
// frustum AABB, camera coordinate carried by the incremental x-step (:286-287).
// Returns the number of voxels updated; *newly gets the newly observed count.
template <class T>
long long integrate(Volume<T>& V, const T* depth, int W, int H, Intr K, const Pose<T>& pose, T min_d,
                    T max_d, long long* newly = nullptr) {
    const T trunc = V.p.trunc, max_w = V.p.max_w, vs = V.p.vs;
    const int res = V.p.res;
    const V3<T> origin{static_cast<T>(V.p.origin[0]), static_cast<T>(V.p.origin[1]),
                       static_cast<T>(V.p.origin[2])};
    const T fx = K.fx, fy = K.fy, cx = K.cx, cy = K.cy;
    const Pose<T> w2c = inverse(pose);
    const M3<T>& R_wc = w2c.R;
    const V3<T> t_wc = w2c.t;
    const M3<T>& R_cw = pose.R;
    const V3<T> t_cw = pose.t;

    V3<T> lo = divs(sub(t_cw, origin), vs);
    V3<T> hi = lo;
    const T z_far = max_d + trunc;
    const T corner_px[4][2] = {{T(0), T(0)},
                               {static_cast<T>(W), T(0)},
                               {T(0), static_cast<T>(H)},
                               {static_cast<T>(W), static_cast<T>(H)}};
    for (const auto& c : corner_px) {
        const V3<T> ray{(c[0] - cx) / fx * z_far, (c[1] - cy) / fy * z_far, z_far};
        const V3<T> v = divs(sub(add(mulv(R_cw, ray), t_cw), origin), vs);
        lo = {smin(lo.x, v.x), smin(lo.y, v.y), smin(lo.z, v.z)};
        hi = {smax(hi.x, v.x), smax(hi.y, v.y), smax(hi.z, v.z)};
    }
    if (!finite3(lo) || !finite3(hi)) return 0;
    auto clampIdx = [res](T f) {
        return static_cast<int>(smax(T(0), smin(static_cast<T>(res - 1), f)));
    };
    const int x0 = clampIdx(std::floor(lo.x)), x1 = clampIdx(std::ceil(hi.x));
    const int y0 = clampIdx(std::floor(lo.y)), y1 = clampIdx(std::ceil(hi.y));
    const int z0 = clampIdx(std::floor(lo.z)), z1 = clampIdx(std::ceil(hi.z));

    const V3<T> step_x{R_wc.m[0][0] * vs, R_wc.m[1][0] * vs, R_wc.m[2][0] * vs};
    long long updated = 0, nobs = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 16) reduction(+ : updated, nobs)
    for (int z = z0; z <= z1; ++z) {
        for (int y = y0; y <= y1; ++y) {
            const V3<T> p{origin.x + static_cast<T>(x0) * vs, origin.y + static_cast<T>(y) * vs,
                          origin.z + static_cast<T>(z) * vs};
            V3<T> c = add(mulv(R_wc, p), t_wc);
            for (int x = x0; x <= x1; ++x, c = add(c, step_x)) {
                if (c.z <= T(0)) continue;
                const T inv_z = T(1) / c.z;
                const T uf = std::floor(fx * c.x * inv_z + cx + static_cast<T>(0.5f));
                const T vf = std::floor(fy * c.y * inv_z + cy + static_cast<T>(0.5f));
                if (!(uf >= T(0) && uf < static_cast<T>(W) && vf >= T(0) && vf < static_cast<T>(H)))
                    continue;
                const int pix = static_cast<int>(vf) * W + static_cast<int>(uf);
                const T d = depth[pix];
                if (!(d >= min_d && d <= max_d)) continue;
                const T sdf = d - c.z;
                if (sdf < -trunc) continue;
                const T tsdf_new = smin(T(1), sdf / trunc);
                Voxel<T>& vx = V.vox[V.idx(x, y, z)];
                const T w_old = vx.weight;
                nobs += (w_old == T(0)) ? 1 : 0;
                vx.tsdf = (vx.tsdf * w_old + tsdf_new) / (w_old + T(1));
                vx.weight = smin(w_old + T(1), max_w);
                ++updated;
            }
        }
    }
    if (newly) *newly = nobs;
    return updated;
}

// sampleTSDF (src/tsdf/TSDFVolume.cpp:492-525): trilinear, valid only when all 8
// corners carry weight > 0. When `touched` is given, every in-bounds corner read is
// marked (distinct-voxel accounting for the raycast byte count).
template <class T>
inline bool sampleTSDF(const Volume<T>& V, V3<T> pw, T& f_out, std::vector<uint8_t>* touched = nullptr) {
    const T vs = V.p.vs;
    const int res = V.p.res;
    const V3<T> v{(pw.x - static_cast<T>(V.p.origin[0])) / vs, (pw.y - static_cast<T>(V.p.origin[1])) / vs,
                  (pw.z - static_cast<T>(V.p.origin[2])) / vs};
    const T fx0 = std::floor(v.x), fy0 = std::floor(v.y), fz0 = std::floor(v.z);
    if (!(fx0 >= T(0) && fy0 >= T(0) && fz0 >= T(0) && fx0 < static_cast<T>(res - 1) &&
          fy0 < static_cast<T>(res - 1) && fz0 < static_cast<T>(res - 1)))
        return false;
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0), z0 = static_cast<int>(fz0);
    const T tx = v.x - fx0, ty = v.y - fy0, tz = v.z - fz0;
    const size_t sx = 1, sy = static_cast<size_t>(res), sz = static_cast<size_t>(res) * res;
    const size_t b = V.idx(x0, y0, z0);
    const size_t c[8] = {b, b + sx, b + sy, b + sx + sy, b + sz, b + sx + sz, b + sy + sz, b + sx + sy + sz};
    if (touched)
        for (size_t ci : c) (*touched)[ci] = 1;
    for (size_t ci : c)
        if (!(V.vox[ci].weight > T(0))) return false;
    const T v00 = V.vox[c[0]].tsdf * (T(1) - tx) + V.vox[c[1]].tsdf * tx;
    const T v10 = V.vox[c[2]].tsdf * (T(1) - tx) + V.vox[c[3]].tsdf * tx;
    const T v01 = V.vox[c[4]].tsdf * (T(1) - tx) + V.vox[c[5]].tsdf * tx;
    const T v11 = V.vox[c[6]].tsdf * (T(1) - tx) + V.vox[c[7]].tsdf * tx;
    const T v0 = v00 * (T(1) - ty) + v10 * ty;
    const T v1 = v01 * (T(1) - ty) + v11 * ty;
    f_out = v0 * (T(1) - tz) + v1 * tz;
    return true;
}

// computeNormal (src/tsdf/TSDFVolume.cpp:478-490): normalised central difference of
// six valid samples at +-voxel_size.
template <class T>
inline bool computeNormal(const Volume<T>& V, V3<T> p, V3<T>& n_out, std::vector<uint8_t>* touched = nullptr) {
    const T h = V.p.vs;
    T s[6];
    const V3<T> off[3] = {{h, 0, 0}, {0, h, 0}, {0, 0, h}};
    for (int a = 0; a < 3; ++a) {
        if (!sampleTSDF(V, add(p, off[a]), s[2 * a], touched) ||
            !sampleTSDF(V, sub(p, off[a]), s[2 * a + 1], touched))
            return false;
    }
    const V3<T> n{s[0] - s[1], s[2] - s[3], s[4] - s[5]};
    const T len = norm(n);
    if (!(len > static_cast<T>(1e-6f)) || !std::isfinite(len)) return false;
    n_out = divs(n, len);
    return true;
}

// One ray of raycast (src/tsdf/TSDFVolume.cpp:384-473): slab clip, adaptive march,
// first valid +-to-- crossing refined by one secant step, then the normal.
// Returns true on a hit; *t_hit_out receives the ray parameter.
template <class T>
inline bool castRay(const Volume<T>& V, const Pose<T>& pose, Intr K, int x, int y, V3<T>& vert, V3<T>& nrm,
                    std::vector<uint8_t>* touched = nullptr, long long* nsamples = nullptr) {
    const T vs = V.p.vs, trunc = V.p.trunc;
    const V3<T> cam = pose.t;
    const V3<T> box_lo{static_cast<T>(V.p.origin[0]), static_cast<T>(V.p.origin[1]),
                       static_cast<T>(V.p.origin[2])};
    const T span = static_cast<T>(V.p.res - 1) * vs;
    const V3<T> box_hi{box_lo.x + span, box_lo.y + span, box_lo.z + span};
    const T min_step = static_cast<T>(0.5f) * vs;
    const T far_step = smax(min_step, static_cast<T>(0.8f) * trunc);

    const V3<T> ray_cam{(static_cast<T>(x) - static_cast<T>(K.cx)) / static_cast<T>(K.fx),
                        (static_cast<T>(y) - static_cast<T>(K.cy)) / static_cast<T>(K.fy), T(1)};
    const T ray_norm = norm(ray_cam);
    const V3<T> dir = divs(mulv(pose.R, ray_cam), ray_norm);
    T t_near = static_cast<T>(V.p.min_depth) * ray_norm;
    T t_far = static_cast<T>(V.p.max_depth) * ray_norm;
    const T cam_a[3] = {cam.x, cam.y, cam.z}, dir_a[3] = {dir.x, dir.y, dir.z};
    const T lo_a[3] = {box_lo.x, box_lo.y, box_lo.z}, hi_a[3] = {box_hi.x, box_hi.y, box_hi.z};
    for (int a = 0; a < 3; ++a) {
        const T inv = T(1) / dir_a[a];
        T t0 = (lo_a[a] - cam_a[a]) * inv;
        T t1 = (hi_a[a] - cam_a[a]) * inv;
        if (t0 > t1) std::swap(t0, t1);
        if (std::isnan(t0) || std::isnan(t1)) {
            if (cam_a[a] < lo_a[a] || cam_a[a] > hi_a[a]) t_far = T(-1);
            continue;
        }
        t_near = smax(t_near, t0);
        t_far = smin(t_far, t1);
    }
    if (!(t_near < t_far)) return false;

    T t_prev = t_near, f_prev = T(1);
    bool prev_valid = sampleTSDF(V, add(cam, mul(dir, t_prev)), f_prev, touched);
    long long ns = 1;
    if (prev_valid && !std::isfinite(f_prev)) { if (nsamples) *nsamples += ns; return false; }
    if (prev_valid && f_prev < T(0)) { if (nsamples) *nsamples += ns; return false; }
    bool found = false;
    T t_hit = 0;
    while (t_prev < t_far) {
        const T step = (!prev_valid || f_prev >= static_cast<T>(0.999f))
                           ? far_step
                           : smax(min_step, static_cast<T>(0.8f) * f_prev * trunc);
        const T t = smin(t_prev + step, t_far);
        T f = T(1);
        const bool valid = sampleTSDF(V, add(cam, mul(dir, t)), f, touched);
        ++ns;
        if (valid && !std::isfinite(f)) break;
        if (valid && prev_valid) {
            if (f_prev > T(0) && f <= T(0)) {
                t_hit = t_prev + (t - t_prev) * (f_prev / (f_prev - f));
                found = true;
                break;
            }
            if (f_prev < T(0) && f >= T(0)) break;
        } else if (valid && f < T(0)) {
            break;
        }
        t_prev = t;
        f_prev = f;
        prev_valid = valid;
        if (t >= t_far) break;
    }
    if (nsamples) *nsamples += ns;
    if (!found) return false;
    const V3<T> hit = add(cam, mul(dir, t_hit));
    V3<T> n;
    if (!computeNormal(V, hit, n, touched)) return false;
    if (!finite3(hit)) return false;
    vert = hit;
    nrm = n;
    return true;
}

// raycast (src/tsdf/TSDFVolume.cpp:348-476). A miss leaves an all-zero output.
template <class T>
void raycast(const Volume<T>& V, const Pose<T>& pose, Intr K, int W, int H, V3<T>* verts, V3<T>* norms,
             std::vector<uint8_t>* touched = nullptr, long long* nsamples = nullptr) {
    long long ns_total = 0;
    if (touched) {
        // Serial when accounting, so the bitset has one writer.
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int o = y * W + x;
                verts[o] = {0, 0, 0};
                norms[o] = {0, 0, 0};
                V3<T> v, n;
                if (castRay(V, pose, K, x, y, v, n, touched, &ns_total)) { verts[o] = v; norms[o] = n; }
            }
    } else {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : ns_total)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const int o = y * W + x;
                verts[o] = {0, 0, 0};
                norms[o] = {0, 0, 0};
                V3<T> v, n;
                if (castRay(V, pose, K, x, y, v, n, nullptr, &ns_total)) { verts[o] = v; norms[o] = n; }
            }
    }
    if (nsamples) *nsamples = ns_total;
}

// ---------------------------------------------------------------------------
// ICP (src/tracking/ICPTracker.cpp:40-354, include/tracking/ICPShared.h:34-110)
// ---------------------------------------------------------------------------
struct IcpParams {
    int iters[3];  // per pyramid level 0, 1, 2 ({10, 5, 4})
    float dist_threshold;
    float angle_threshold;  // degrees
};

template <class T> struct IcpResult {
    T pose[16];
    T error;
    int inliers;
    int converged;
    int tracking_ok;
    T final_step;
    T information[36];
    int valid_live, valid_model, projected, dist_filtered, angle_filtered;
    int steps;  // accepted steps, packed one byte per level (level l at bit 8*l)
};

// One live pyramid level: W x H vertices / normals in camera coordinates.
template <class T> struct Level {
    const V3<T>* v;
    const V3<T>* n;
    int W, H;
};

constexpr float kHuberK = 0.02f;

// Fixed virtual-thread partition of rows [1, H-1) into icpChunks() contiguous
// blocks, folded in block order (the reference's schedule(static) + ordered fold,
// ICPTracker.cpp:228, 334-351, with the thread count pinned so the sum order does
// not depend on the host). Default 8; settable to model other reference thread counts.
inline int& icpChunks() {
    static int n = 8;
    return n;
}
constexpr int kIcpMaxChunks = 64;

struct IcpSystemCounters {
    int valid_live, valid_model, projected, dist_filtered, angle_filtered, inliers;
};

// buildLinearSystem (ICPTracker.cpp:136-354). A and b are full 6x6 / 6 (row-major A).
template <class T>
bool buildLinearSystem(const Level<T>& live, const V3<T>* mv, const V3<T>* mn, int MW, int MH, Intr K0,
                       const Pose<T>& pose, const Pose<T>& ref_pose, const IcpParams& prm, T A[36], T b[6],
                       T& residual, IcpSystemCounters& cnt) {
    const int W = live.W, H = live.H;
    const Intr K = scaleIntrinsics(K0, 640, 480, MW, MH);
    const T angle_cos = static_cast<T>(static_cast<float>(std::cos(prm.angle_threshold * M_PI / 180.0f)));
    const M3<T>& R_cw = pose.R;
    const V3<T> t_cw = pose.t;
    const Pose<T> ref_inv = inverse(ref_pose);
    const M3<T> R_rel = mulm(ref_inv.R, R_cw);
    const V3<T> t_rel = add(mulv(ref_inv.R, t_cw), ref_inv.t);
    const T dist_sq_max = static_cast<T>(prm.dist_threshold) * static_cast<T>(prm.dist_threshold);
    const T kK = static_cast<T>(kHuberK);

    struct Acc {
        T A[21], b[6], res;
        int count, vl, vm, pr, df, af;
    };
    const int nch = std::max(1, std::min(icpChunks(), kIcpMaxChunks));
    Acc acc[kIcpMaxChunks];
    for (auto& a : acc) {
        for (T& x : a.A) x = 0;
        for (T& x : a.b) x = 0;
        a.res = 0;
        a.count = a.vl = a.vm = a.pr = a.df = a.af = 0;
    }
    const int rows = std::max(0, H - 2);
#pragma omp parallel for schedule(static, 1)
    for (int ch = 0; ch < nch; ++ch) {
        const int base = rows / nch, extra = rows % nch;
        const int r0 = 1 + ch * base + std::min(ch, extra);
        const int r1 = r0 + base + (ch < extra ? 1 : 0);
        Acc& a = acc[ch];
        for (int y = r0; y < r1; ++y) {
            for (int x = 1; x < W - 1; ++x) {
                const int idx = y * W + x;
                const V3<T> lv = live.v[idx];
                if (lv.z <= static_cast<T>(0.001f) || !finite3(lv)) continue;
                const V3<T> vr = add(mulv(R_rel, lv), t_rel);
                if (vr.z <= static_cast<T>(0.001f)) continue;
                a.vl++;
                const T inv_z = T(1) / vr.z;
                const T mxf = static_cast<T>(K.fx) * vr.x * inv_z + static_cast<T>(K.cx);
                const T myf = static_cast<T>(K.fy) * vr.y * inv_z + static_cast<T>(K.cy);
                int mx = 0, my = 0;
                if (!floorToInt(mxf + static_cast<T>(0.5f), &mx) || !floorToInt(myf + static_cast<T>(0.5f), &my))
                    continue;
                if (mx < 0 || mx >= MW || my < 0 || my >= MH) continue;
                a.pr++;
                const int midx = my * MW + mx;
                const V3<T> mvw = mv[midx], mnw = mn[midx];
                if (!(finite3(mvw) && sqnorm(mvw) > static_cast<T>(1e-12f)) ||
                    !(finite3(mnw) && sqnorm(mnw) > static_cast<T>(0.9f)))
                    continue;
                a.vm++;
                const V3<T> vlw = add(mulv(R_cw, lv), t_cw);
                const T dist_sq = sqnorm(sub(vlw, mvw));
                if (dist_sq > dist_sq_max) { a.df++; continue; }
                const V3<T> ln = live.n[idx];
                if (!(finite3(ln) && sqnorm(ln) > static_cast<T>(0.9f))) continue;
                const V3<T> nlw = mulv(R_cw, ln);
                const T dp = std::abs(dot(nlw, mnw));
                if (dp < angle_cos) { a.af++; continue; }
                const T err = dot(mnw, sub(vlw, mvw));
                if (!std::isfinite(err)) continue;
                const V3<T> nlc = mulvT(R_cw, mnw);
                const V3<T> cr = cross(lv, nlc);
                const T J[6] = {nlc.x, nlc.y, nlc.z, cr.x, cr.y, cr.z};
                bool fin = true;
                for (T j : J) fin = fin && std::isfinite(j);
                if (!fin) continue;
                const T ae = std::abs(err);
                const T w = (ae <= kK) ? T(1) : kK / ae;
                const T we = err * w;
                const T loss = (ae <= kK) ? ae * ae : T(2) * kK * ae - kK * kK;
                int k = 0;
                for (int i = 0; i < 6; ++i) {
                    for (int j = i; j < 6; ++j) a.A[k++] += w * J[i] * J[j];
                    a.b[i] -= J[i] * we;
                }
                a.res += loss;
                a.count++;
            }
        }
    }
    for (int i = 0; i < 36; ++i) A[i] = 0;
    for (int i = 0; i < 6; ++i) b[i] = 0;
    residual = 0;
    cnt = IcpSystemCounters{0, 0, 0, 0, 0, 0};
    for (int ch = 0; ch < nch; ++ch) {
        const Acc& a = acc[ch];
        int k = 0;
        for (int i = 0; i < 6; ++i) {
            for (int j = i; j < 6; ++j) {
                const T val = a.A[k++];
                A[i * 6 + j] += val;
                if (i != j) A[j * 6 + i] += val;
            }
            b[i] += a.b[i];
        }
        residual += a.res;
        cnt.inliers += a.count;
        cnt.valid_live += a.vl;
        cnt.valid_model += a.vm;
        cnt.projected += a.pr;
        cnt.dist_filtered += a.df;
        cnt.angle_filtered += a.af;
    }
    return cnt.inliers > 0;
}

// Cyclic Jacobi eigenvalues of a symmetric n x n matrix, in double (textbook method,
// Golub & Van Loan 8.5). Eigen's SelfAdjointEigenSolver is the reference's routine
// (ICPShared.h:86-100); only the extreme eigenvalues are used.
inline void jacobiEigenvalues(const double* Ain, int n, double* eig) {
    double a[36];
    for (int i = 0; i < n * n; ++i) a[i] = Ain[i];
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
        if (off == 0) break;
        double diag = 0;
        for (int i = 0; i < n; ++i) diag += a[i * n + i] * a[i * n + i];
        if (off <= 1e-32 * diag) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (apq == 0) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    for (int i = 0; i < n; ++i) eig[i] = a[i * n + i];
    std::sort(eig, eig + n);
}

// dampingForHessian (include/tracking/ICPShared.h:86-100)
template <class T> inline T dampingForHessian(const T A[36], int level) {
    double Ad[36], eig[6];
    for (int i = 0; i < 36; ++i) Ad[i] = static_cast<double>(A[i]);
    jacobiEigenvalues(Ad, 6, eig);
    const double mn = eig[0], mx = eig[5];
    if (!std::isfinite(mn) || !std::isfinite(mx) || mn < static_cast<double>(1e-4f)) return T(1);
    if (mx > static_cast<double>(1e7f) * mn) return T(1);
    return level == 0 ? static_cast<T>(0.01f) : static_cast<T>(0.1f);
}

// Symmetric LDL^T with diagonal pivoting, then solve (the reference's A.ldlt().solve,
// ICPTracker.cpp:93). Returns false when a pivot vanishes.
template <class T> inline bool ldltSolve(const T Ain[36], const T bin[6], T x[6]) {
    T a[36], b[6];
    int perm[6];
    for (int i = 0; i < 36; ++i) a[i] = Ain[i];
    for (int i = 0; i < 6; ++i) { b[i] = bin[i]; perm[i] = i; }
    // Symmetric pivoting: bring the largest remaining diagonal to position k.
    for (int k = 0; k < 6; ++k) {
        int piv = k;
        T best = std::abs(a[k * 6 + k]);
        for (int i = k + 1; i < 6; ++i)
            if (std::abs(a[i * 6 + i]) > best) { best = std::abs(a[i * 6 + i]); piv = i; }
        if (piv != k) {
            for (int j = 0; j < 6; ++j) std::swap(a[k * 6 + j], a[piv * 6 + j]);
            for (int j = 0; j < 6; ++j) std::swap(a[j * 6 + k], a[j * 6 + piv]);
            std::swap(b[k], b[piv]);
            std::swap(perm[k], perm[piv]);
        }
        const T d = a[k * 6 + k];
        if (d == T(0) || !std::isfinite(d)) return false;
        T col[6];
        for (int i = k + 1; i < 6; ++i) col[i] = a[i * 6 + k];
        for (int i = k + 1; i < 6; ++i) {
            const T l = col[i] / d;
            for (int j = k + 1; j <= i; ++j) a[i * 6 + j] -= l * col[j];
            a[i * 6 + k] = l;
        }
        for (int i = k + 1; i < 6; ++i)
            for (int j = k + 1; j < i; ++j) a[j * 6 + i] = a[i * 6 + j];
    }
    // Forward (unit L), diagonal, backward (L^T).
    T y[6];
    for (int i = 0; i < 6; ++i) {
        T s = b[i];
        for (int j = 0; j < i; ++j) s -= a[i * 6 + j] * y[j];
        y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] = y[i] / a[i * 6 + i];
    T z[6];
    for (int i = 5; i >= 0; --i) {
        T s = y[i];
        for (int j = i + 1; j < 6; ++j) s -= a[j * 6 + i] * z[j];
        z[i] = s;
    }
    for (int i = 0; i < 6; ++i) x[perm[i]] = z[i];
    return true;
}

// Nearest rotation to M (projectToSO3, include/tracking/ICPShared.h:102-110): U V^T
// from a one-sided Jacobi SVD in double (Hestenes), with the column of U that belongs
// to the smallest singular value negated when det(U V^T) < 0.
inline void projectToSO3d(const double M[9], double R[9]) {
    double U[9], V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i) U[i] = M[i];  // row-major; columns are orthogonalised
    for (int sweep = 0; sweep < 60; ++sweep) {
        double offmax = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int k = 0; k < 3; ++k) {
                    alpha += U[k * 3 + p] * U[k * 3 + p];
                    beta += U[k * 3 + q] * U[k * 3 + q];
                    gamma += U[k * 3 + p] * U[k * 3 + q];
                }
                const double denom = std::sqrt(alpha * beta);
                if (denom == 0) continue;
                offmax = std::max(offmax, std::abs(gamma) / denom);
                if (std::abs(gamma) <= 1e-300) continue;
                const double zeta = (beta - alpha) / (2 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t), s = c * t;
                for (int k = 0; k < 3; ++k) {
                    const double up = U[k * 3 + p], uq = U[k * 3 + q];
                    U[k * 3 + p] = c * up - s * uq;
                    U[k * 3 + q] = s * up + c * uq;
                    const double vp = V[k * 3 + p], vq = V[k * 3 + q];
                    V[k * 3 + p] = c * vp - s * vq;
                    V[k * 3 + q] = s * vp + c * vq;
                }
            }
        if (offmax < 1e-15) break;
    }
    double sig[3];
    for (int j = 0; j < 3; ++j) {
        sig[j] = std::sqrt(U[0 * 3 + j] * U[0 * 3 + j] + U[1 * 3 + j] * U[1 * 3 + j] + U[2 * 3 + j] * U[2 * 3 + j]);
        if (sig[j] > 0)
            for (int k = 0; k < 3; ++k) U[k * 3 + j] /= sig[j];
    }
    // Rank-deficient columns: complete U to an orthonormal basis.
    for (int j = 0; j < 3; ++j) {
        if (sig[j] > 0) continue;
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        const double ca[3] = {U[a], U[3 + a], U[6 + a]}, cb[3] = {U[b], U[3 + b], U[6 + b]};
        double c[3] = {ca[1] * cb[2] - ca[2] * cb[1], ca[2] * cb[0] - ca[0] * cb[2], ca[0] * cb[1] - ca[1] * cb[0]};
        const double l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; ++k) U[k * 3 + j] = l > 0 ? c[k] / l : (k == j ? 1.0 : 0.0);
    }
    auto compose = [&](double* out) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) out[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] + U[i * 3 + 2] * V[j * 3 + 2];
    };
    compose(R);
    const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                       R[2] * (R[3] * R[7] - R[4] * R[6]);
    if (det < 0) {
        int js = 0;
        for (int j = 1; j < 3; ++j)
            if (sig[j] < sig[js]) js = j;
        for (int k = 0; k < 3; ++k) U[k * 3 + js] = -U[k * 3 + js];
        compose(R);
    }
}

// trackLevel (ICPTracker.cpp:68-134)
template <class T>
IcpResult<T> trackLevel(const Level<T>& live, const V3<T>* mv, const V3<T>* mn, int MW, int MH, Intr K,
                        const Pose<T>& est, const Pose<T>& ref, const IcpParams& prm, int level, int max_iter) {
    IcpResult<T> r{};
    Pose<T> pose = est;
    r.final_step = std::numeric_limits<T>::infinity();
    r.error = 0;
    for (int it = 0; it < max_iter; ++it) {
        T A[36], b[6], residual;
        IcpSystemCounters c;
        const bool ok = buildLinearSystem(live, mv, mn, MW, MH, K, pose, ref, prm, A, b, residual, c);
        r.valid_live = c.valid_live;
        r.valid_model = c.valid_model;
        r.projected = c.projected;
        r.dist_filtered = c.dist_filtered;
        r.angle_filtered = c.angle_filtered;
        if (!ok) break;
        if (c.inliers < 10) break;
        for (int i = 0; i < 36; ++i) r.information[i] = A[i];
        const T lam = dampingForHessian(A, level);
        for (int i = 0; i < 6; ++i) A[i * 6 + i] += lam;
        T x[6];
        if (!ldltSolve(A, b, x)) break;
        bool fin = true;
        for (T v : x) fin = fin && std::isfinite(v);
        if (!fin) break;
        const T tn = std::sqrt(sum3(x[0] * x[0], x[1] * x[1], x[2] * x[2]));
        const T rn = std::sqrt(sum3(x[3] * x[3], x[4] * x[4], x[5] * x[5]));
        if (tn > static_cast<T>(0.2f) || rn > static_cast<T>(0.5f)) break;
        T sq = 0;
        for (int i = 0; i < 6; ++i) sq += x[i] * x[i];
        const T step_norm = std::sqrt(sq);
        const T rx = x[3], ry = x[4], rz = x[5];
        M3<T> Rd{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
        const T angle = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (angle > static_cast<T>(1e-4f)) {
            // Rodrigues' formula R = cI + s[k]x + (1-c) k k^T for the unit axis k.
            const T nn = std::sqrt(sum3(rx * rx, ry * ry, rz * rz));
            const V3<T> k{rx / nn, ry / nn, rz / nn};
            const T s = std::sin(angle), co = std::cos(angle), oc = T(1) - co;
            Rd.m[0][0] = oc * k.x * k.x + co;
            Rd.m[1][1] = oc * k.y * k.y + co;
            Rd.m[2][2] = oc * k.z * k.z + co;
            Rd.m[0][1] = oc * k.x * k.y - s * k.z;
            Rd.m[1][0] = oc * k.x * k.y + s * k.z;
            Rd.m[0][2] = oc * k.x * k.z + s * k.y;
            Rd.m[2][0] = oc * k.x * k.z - s * k.y;
            Rd.m[1][2] = oc * k.y * k.z - s * k.x;
            Rd.m[2][1] = oc * k.y * k.z + s * k.x;
        } else {
            Rd.m[0][1] = -rz; Rd.m[0][2] = ry;
            Rd.m[1][0] = rz;  Rd.m[1][2] = -rx;
            Rd.m[2][0] = -ry; Rd.m[2][1] = rx;
        }
        // pose <- pose * [Rd | t] (Eigen 4x4 packet order), then re-projected to SO(3).
        Pose<T> np;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j)
                np.R.m[i][j] = ((pose.R.m[i][0] * Rd.m[0][j] + pose.R.m[i][1] * Rd.m[1][j]) + pose.R.m[i][2] * Rd.m[2][j]);
        }
        const T tv[3] = {x[0], x[1], x[2]};
        const T pt[3] = {pose.t.x, pose.t.y, pose.t.z};
        T nt[3];
        for (int i = 0; i < 3; ++i)
            nt[i] = ((pose.R.m[i][0] * tv[0] + pose.R.m[i][1] * tv[1]) + pose.R.m[i][2] * tv[2]) + pt[i];
        np.t = {nt[0], nt[1], nt[2]};
        double Md[9], Rp[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Md[i * 3 + j] = static_cast<double>(np.R.m[i][j]);
        projectToSO3d(Md, Rp);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) np.R.m[i][j] = static_cast<T>(Rp[i * 3 + j]);
        pose = np;
        r.final_step = step_norm;
        r.steps += 1 << (8 * level);
        r.error = residual / static_cast<T>(std::max(c.inliers, 1));
        r.inliers = c.inliers;
        if (step_norm < static_cast<T>(5e-5f)) {
            r.converged = 1;
            break;
        }
    }
    poseTo(pose, r.pose);
    return r;
}

// track (ICPTracker.cpp:40-66). levels[l] is pyramid level l (0 = full resolution).
template <class T>
IcpResult<T> track(const Level<T> levels[3], const V3<T>* mv, const V3<T>* mn, int MW, int MH, Intr K,
                   const Pose<T>& est, const Pose<T>& ref, const IcpParams& prm) {
    IcpResult<T> result{};
    poseTo(est, result.pose);
    result.final_step = std::numeric_limits<T>::infinity();
    T last_final = std::numeric_limits<T>::infinity();
    for (int level = 2; level >= 0; --level) {
        if (prm.iters[level] <= 0) continue;
        const Pose<T> cur = poseFrom<T>(result.pose);
        IcpResult<T> lr = trackLevel(levels[level], mv, mn, MW, MH, K, cur, ref, prm, level, prm.iters[level]);
        if (std::isfinite(lr.final_step)) last_final = lr.final_step;
        lr.final_step = last_final;
        const int steps_before = result.steps;
        result = lr;
        result.steps += steps_before;
    }
    bool fin = true;
    for (T v : result.pose) fin = fin && std::isfinite(v);
    result.tracking_ok = fin && result.inliers > 100 &&
                         (result.converged || result.final_step <= static_cast<T>(1e-3f));
    return result;
}

}  // namespace orc
