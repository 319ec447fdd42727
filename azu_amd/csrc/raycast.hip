// Raycast on gfx950 (the reference's TSDFVolume::raycast / sampleTSDF / computeNormal,
// src/tsdf/TSDFVolume.cpp:348-525): per pixel, slab-clip the ray to the volume box,
// march with the adaptive step (far step while unobserved or f >= 0.999, else
// max(vs/2, 0.8 f tau)), stop at the first observed + -> - crossing, refine it with one
// secant step and take the normal from six trilinear samples at +-vs.
//
// Layout: a 256-thread block owns a 16x16 pixel tile, each wave an 8x8 sub-tile, so
// the rays of a wave march through neighbouring voxels. Blocks are remapped so that
// each of the 8 XCDs (block b runs on XCD b % 8) handles one contiguous band of
// tiles and its L2 holds the voxels of that band. Each trilinear sample reads its 8
// corners as 4 x-adjacent pairs (16 B per pair for FP32, 8 B for FP16/U16).
//
// Exact = true (FP32 parity path): IEEE division where the reference divides, no
// contraction, bit-identical to the CPU restatement. Exact = false (FP16 form):
// multiplications by 1/vs and fmaf interpolation.
#include "common.hpp"

#ifndef RAY_SPEC32
#define RAY_SPEC32 4
#endif
#ifndef RAY_SPEC16
#define RAY_SPEC16 0
#endif

namespace azu {

template <class S> struct PairIO;
template <> struct PairIO<StoreF32> {
    struct P {
        float2 a, b;
    };
    __device__ static P load(const float2* p) {
        P r;
        __builtin_memcpy(&r, p, sizeof(P));
        return r;
    }
};
template <> struct PairIO<StoreF16> {
    struct P {
        uint32_t a, b;
    };
    __device__ static P load(const uint32_t* p) {
        P r;
        __builtin_memcpy(&r, p, sizeof(P));
        return r;
    }
};

template <class S> struct Vol {
    const typename S::Word* vox;
    int res;
    float vs, inv_vs;
    float ox, oy, oz;
    FixedDiv dvs;  // division by vs
};

// sampleTSDF (TSDFVolume.cpp:492-525) in two halves, so a ray can issue the reads of
// several samples before it evaluates the first: sampleLoad locates the cell and reads
// its 8 corners (nothing is read outside the volume), sampleEval applies the weight
// gate and interpolates.
template <class S> struct Cell {
    typename PairIO<S>::P q0, q1, q2, q3;
    float tx, ty, tz;
    bool inside;
};

template <class S, bool E, int FD>
__device__ __forceinline__ Cell<S> sampleLoad(const Vol<S>& V, F3 p) {
    Cell<S> c;
    float gx, gy, gz;
    if (E) {
        gx = divBy<FD>(V.dvs, p.x - V.ox);
        gy = divBy<FD>(V.dvs, p.y - V.oy);
        gz = divBy<FD>(V.dvs, p.z - V.oz);
    } else {
        gx = (p.x - V.ox) * V.inv_vs;
        gy = (p.y - V.oy) * V.inv_vs;
        gz = (p.z - V.oz) * V.inv_vs;
    }
    const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
    const float top = (float)(V.res - 1);
    c.inside = fx0 >= 0.0f && fy0 >= 0.0f && fz0 >= 0.0f && fx0 < top && fy0 < top && fz0 < top;
    const int x0 = c.inside ? (int)fx0 : 0, y0 = c.inside ? (int)fy0 : 0, z0 = c.inside ? (int)fz0 : 0;
    c.tx = gx - fx0;
    c.ty = gy - fy0;
    c.tz = gz - fz0;
    const size_t sy = (size_t)V.res, sz = (size_t)V.res * (size_t)V.res;
    const size_t b = (size_t)z0 * sz + (size_t)y0 * sy + (size_t)x0;
    using IO = PairIO<S>;
    c.q0 = IO::load(V.vox + b);
    c.q1 = IO::load(V.vox + b + sy);
    c.q2 = IO::load(V.vox + b + sz);
    c.q3 = IO::load(V.vox + b + sy + sz);
    return c;
}

template <class S, bool E>
__device__ __forceinline__ bool sampleEval(const Cell<S>& c, float& f) {
    if (!c.inside) return false;
    if (!S::observed8(c.q0.a, c.q0.b, c.q1.a, c.q1.b, c.q2.a, c.q2.b, c.q3.a, c.q3.b)) return false;
    const float tx = c.tx, ty = c.ty, tz = c.tz;
    const float c0 = S::tsdf(c.q0.a), c1 = S::tsdf(c.q0.b), c2 = S::tsdf(c.q1.a), c3 = S::tsdf(c.q1.b);
    const float c4 = S::tsdf(c.q2.a), c5 = S::tsdf(c.q2.b), c6 = S::tsdf(c.q3.a), c7 = S::tsdf(c.q3.b);
    if (E) {
        const float v00 = c0 * (1.0f - tx) + c1 * tx;
        const float v10 = c2 * (1.0f - tx) + c3 * tx;
        const float v01 = c4 * (1.0f - tx) + c5 * tx;
        const float v11 = c6 * (1.0f - tx) + c7 * tx;
        const float v0 = v00 * (1.0f - ty) + v10 * ty;
        const float v1 = v01 * (1.0f - ty) + v11 * ty;
        f = v0 * (1.0f - tz) + v1 * tz;
    } else {
        const float v00 = fmaf(c1 - c0, tx, c0), v10 = fmaf(c3 - c2, tx, c2);
        const float v01 = fmaf(c5 - c4, tx, c4), v11 = fmaf(c7 - c6, tx, c6);
        const float v0 = fmaf(v10 - v00, ty, v00), v1 = fmaf(v11 - v01, ty, v01);
        f = fmaf(v1 - v0, tz, v0);
    }
    return true;
}

template <class S, bool E, int FD>
__device__ __forceinline__ bool sampleVol(const Vol<S>& V, F3 p, float& f) {
    float gx, gy, gz;
    if (E) {
        gx = divBy<FD>(V.dvs, p.x - V.ox);
        gy = divBy<FD>(V.dvs, p.y - V.oy);
        gz = divBy<FD>(V.dvs, p.z - V.oz);
    } else {
        gx = (p.x - V.ox) * V.inv_vs;
        gy = (p.y - V.oy) * V.inv_vs;
        gz = (p.z - V.oz) * V.inv_vs;
    }
    const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
    const float top = (float)(V.res - 1);
    if (!(fx0 >= 0.0f && fy0 >= 0.0f && fz0 >= 0.0f && fx0 < top && fy0 < top && fz0 < top)) return false;
    Cell<S> c;
    c.inside = true;
    c.tx = gx - fx0;
    c.ty = gy - fy0;
    c.tz = gz - fz0;
    const size_t sy = (size_t)V.res, sz = (size_t)V.res * (size_t)V.res;
    const size_t b = (size_t)(int)fz0 * sz + (size_t)(int)fy0 * sy + (size_t)(int)fx0;
    using IO = PairIO<S>;
    c.q0 = IO::load(V.vox + b);
    c.q1 = IO::load(V.vox + b + sy);
    c.q2 = IO::load(V.vox + b + sz);
    c.q3 = IO::load(V.vox + b + sy + sz);
    return sampleEval<S, E>(c, f);
}

// computeNormal (TSDFVolume.cpp:478-490)
template <class S, bool E, int FD>
__device__ __forceinline__ bool normalVol(const Vol<S>& V, F3 p, F3& n) {
    const float h = V.vs;
    float s0, s1, s2, s3, s4, s5;
    if (!sampleVol<S, E, FD>(V, F3{p.x + h, p.y + 0.0f, p.z + 0.0f}, s0) ||
        !sampleVol<S, E, FD>(V, F3{p.x - h, p.y - 0.0f, p.z - 0.0f}, s1) ||
        !sampleVol<S, E, FD>(V, F3{p.x + 0.0f, p.y + h, p.z + 0.0f}, s2) ||
        !sampleVol<S, E, FD>(V, F3{p.x - 0.0f, p.y - h, p.z - 0.0f}, s3) ||
        !sampleVol<S, E, FD>(V, F3{p.x + 0.0f, p.y + 0.0f, p.z + h}, s4) ||
        !sampleVol<S, E, FD>(V, F3{p.x - 0.0f, p.y - 0.0f, p.z - h}, s5))
        return false;
    const F3 g{s0 - s1, s2 - s3, s4 - s5};
    const float len = f3norm(g);
    if (!(len > 1e-6f) || !isfinite(len)) return false;
    n = f3div(g, len);
    return true;
}

struct RayOut {
    float* verts;  // W*H packed xyz
    float* norms;
    float* t_hit;  // optional: ray parameter of the hit (NaN on a miss)
};

// Far-step samples whose reads one ray issues together, per storage form.
template <class S> constexpr int kSpecOf = S::kBytes == 8 ? RAY_SPEC32 : RAY_SPEC16;

template <class S, bool E, int FD>
__global__ __launch_bounds__(256) void raycastKernel(Vol<S> V, VolParams p, PoseArg pa, Intr K, int W, int H,
                                                     RayOut out) {
    // XCD-aware tile order: XCD k (blocks b with b % 8 == k) gets a contiguous band.
    const int ntx = (W + 15) / 16, nty = (H + 15) / 16, nt = ntx * nty;
    const int b = blockIdx.x, xcd = b & 7, k = b >> 3;
    const int per = nt >> 3, rem = nt & 7;
    if (k >= per + (xcd < rem ? 1 : 0)) return;
    const int tile = xcd * per + min(xcd, rem) + k;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x = (tile % ntx) * 16 + (wv & 1) * 8 + (lane & 7);
    const int y = (tile / ntx) * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (x >= W || y >= H) return;
    const int o = y * W + x;

    const Rigid pose = rigidFrom(pa);
    const float vs = p.vs, trunc = p.trunc;
    const F3 cam = pose.t;
    const float span = (float)(p.res - 1) * vs;
    const float lo_a[3] = {p.origin[0], p.origin[1], p.origin[2]};
    const float hi_a[3] = {p.origin[0] + span, p.origin[1] + span, p.origin[2] + span};
    const float min_step = 0.5f * vs;
    const float far_step = smax(min_step, 0.8f * trunc);

    const F3 ray_cam{((float)x - K.cx) / K.fx, ((float)y - K.cy) / K.fy, 1.0f};
    const float ray_norm = f3norm(ray_cam);
    const F3 dir = f3div(mulv(pose.R, ray_cam), ray_norm);
    float t_near = p.min_depth * ray_norm;
    float t_far = p.max_depth * ray_norm;
    const float cam_a[3] = {cam.x, cam.y, cam.z}, dir_a[3] = {dir.x, dir.y, dir.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / dir_a[a];
        float t0 = (lo_a[a] - cam_a[a]) * inv;
        float t1 = (hi_a[a] - cam_a[a]) * inv;
        if (t0 > t1) {
            const float s = t0;
            t0 = t1;
            t1 = s;
        }
        if (__builtin_isnan(t0) || __builtin_isnan(t1)) {
            if (cam_a[a] < lo_a[a] || cam_a[a] > hi_a[a]) t_far = -1.0f;
            continue;
        }
        t_near = smax(t_near, t0);
        t_far = smin(t_far, t1);
    }

    bool found = false;
    float t_hit = 0.0f;
    F3 hit{0, 0, 0}, nrm{0, 0, 0};
    if (t_near < t_far) {
        float t_prev = t_near, f_prev = 1.0f;
        bool prev_valid = sampleVol<S, E, FD>(V, f3add(cam, f3mul(dir, t_prev)), f_prev);
        bool go = !(prev_valid && (!isfinite(f_prev) || f_prev < 0.0f));
        while (go && t_prev < t_far) {
            if (kSpecOf<S> > 0 && (!prev_valid || f_prev >= 0.999f)) {
                // A run of far steps: while every sample is unobserved or >= 0.999 the
                // step stays far_step, so the next kSpec parameters follow from t_prev
                // alone (the same recurrence as one step at a time). Their reads are
                // issued together, then the samples are taken in order; the run ends at
                // the first sample that changes the step, and the rest are discarded.
                constexpr int kSpec = kSpecOf<S> > 0 ? kSpecOf<S> : 1;
                float ts[kSpec];
                Cell<S> cs[kSpec];
                float tt = t_prev;
#pragma unroll
                for (int k = 0; k < kSpec; ++k) {
                    tt = smin(tt + far_step, t_far);
                    ts[k] = tt;
                }
#pragma unroll
                for (int k = 0; k < kSpec; ++k) cs[k] = sampleLoad<S, E, FD>(V, f3add(cam, f3mul(dir, ts[k])));
#pragma unroll
                for (int k = 0; k < kSpec; ++k) {
                    const float t = ts[k];
                    float f = 1.0f;
                    const bool valid = sampleEval<S, E>(cs[k], f);
                    if (valid && !isfinite(f)) {
                        go = false;
                        break;
                    }
                    if (valid && prev_valid) {
                        if (f_prev > 0.0f && f <= 0.0f) {
                            t_hit = t_prev + (t - t_prev) * (f_prev / (f_prev - f));
                            found = true;
                            go = false;
                            break;
                        }
                        if (f_prev < 0.0f && f >= 0.0f) {
                            go = false;
                            break;
                        }
                    } else if (valid && f < 0.0f) {
                        go = false;
                        break;
                    }
                    t_prev = t;
                    f_prev = f;
                    prev_valid = valid;
                    if (t >= t_far) {
                        go = false;
                        break;
                    }
                    if (valid && f < 0.999f) break;
                }
                continue;
            }
            const float step =
                (!prev_valid || f_prev >= 0.999f) ? far_step : smax(min_step, 0.8f * f_prev * trunc);
            const float t = smin(t_prev + step, t_far);
            float f = 1.0f;
            const bool valid = sampleVol<S, E, FD>(V, f3add(cam, f3mul(dir, t)), f);
            if (valid && !isfinite(f)) break;
            if (valid && prev_valid) {
                if (f_prev > 0.0f && f <= 0.0f) {
                    t_hit = t_prev + (t - t_prev) * (f_prev / (f_prev - f));
                    found = true;
                    break;
                }
                if (f_prev < 0.0f && f >= 0.0f) break;
            } else if (valid && f < 0.0f) {
                break;
            }
            t_prev = t;
            f_prev = f;
            prev_valid = valid;
            if (t >= t_far) break;
        }
        if (found) {
            hit = f3add(cam, f3mul(dir, t_hit));
            found = normalVol<S, E, FD>(V, hit, nrm) && f3finite(hit);
        }
    }
    if (!found) {
        hit = F3{0, 0, 0};
        nrm = F3{0, 0, 0};
    }
    float* vo = out.verts + 3 * (size_t)o;
    float* no = out.norms + 3 * (size_t)o;
    vo[0] = hit.x;
    vo[1] = hit.y;
    vo[2] = hit.z;
    no[0] = nrm.x;
    no[1] = nrm.y;
    no[2] = nrm.z;
    if (out.t_hit) out.t_hit[o] = found ? t_hit : __builtin_nanf("");
}

template <class S>
Vol<S> makeVol(const void* vox, const VolParams& p, const FixedDiv& dvs) {
    Vol<S> V;
    V.vox = static_cast<const typename S::Word*>(vox);
    V.res = p.res;
    V.vs = p.vs;
    V.inv_vs = 1.0f / p.vs;
    V.ox = p.origin[0];
    V.oy = p.origin[1];
    V.oz = p.origin[2];
    V.dvs = dvs;
    return V;
}

template <class S, bool E, int FD>
void launchRay(const void* vox, const VolParams& p, const FixedDiv& dvs, const PoseArg& pa, Intr K, int W, int H,
               const RayOut& out, hipStream_t st) {
    const int nt = ((W + 15) / 16) * ((H + 15) / 16);
    const unsigned blocks = (unsigned)(((nt + 7) / 8) * 8);
    hipLaunchKernelGGL((raycastKernel<S, E, FD>), dim3(blocks), dim3(256), 0, st, makeVol<S>(vox, p, dvs), p, pa, K,
                       W, H, out);
}

// dvs: division by the voxel size, verified exact (dvs.fast) or IEEE. The divided
// values are sample coordinates minus the volume origin; when no origin component is
// below 2^-60 in magnitude they cannot land in (0, 2^-100) and the unguarded form runs.
hipError_t launchRaycast(int format, bool exact, const void* vox, const VolParams& p, const FixedDiv& dvs,
                         const PoseArg& pa, Intr K, int W, int H, float* verts, float* norms, float* t_hit,
                         hipStream_t st) {
    const RayOut out{verts, norms, t_hit};
    const bool far_origin = fabsf(p.origin[0]) >= 0x1p-60f && fabsf(p.origin[1]) >= 0x1p-60f &&
                            fabsf(p.origin[2]) >= 0x1p-60f;
    const int fd = dvs.fast == 0 ? 0 : (far_origin ? 2 : 1);
    if (format == kFmtF32) {
        if (!exact) launchRay<StoreF32, false, 0>(vox, p, dvs, pa, K, W, H, out, st);
        else if (fd == 2) launchRay<StoreF32, true, 2>(vox, p, dvs, pa, K, W, H, out, st);
        else if (fd == 1) launchRay<StoreF32, true, 1>(vox, p, dvs, pa, K, W, H, out, st);
        else launchRay<StoreF32, true, 0>(vox, p, dvs, pa, K, W, H, out, st);
    } else {
        // The FP16 march is instruction-bound: the guarded form (a branch per division)
        // measures slower than the IEEE division there, so it takes only the unguarded one.
        if (!exact) launchRay<StoreF16, false, 0>(vox, p, dvs, pa, K, W, H, out, st);
        else if (fd == 2) launchRay<StoreF16, true, 2>(vox, p, dvs, pa, K, W, H, out, st);
        else launchRay<StoreF16, true, 0>(vox, p, dvs, pa, K, W, H, out, st);
    }
    return hipGetLastError();
}
}  // namespace azu
