// Raycast on gfx950 (the reference's TSDFVolume::raycast / sampleTSDF / computeNormal,
// src/tsdf/TSDFVolume.cpp:348-525): per pixel, slab-clip the ray to the volume box,
// march with the adaptive step (far step while unobserved or f >= 0.999, else
// max(vs/2, 0.8 f tau)), stop at the first observed + -> - crossing, refine it with one
// secant step and take the normal from six trilinear samples at +-vs.
//
// Layout: a 256-thread block owns a 16x16 pixel tile, each wave an 8x8 sub-tile, so
// the rays of a wave march through neighbouring voxels. For FP16 words the blocks are
// remapped so that each of the 8 XCDs (block b runs on XCD b % 8) handles one
// contiguous band of tiles; FP32 tiles run in launch order. Each trilinear sample reads
// its 8 corners as 4 x-adjacent pairs (16 B per pair for FP32, 8 B for FP16/U16).
//
// Exact = true (FP32 parity path): IEEE division where the reference divides, no
// contraction, bit-identical to the CPU restatement. Exact = false (FP16 form):
// multiplications by 1/vs and fmaf interpolation.
#include <algorithm>

#include "common.hpp"

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
    __device__ static P loadB(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __builtin_bit_cast(P, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
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
    __device__ static P loadB(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __builtin_bit_cast(P, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    }
};

template <class S> struct Vol {
    const typename S::Word* vox;
    int res;
    float vs, inv_vs;
    float ox, oy, oz;
    FixedDiv dvs;  // division by vs
};

// FP32 volumes of at most 2^31 bytes are read through a buffer resource over the whole
// volume (wave-uniform), a 32-bit byte offset per sample and uniform row / plane
// offsets for the cell's four corner pairs; an offset of 2^31 (a cell outside the
// volume) reads 0 with no memory access.
struct VolBuf {
    __amdgpu_buffer_rsrc_t rs;
    unsigned sy, sz;  // bytes between rows / planes
};

// Empty-space skipping: per pixel, a lower bound on the distance from the camera to
// every marked brick whose span box the pixel's ray can meet (brickSplat), and a flag
// saying whether the bounds hold for this call. With the bound, every sample at
// t < dist / |dir| lies in no marked brick: each of its corners lies in a clear brick's
// span (or the cell is outside the volume), so the sample is far, and the march takes
// its far step there without reading it.
struct RaySkip {
    // Per 2 x 2 pixel cell, two float bit patterns: the least distance from the camera of
    // any marked brick the cell's rays can meet (+inf between calls) and the greatest (0
    // between calls); or null.
    unsigned* dsafe;
    const int* ok;
};

// sampleTSDF (TSDFVolume.cpp:492-525) in two halves: sampleLoad locates the cell and
// reads its 8 corners (nothing is read outside the volume), sampleEval applies the
// weight gate and interpolates.
template <class S> struct Cell {
    typename PairIO<S>::P q0, q1, q2, q3;
    float tx, ty, tz;
    bool inside;
};

template <class S, bool E, int FD, bool B>
__device__ __forceinline__ Cell<S> sampleLoad(const Vol<S>& V, const VolBuf& R, F3 p) {
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
    using IO = PairIO<S>;
    if (B) {
        // 24-bit multiplies: the volume spans < 2^31 bytes, so res < 2^11.
        const unsigned off =
            c.inside ? ((unsigned)__umul24((unsigned)__umul24((unsigned)z0, (unsigned)V.res) + (unsigned)y0,
                                           (unsigned)V.res) +
                        (unsigned)x0) *
                           (unsigned)S::kBytes
                     : 0x80000000u;
        c.q0 = IO::loadB(R.rs, off, 0u);
        c.q1 = IO::loadB(R.rs, off, R.sy);
        c.q2 = IO::loadB(R.rs, off, R.sz);
        c.q3 = IO::loadB(R.rs, off, R.sy + R.sz);
        return c;
    }
    // The in-plane index by a 24-bit multiply (y0, x0 < res < 2^16), the plane offset as one
    // 32 x 32 -> 64-bit product (res^2 < 2^32).
    const unsigned plane = (unsigned)V.res * (unsigned)V.res;
    const size_t sy = (size_t)V.res, sz = plane;
    const unsigned in_plane = (unsigned)__umul24((unsigned)y0, (unsigned)V.res) + (unsigned)x0;
    const size_t b = (size_t)(unsigned)z0 * plane + (size_t)in_plane;
    c.q0 = IO::load(V.vox + b);
    c.q1 = IO::load(V.vox + b + sy);
    c.q2 = IO::load(V.vox + b + sz);
    c.q3 = IO::load(V.vox + b + sy + sz);
    return c;
}

template <class S, bool E>
__device__ __forceinline__ bool sampleEvalValue(const Cell<S>& c, float& f);

template <class S, bool E>
__device__ __forceinline__ bool sampleEval(const Cell<S>& c, float& f) {
    if (!c.inside) return false;
    if (!S::observed8(c.q0.a, c.q0.b, c.q1.a, c.q1.b, c.q2.a, c.q2.b, c.q3.a, c.q3.b)) return false;
    return sampleEvalValue<S, E>(c, f);
}

// The trilinear value of a cell (no gates).
template <class S, bool E>
__device__ __forceinline__ bool sampleEvalValue(const Cell<S>& c, float& f) {
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

// sampleEval without branches: the value is computed from whatever was read and
// replaced by 1 where the sample is invalid (outside, or a corner unobserved).
template <class S, bool E>
__device__ __forceinline__ bool sampleEvalFlat(const Cell<S>& c, float& f) {
    // (bitwise: both operands are evaluated, no branch)
    const bool valid =
        c.inside & S::observed8(c.q0.a, c.q0.b, c.q1.a, c.q1.b, c.q2.a, c.q2.b, c.q3.a, c.q3.b);
    Cell<S> d = c;
    d.inside = true;
    float v = 1.0f;
    (void)sampleEvalValue<S, E>(d, v);
    f = valid ? v : 1.0f;
    return valid;
}

template <class S, bool E, int FD, bool B>
__device__ __forceinline__ bool sampleVol(const Vol<S>& V, const VolBuf& R, F3 p, float& f) {
    if (B) return sampleEval<S, E>(sampleLoad<S, E, FD, true>(V, R, p), f);
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

// computeNormal (TSDFVolume.cpp:478-490). The six samples' reads are all issued before
// the first is evaluated (three at a time for the 16-byte FP32 pairs), and the normal is
// withdrawn when any of them is invalid: the reference stops at the first invalid one, and
// either way there is no normal.
template <class S, bool E, int FD, bool B>
__device__ __forceinline__ bool normalVol(const Vol<S>& V, const VolBuf& R, F3 p, F3& n) {
    const float h = V.vs;
    const F3 q[6] = {F3{p.x + h, p.y + 0.0f, p.z + 0.0f}, F3{p.x - h, p.y - 0.0f, p.z - 0.0f},
                     F3{p.x + 0.0f, p.y + h, p.z + 0.0f}, F3{p.x - 0.0f, p.y - h, p.z - 0.0f},
                     F3{p.x + 0.0f, p.y + 0.0f, p.z + h}, F3{p.x - 0.0f, p.y - 0.0f, p.z - h}};
    constexpr int kBatch = S::kBytes == 8 ? 3 : 6;
    float s[6];
    bool ok = true;
#pragma unroll
    for (int b0 = 0; b0 < 6; b0 += kBatch) {
        Cell<S> c[kBatch];
#pragma unroll
        for (int i = 0; i < kBatch; ++i) c[i] = sampleLoad<S, E, FD, B>(V, R, q[b0 + i]);
#pragma unroll
        for (int i = 0; i < kBatch; ++i) ok = sampleEval<S, E>(c[i], s[b0 + i]) && ok;
    }
    if (!ok) return false;
    const F3 g{s[0] - s[1], s[2] - s[3], s[4] - s[5]};
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


template <class S, bool E, int FD, bool B>
__global__ __launch_bounds__(256) void raycastKernel(Vol<S> V, VolParams p, PoseArg pa, Intr K, int W, int H,
                                                     RayOut out, RaySkip skip) {
    VolBuf R;
    if (B) {
        R.rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<typename S::Word*>(V.vox), (short)0,
                                                 (int)((unsigned)V.res * (unsigned)V.res * (unsigned)V.res *
                                                       (unsigned)S::kBytes),
                                                 0x00020000);
        R.sy = (unsigned)V.res * (unsigned)S::kBytes;
        R.sz = R.sy * (unsigned)V.res;
    }
    // Tile order. FP32 words: launch order, so that consecutive tiles run on different
    // XCDs (blocks are dealt to the XCDs in turn): the long rays cluster in a few image
    // regions, and a contiguous band per XCD leaves their memory traffic to one XCD's L2
    // (measured 338 -> 297 us per launch at 512^3). FP16 words: the band per XCD (block b
    // on XCD b % 8), which measures a few microseconds faster there.
    const int ntx = (W + 15) / 16, nty = (H + 15) / 16, nt = ntx * nty;
    int tile = blockIdx.x;
    if (S::kBytes == 8) {
        if (tile >= nt) return;
    } else {
        const int b = blockIdx.x, xcd = b & 7, k = b >> 3;
        const int per = nt >> 3, rem = nt & 7;
        if (k >= per + (xcd < rem ? 1 : 0)) return;
        tile = xcd * per + min(xcd, rem) + k;
    }
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

    // A step of min_step must move t everywhere on [t_near, t_far]; where it cannot (a
    // camera so far away that the float spacing of t exceeds twice the step) the march
    // would never end, and the ray is left a miss.
    if (!(fabsf(t_far) + min_step > fabsf(t_far)) || !(fabsf(t_near) + min_step > fabsf(t_near))) t_far = -1.0f;
    // The skip bounds of this pixel's 2 x 2 cell (the cell lies in this wave's 8 x 8 tile;
    // its first pixel resets them at the end, after every lane has read them). Samples
    // before t_safe and after t_end are far.
    const int cell = (y >> 1) * ((W + 1) >> 1) + (x >> 1);
    float t_safe = -__builtin_inff(), t_end = __builtin_inff();
    if (skip.dsafe) {
        const uint2 d = *reinterpret_cast<const uint2*>(skip.dsafe + 2 * (size_t)cell);
        if (*skip.ok) {
            const float dn = f3norm(dir);
            t_safe = __uint_as_float(d.x) / dn;
            t_end = __uint_as_float(d.y) / dn;
        }
    }
    bool found = false;
    float t_hit = 0.0f;
    F3 hit{0, 0, 0}, nrm{0, 0, 0};
    if (t_near < t_far) {
        float t_prev = t_near, f_prev = 1.0f;
        // Every sample lies at t >= t_near; past t_end all are far, so none can hit.
        bool prev_valid = false, go = t_near < t_end;
        if (go && t_near < t_safe) {
            // Samples before t_safe are far: advance by the far-step recurrence without
            // taking them, then take the last one skipped (its value enters the next
            // step's crossing test).
            for (;;) {
                const float t = smin(t_prev + far_step, t_far);
                if (!(t < t_safe)) break;
                t_prev = t;
                if (t_prev >= t_far) {
                    go = false;
                    break;
                }
            }
        }
        if (go) {
            prev_valid = sampleVol<S, E, FD, B>(V, R, f3add(cam, f3mul(dir, t_prev)), f_prev);
            go = !(prev_valid && (!isfinite(f_prev) || f_prev < 0.0f));
        }
        // Once t_prev >= t_end every later sample is far: no hit can follow. Each step
        // carries its outcomes as flags and the loop has one exit test; the hit step's
        // bracket (t_a, f_a) -> (t_prev, f_prev) is refined after the loop.
        bool more = go && t_prev < t_far && t_prev < t_end;
        float t_a = 0.0f, f_a = 0.0f;
        while (more) {
            const float step =
                (!prev_valid || f_prev >= 0.999f) ? far_step : smax(min_step, 0.8f * f_prev * trunc);
            const float t = smin(t_prev + step, t_far);
            // A non-finite value stops the march; with both samples valid, + -> - is the
            // hit and - -> + stops; a valid negative first sample after an invalid one
            // stops. A step that does not stop leaves t < t_far.
            float f;
            const bool valid = sampleEvalFlat<S, E>(sampleLoad<S, E, FD, B>(V, R, f3add(cam, f3mul(dir, t))), f);
            // (bitwise: every flag is evaluated, no branch)
            const bool fin = isfinite(f), both = valid & prev_valid;
            const bool hit = both & fin & (f_prev > 0.0f) & (f <= 0.0f);
            const bool stop = hit | (valid & !fin) | (both & (f_prev < 0.0f) & (f >= 0.0f)) |
                              (valid & !prev_valid & (f < 0.0f)) | (t >= t_far);
            t_a = t_prev;
            f_a = f_prev;
            t_prev = t;
            f_prev = f;
            prev_valid = valid;
            found = hit;
            more = !stop & (t < t_end);
        }
        if (found) t_hit = t_a + (t_prev - t_a) * (f_a / (f_a - f_prev));
        if (found) {
            hit = f3add(cam, f3mul(dir, t_hit));
            found = normalVol<S, E, FD, B>(V, R, hit, nrm) && f3finite(hit);
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
    if (skip.dsafe && ((x | y) & 1) == 0) *reinterpret_cast<uint2*>(skip.dsafe + 2 * (size_t)cell) = make_uint2(0x7f800000u, 0u);
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
void launchRay(const void* vox, const VolParams& p, const FixedDiv& dvs, const RaySkip& skip, const PoseArg& pa,
               Intr K, int W, int H, const RayOut& out, hipStream_t st) {
    const int nt = ((W + 15) / 16) * ((H + 15) / 16);
    const unsigned blocks = (unsigned)(S::kBytes == 8 ? nt : ((nt + 7) / 8) * 8);
    // The buffer form measures faster for the FP32 words (344 against 412 us per launch
    // at 512^3) and slightly slower for the FP16 ones (224 against 218 us).
    const bool small = S::kBytes == 8 && (size_t)p.res * p.res * p.res * S::kBytes <= ((size_t)1 << 31);
    if (small)
        hipLaunchKernelGGL((raycastKernel<S, E, FD, true>), dim3(blocks), dim3(256), 0, st, makeVol<S>(vox, p, dvs),
                           p, pa, K, W, H, out, skip);
    else
        hipLaunchKernelGGL((raycastKernel<S, E, FD, false>), dim3(blocks), dim3(256), 0, st,
                           makeVol<S>(vox, p, dvs), p, pa, K, W, H, out, skip);
}

constexpr int kSplatPer = 16;  // bricks per wave in brickSplat

// One marked brick's share of brickSplat: its distance bounds (float bit patterns) and the
// rectangle of 2 x 2 pixel cells (first cell, width, cell count; area 0: none) its widened
// span box can reach.
__device__ void splatRect(const Rigid& P, const M33& Ri, const VolParams& p, Intr K, int W, int H, int nb, size_t bi,
                          int& cx0_out, int& cy0_out, int& w_out, int& area_out, unsigned& dlow_out,
                          unsigned& dhigh_out) {
    const int b[3] = {(int)(bi % nb), (int)((bi / nb) % nb), (int)(bi / ((size_t)nb * nb))};
    const F3 cam = P.t;
    const float cam_a[3] = {cam.x, cam.y, cam.z};
    const float M = (fabsf(cam.x) + fabsf(cam.y) + fabsf(cam.z) +
                     2.0f * (fabsf(p.origin[0]) + fabsf(p.origin[1]) + fabsf(p.origin[2]))) / p.vs +
                    12.0f * (float)p.res + 32.0f;
    const float e = 0.01f + M * 0x1p-17f;
    float lo[3], hi[3], d2 = 0.0f, g2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = p.origin[a] + p.vs * ((float)(kBrick * b[a]) - e);
        hi[a] = p.origin[a] + p.vs * ((float)(kBrick * b[a] + kBrick) + e);
        const float d = fmaxf(fmaxf(lo[a] - cam_a[a], cam_a[a] - hi[a]), 0.0f);
        const float g = fmaxf(cam_a[a] - lo[a], hi[a] - cam_a[a]);
        d2 += d * d;
        g2 += g * g;
    }
    const unsigned dlow = __float_as_uint(sqrtf(d2) * (1.0f - 1e-5f));
    const unsigned dhigh = __float_as_uint(sqrtf(g2) * (1.0f + 1e-5f));
    F3 q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        q[k] = mulv(Ri, F3{((k & 1) ? hi[0] : lo[0]) - cam.x, ((k & 2) ? hi[1] : lo[1]) - cam.y,
                           ((k & 4) ? hi[2] : lo[2]) - cam.z});
    const float zc = 0.5f * p.min_depth;
    float umin = __builtin_inff(), umax = -__builtin_inff(), vmin = __builtin_inff(), vmax = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (q[k].z >= zc) {
            const float u = K.fx * q[k].x / q[k].z + K.cx, v = K.fy * q[k].y / q[k].z + K.cy;
            umin = fminf(umin, u);
            umax = fmaxf(umax, u);
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
        }
#pragma unroll
        for (int bit = 1; bit < 8; bit <<= 1) {
            if (k & bit) continue;
            const F3 qa = q[k], qb = q[k | bit];
            if ((qa.z >= zc) == (qb.z >= zc)) continue;
            const float s = (zc - qa.z) / (qb.z - qa.z);
            const float x = qa.x + (qb.x - qa.x) * s, y = qa.y + (qb.y - qa.y) * s;
            const float u = K.fx * x / zc + K.cx, v = K.fy * y / zc + K.cy;
            umin = fminf(umin, u);
            umax = fmaxf(umax, u);
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
        }
    }
    area_out = 0;
    if (!(umin <= umax)) return;  // nothing in front of the plane
    const int x0 = (int)fmaxf(floorf(fmaxf(umin, -4.0f)) - 1.0f, 0.0f);
    const int x1 = (int)fminf(ceilf(fminf(umax, (float)W + 4.0f)) + 1.0f, (float)(W - 1));
    const int y0 = (int)fmaxf(floorf(fmaxf(vmin, -4.0f)) - 1.0f, 0.0f);
    const int y1 = (int)fminf(ceilf(fminf(vmax, (float)H + 4.0f)) + 1.0f, (float)(H - 1));
    if (x0 > x1 || y0 > y1) return;
    cx0_out = x0 >> 1;
    cy0_out = y0 >> 1;
    w_out = (x1 >> 1) - cx0_out + 1;
    area_out = w_out * ((y1 >> 1) - cy0_out + 1);
    dlow_out = dlow;
    dhigh_out = dhigh;
}

// dvs: division by the voxel size, verified exact (dvs.fast) or IEEE. The divided
// values are sample coordinates minus the volume origin; when no origin component is
// below 2^-60 in magnitude they cannot land in (0, 2^-100) and the unguarded form runs.
// The per-pixel bounds of RaySkip (brickSplat, splatRect). Per marked brick: its span box,
// widened by e voxels (e covers the rounding of the march's grid coordinates, which stays
// below 4 2^-24 M voxels for a bound M on the magnitudes involved), is projected through
// the inverse camera rotation. Only its part in front of the plane z = min_depth / 2 can
// hold a sample (every sample lies at camera depth >= min_depth), so the corners there
// and the box edges' crossings of the plane are projected; the perspective image of that
// convex part is their hull, and the pixel rectangle around it, widened by one pixel,
// takes the minimum of the box's distance from the camera (less 1e-5 of it) and the
// maximum of its farthest point's distance (plus 1e-5 of it) in every 2 x 2 pixel cell
// it touches, one cell per lane. A sample at parameter t is the point cam + dir t, at
// distance t |dir|; a sample that is not far lies in some marked brick's widened box, in
// front of the plane, so its pixel is in that brick's rectangle and its distance between
// the two bounds.
// ok = 0 (take every sample) unless the pose is finite and invertible and min_depth > 0.
__global__ __launch_bounds__(256) void brickSplat(const unsigned char* __restrict__ bricks, int nb, VolParams p,
                                                  PoseArg pa, Intr K, int W, int H, unsigned* __restrict__ dsafe,
                                                  int* __restrict__ ok) {
    const Rigid P = rigidFrom(pa);
    const M33& R = P.R;
    M33 A;  // adjugate of R
    A.m[0][0] = R.m[1][1] * R.m[2][2] - R.m[1][2] * R.m[2][1];
    A.m[0][1] = R.m[0][2] * R.m[2][1] - R.m[0][1] * R.m[2][2];
    A.m[0][2] = R.m[0][1] * R.m[1][2] - R.m[0][2] * R.m[1][1];
    A.m[1][0] = R.m[1][2] * R.m[2][0] - R.m[1][0] * R.m[2][2];
    A.m[1][1] = R.m[0][0] * R.m[2][2] - R.m[0][2] * R.m[2][0];
    A.m[1][2] = R.m[0][2] * R.m[1][0] - R.m[0][0] * R.m[1][2];
    A.m[2][0] = R.m[1][0] * R.m[2][1] - R.m[1][1] * R.m[2][0];
    A.m[2][1] = R.m[0][1] * R.m[2][0] - R.m[0][0] * R.m[2][1];
    A.m[2][2] = R.m[0][0] * R.m[1][1] - R.m[0][1] * R.m[1][0];
    const float det = R.m[0][0] * A.m[0][0] + R.m[0][1] * A.m[1][0] + R.m[0][2] * A.m[2][0];
    bool valid = isfinite(det) && fabsf(det) > 1e-3f && p.min_depth > 0.0f && f3finite(P.t);
    M33 Ri;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Ri.m[r][c] = A.m[r][c] / det;
            valid = valid && isfinite(Ri.m[r][c]);
        }
    if (blockIdx.x == 0 && threadIdx.x == 0) *ok = valid ? 1 : 0;
    if (!valid) return;

    // Wave w takes bricks kSplatPer w ... kSplatPer (w + 1) - 1, one lane each for the test
    // and the projection; then all lanes splat each marked brick's cells in turn.
    const int lane = threadIdx.x & 63;
    const size_t bi = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kSplatPer + (size_t)lane;
    const bool marked = lane < kSplatPer && bi < (size_t)nb * nb * nb && bricks[bi] != 0;
    if (__ballot(marked) == 0ull) return;  // whole wave
    int cx0 = 0, cy0 = 0, w = 0, area = 0;
    unsigned dlow = 0u, dhigh = 0u;
    if (marked) splatRect(P, Ri, p, K, W, H, nb, bi, cx0, cy0, w, area, dlow, dhigh);
    const int cw = (W + 1) >> 1;
    for (unsigned long long m = __ballot(area > 0); m != 0ull; m &= m - 1ull) {
        const int l = __builtin_ctzll(m);
        const int bx0 = __builtin_amdgcn_readlane(cx0, l), by0 = __builtin_amdgcn_readlane(cy0, l);
        const int bw = __builtin_amdgcn_readlane(w, l), ba = __builtin_amdgcn_readlane(area, l);
        const unsigned bl = (unsigned)__builtin_amdgcn_readlane((int)dlow, l);
        const unsigned bh = (unsigned)__builtin_amdgcn_readlane((int)dhigh, l);
        for (int k = lane; k < ba; k += 64) {
            unsigned* c = dsafe + 2 * ((size_t)(by0 + k / bw) * cw + (bx0 + k % bw));
            atomicMin(c, bl);
            atomicMax(c + 1, bh);
        }
    }
}

__global__ void fillPairs(uint2* __restrict__ p, size_t n, uint2 v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// The skip bounds' scratch in its resting state: cells cells of (+inf, 0).
hipError_t launchSkipInit(unsigned* dsafe, size_t cells, hipStream_t st) {
    hipLaunchKernelGGL(fillPairs, dim3((unsigned)std::min<size_t>((cells + 255) / 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<uint2*>(dsafe), cells, make_uint2(0x7f800000u, 0u));
    return hipGetLastError();
}

// The per-pixel skip bounds of a pose into dsafe / ok (brickSplat over the marked bricks).
hipError_t launchBrickSplat(const unsigned char* bricks, const VolParams& p, const PoseArg& pa, Intr K, int W, int H,
                            unsigned* dsafe, int* ok, hipStream_t st) {
    const int nb = brickCount(p.res);
    const unsigned blocks = (unsigned)(((size_t)nb * nb * nb + 4 * kSplatPer - 1) / (4 * kSplatPer));
    hipLaunchKernelGGL(brickSplat, dim3(blocks), dim3(256), 0, st, bricks, nb, p, pa, K, W, H, dsafe, ok);
    return hipGetLastError();
}

// bricks: the volume's brick map, or null to take every sample; dsafe (two words per
// 2 x 2 pixel cell, ((W + 1) / 2) ((H + 1) / 2) cells, 8-byte aligned, in the resting
// state of launchSkipInit on entry and on return) and ok (one int): the skip bounds'
// scratch.
hipError_t launchRaycast(int format, bool exact, const void* vox, const VolParams& p, const FixedDiv& dvs,
                         const unsigned char* bricks, unsigned* dsafe, int* ok, const PoseArg& pa, Intr K, int W,
                         int H, float* verts, float* norms, float* t_hit, hipStream_t st) {
    const RayOut out{verts, norms, t_hit};
    RaySkip skip{nullptr, ok};
    if (bricks && dsafe && ok) {
        const hipError_t e = launchBrickSplat(bricks, p, pa, K, W, H, dsafe, ok, st);
        if (e != hipSuccess) return e;
        skip.dsafe = dsafe;
    }
    const bool far_origin = fabsf(p.origin[0]) >= 0x1p-60f && fabsf(p.origin[1]) >= 0x1p-60f &&
                            fabsf(p.origin[2]) >= 0x1p-60f;
    const int fd = dvs.fast == 0 ? 0 : (far_origin ? 2 : 1);
    if (format == kFmtF32) {
        if (!exact) launchRay<StoreF32, false, 0>(vox, p, dvs, skip, pa, K, W, H, out, st);
        else if (fd == 2) launchRay<StoreF32, true, 2>(vox, p, dvs, skip, pa, K, W, H, out, st);
        else if (fd == 1) launchRay<StoreF32, true, 1>(vox, p, dvs, skip, pa, K, W, H, out, st);
        else launchRay<StoreF32, true, 0>(vox, p, dvs, skip, pa, K, W, H, out, st);
    } else {
        // The FP16 march is instruction-bound: the guarded form (a branch per division)
        // measures slower than the IEEE division there, so it takes only the unguarded one.
        if (!exact) launchRay<StoreF16, false, 0>(vox, p, dvs, skip, pa, K, W, H, out, st);
        else if (fd == 2) launchRay<StoreF16, true, 2>(vox, p, dvs, skip, pa, K, W, H, out, st);
        else launchRay<StoreF16, true, 0>(vox, p, dvs, skip, pa, K, W, H, out, st);
    }
    return hipGetLastError();
}
}  // namespace azu
