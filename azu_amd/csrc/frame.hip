// Live frame geometry on gfx950: the reference's buildFrameData / updateVerticesFromDepth
// / computeNormals / buildFramePyramid (src/sensor/FrameData.cpp:113-345).
//
//  * liveLevel0: per pixel, millimetres, metres or raw Kinect codes -> band-gated metres, back-projected
//    vertex, and the 4-neighbour cross-product normal with the depth-jump guard. The
//    neighbours' vertices are recomputed from their depths (same formula, same bits),
//    so one kernel produces the whole level.
//  * livePyramid: per level-1 pixel, the 2x2 level-0 block beneath (2x2 mean of valid
//    samples, or the first nearest sample when the block straddles a depth jump); the
//    four level-1 pixels of a 2x2 block share a lane quad, which then forms the level-2
//    pixel from them. Levels 1 and 2 keep the averaged, renormalised normals like the CPU path.
// Vectors are packed xyz float triples; invalid entries are zero.
#include "common.hpp"

namespace azu {

__device__ __forceinline__ float depthJump(float d) { return smax(0.03f, d * 0.05f); }

// Source encodings of the live depth: float metres, uint16 millimetres, or uint16 raw
// Kinect v1 11-bit codes (DepthValidity.h:47-93: raw 0 and raw >= 2047 are invalid,
// metres = 1 / (raw * -0.0030711016 + 3.3309495161), which has its pole at raw ~1084.6
// and is negative beyond it; the band gate rejects both).
enum LiveSrc { kSrcMetres = 0, kSrcMm = 1, kSrcRaw = 2 };

template <int SRC>
__device__ __forceinline__ float liveDepth(const void* src, int i, float min_d, float max_d) {
    float m;
    bool nz;
    if (SRC == kSrcRaw) {
        const uint16_t v = static_cast<const uint16_t*>(src)[i];
        nz = v != 0 && v < 2047;
        m = nz ? 1.0f / ((float)v * -0.0030711016f + 3.3309495161f) : 0.0f;
    } else if (SRC == kSrcMm) {
        const uint16_t v = static_cast<const uint16_t*>(src)[i];
        m = (float)v * 0.001f;
        nz = v != 0;
    } else {
        m = static_cast<const float*>(src)[i];
        nz = m != 0.0f;
    }
    return (nz && isfinite(m) && m >= min_d && m <= max_d) ? m : 0.0f;
}

__device__ __forceinline__ F3 backProject(int x, int y, float d, float fx_inv, float fy_inv, Intr K) {
    if (!(d > 0.0f)) return F3{0, 0, 0};
    return F3{((float)x - K.cx) * fx_inv * d, ((float)y - K.cy) * fy_inv * d, d};
}

__device__ __forceinline__ void st3(float* p, int i, F3 v) {
    p[3 * i + 0] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ F3 ld3(const float* p, int i) { return F3{p[3 * i + 0], p[3 * i + 1], p[3 * i + 2]}; }

template <int SRC>
__global__ __launch_bounds__(256) void liveLevel0(const void* __restrict__ src, int W, int H, Intr K, float min_d,
                                                  float max_d, float* __restrict__ d0, float* __restrict__ v0,
                                                  float* __restrict__ n0) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const int c = y * W + x;
    const float fx_inv = 1.0f / K.fx, fy_inv = 1.0f / K.fy;
    const float dc = liveDepth<SRC>(src, c, min_d, max_d);
    d0[c] = dc;
    st3(v0, c, backProject(x, y, dc, fx_inv, fy_inv, K));
    F3 n{0, 0, 0};
    if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
        const float dr = liveDepth<SRC>(src, c + 1, min_d, max_d);
        const float dl = liveDepth<SRC>(src, c - 1, min_d, max_d);
        const float du = liveDepth<SRC>(src, c - W, min_d, max_d);
        const float dd = liveDepth<SRC>(src, c + W, min_d, max_d);
        if (dc > 0.0f && dr > 0.0f && dl > 0.0f && du > 0.0f && dd > 0.0f) {
            const float jump = depthJump(dc);
            if (!(fabsf(dc - dr) > jump || fabsf(dc - dl) > jump || fabsf(dc - du) > jump || fabsf(dc - dd) > jump)) {
                const F3 vr = backProject(x + 1, y, dr, fx_inv, fy_inv, K);
                const F3 vl = backProject(x - 1, y, dl, fx_inv, fy_inv, K);
                const F3 vu = backProject(x, y - 1, du, fx_inv, fy_inv, K);
                const F3 vd = backProject(x, y + 1, dd, fx_inv, fy_inv, K);
                const F3 nn = f3cross(f3sub(vr, vl), f3sub(vd, vu));
                const float len = f3norm(nn);
                if (len > 1e-6f) n = f3div(nn, len);
            }
        }
    }
    st3(n0, c, n);
}

struct Px {
    float d;
    F3 v, n;
};

// downsample (FrameData.cpp:262-334) of one 2x2 block whose samples are given in
// (dy, dx) scan order.
__device__ __forceinline__ Px down2x2(const Px s[4]) {
    F3 sum_v{0, 0, 0}, sum_n{0, 0, 0};
    float sum_d = 0.0f, d_min = 0.0f, d_max = 0.0f;
    int count = 0, imin = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float d = s[i].d;
        if (d > 0.0f) {
            sum_v = f3add(sum_v, s[i].v);
            sum_n = f3add(sum_n, s[i].n);
            sum_d = sum_d + d;
            if (count == 0 || d < d_min) {
                d_min = d;
                imin = i;
            }
            if (count == 0 || d > d_max) d_max = d;
            ++count;
        }
    }
    Px r{0.0f, {0, 0, 0}, {0, 0, 0}};
    if (count == 0) return r;
    if (d_max - d_min > depthJump(d_min)) {
        r.v = s[imin].v;
        r.d = s[imin].d;
        const float nl = f3norm(s[imin].n);
        r.n = (nl > 1e-6f) ? f3div(s[imin].n, nl) : F3{0, 0, 0};
    } else {
        const float inv = 1.0f / (float)count;
        r.v = f3mul(sum_v, inv);
        r.d = sum_d * inv;
        const float nl = f3norm(sum_n);
        r.n = (nl > 1e-6f) ? f3div(sum_n, nl) : F3{0, 0, 0};
    }
    return r;
}

// One thread per level-1 pixel: a block covers 16 x 4 of the 2x2 level-1 blocks, and the
// four pixels of a block sit in one lane quad (lane 4 g + 2 j + i for pixel (i, j) of
// block g), so the quad's first lane forms the level-2 pixel from the other three's
// results without a round trip through memory.
__device__ __forceinline__ Px shflPx(const Px& a, int src) {
    Px r;
    r.d = __shfl(a.d, src);
    r.v = F3{__shfl(a.v.x, src), __shfl(a.v.y, src), __shfl(a.v.z, src)};
    r.n = F3{__shfl(a.n.x, src), __shfl(a.n.y, src), __shfl(a.n.z, src)};
    return r;
}

__global__ __launch_bounds__(256) void livePyramid(int W0, int H0, const float* __restrict__ d0,
                                                   const float* __restrict__ v0, const float* __restrict__ n0,
                                                   float* __restrict__ d1, float* __restrict__ v1,
                                                   float* __restrict__ n1, float* __restrict__ d2,
                                                   float* __restrict__ v2, float* __restrict__ n2) {
    const int W1 = W0 / 2, H1 = H0 / 2, W2 = W1 / 2, H2 = H1 / 2;
    const int q = threadIdx.x & 3, g = threadIdx.x >> 2;
    const int gx = blockIdx.x * 16 + (g & 15), gy = blockIdx.y * 4 + (g >> 4);  // 2x2 block of level 1
    const int x1 = 2 * gx + (q & 1), y1 = 2 * gy + (q >> 1);
    Px r{0.0f, {0, 0, 0}, {0, 0, 0}};
    if (x1 < W1 && y1 < H1) {
        Px s[4];
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int si = (2 * y1 + dy) * W0 + (2 * x1 + dx);
                s[dy * 2 + dx] = Px{d0[si], ld3(v0, si), ld3(n0, si)};
            }
        r = down2x2(s);
        const int o = y1 * W1 + x1;
        d1[o] = r.d;
        st3(v1, o, r.v);
        st3(n1, o, r.n);
    }
    // Every lane takes part in the exchange (pixels outside level 1 contribute zeros, as
    // the level-2 pixel of a block reaching outside is not formed).
    const int base = (int)(threadIdx.x & 63) & ~3;
    Px l1[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) l1[k] = shflPx(r, base + k);
    if (q == 0 && gx < W2 && gy < H2) {
        const Px r2 = down2x2(l1);
        const int o = gy * W2 + gx;
        d2[o] = r2.d;
        st3(v2, o, r2.v);
        st3(n2, o, r2.n);
    }
}

hipError_t launchLive(int src_kind, const void* src, int W, int H, Intr K, float min_d, float max_d, float* const d[3],
                      float* const v[3], float* const n[3], hipStream_t st) {
    const dim3 g0((W + 15) / 16, (H + 15) / 16);
    if (src_kind == kSrcMm)
        hipLaunchKernelGGL(liveLevel0<kSrcMm>, g0, dim3(256), 0, st, src, W, H, K, min_d, max_d, d[0], v[0], n[0]);
    else if (src_kind == kSrcRaw)
        hipLaunchKernelGGL(liveLevel0<kSrcRaw>, g0, dim3(256), 0, st, src, W, H, K, min_d, max_d, d[0], v[0], n[0]);
    else
        hipLaunchKernelGGL(liveLevel0<kSrcMetres>, g0, dim3(256), 0, st, src, W, H, K, min_d, max_d, d[0], v[0],
                           n[0]);
    const int W1 = W / 2, H1 = H / 2;
    const dim3 g1(((W1 + 1) / 2 + 15) / 16, ((H1 + 1) / 2 + 3) / 4);
    hipLaunchKernelGGL(livePyramid, g1, dim3(256), 0, st, W, H, d[0], v[0], n[0], d[1], v[1], n[1], d[2], v[2],
                       n[2]);
    return hipGetLastError();
}

}  // namespace azu

namespace azu {

hipError_t launchDownsample(int W, int H, float* const d[3], float* const v[3], float* const n[3], hipStream_t st) {
    const int W1 = W / 2, H1 = H / 2;
    const dim3 g1(((W1 + 1) / 2 + 15) / 16, ((H1 + 1) / 2 + 3) / 4);
    hipLaunchKernelGGL(livePyramid, g1, dim3(256), 0, st, W, H, d[0], v[0], n[0], d[1], v[1], n[1], d[2], v[2],
                       n[2]);
    return hipGetLastError();
}

}  // namespace azu
