// TSDF integration on gfx950: voxel-projective update over the camera frustum's
// bounding box (the reference's integrateCPU, src/tsdf/TSDFVolume.cpp:214-346).
//
// Two kernels:
//
//  * integrateExact<Store>: the FP32 parity path. The reference carries each row's
//    camera coordinate by an incremental x-step (c += R_wc.col(0) * vs,
//    TSDFVolume.cpp:286-287), so voxel x's coordinate is the result of (x - x0)
//    sequential float adds. To reproduce it bit for bit while keeping HBM access
//    coalesced, every wave owns 32 rows (consecutive z at one y): in phase 1 lanes r and
//    r + 32 walk row r through the two halves of a 32-column chunk, run the projection,
//    depth gather and the gates, and leave tsdf_new (or NaN = skip) in LDS; in phase 2
//    the wave reads the chunk transposed, lane = column, so each read-modify-write
//    covers 32 consecutive voxels of two rows. 32-row tiles keep a block's LDS at 17 KB,
//    so that up to 8 waves a SIMD fit (64-row tiles allowed 4 and measured 390 against
//    260 us when held to 2). Division and square root are IEEE-exact, nothing contracts.
//
//    integrateExact runs for both storage forms by default.
//
//  * integrateFast<Store>: the opt-in FAST mode, for either storage form. The same tiles
//    and tile classes, but lane = x column: the camera coordinate is an affine function
//    of (x, z) evaluated with fmaf and the reciprocals are the hardware v_rcp_f32, so
//    the lanes are independent voxels and each row of a tile is one coalesced access.
//    Not bit-identical to the reference; its accuracy is stated against the FP64 oracle.
//
// Both read the pose by value or from device memory (the ICP output), so a frame can
// run with no host synchronisation. Updated / newly observed voxel counts are
// accumulated per wave (ballot popcount) into a 2-entry device counter when given.
#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>

#include "common.hpp"

namespace azu {

struct IntFrame {
    M33 R_wc;
    F3 t_wc;
    M33 R_cw;
    F3 t_cw;
    F3 step;
    int x0, x1, y0, y1, z0, z1;
    bool ok;
};

// Frustum AABB and camera transform exactly as the CPU path computes them
// (TSDFVolume.cpp:238-266).
__device__ __forceinline__ IntFrame integrateSetup(const VolParams& p, const PoseArg& pa, Intr K, int W, int H,
                                                   float max_d) {
    IntFrame f;
    const Rigid pose = rigidFrom(pa);
    const Rigid w2c = rigidInverse(pose);
    f.R_wc = w2c.R;
    f.t_wc = w2c.t;
    f.R_cw = pose.R;
    f.t_cw = pose.t;
    const float vs = p.vs;
    const F3 origin{p.origin[0], p.origin[1], p.origin[2]};
    F3 lo = f3div(f3sub(pose.t, origin), vs);
    F3 hi = lo;
    const float z_far = max_d + p.trunc;
    const float cpx[4][2] = {{0.0f, 0.0f}, {(float)W, 0.0f}, {0.0f, (float)H}, {(float)W, (float)H}};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const F3 ray{(cpx[i][0] - K.cx) / K.fx * z_far, (cpx[i][1] - K.cy) / K.fy * z_far, z_far};
        const F3 v = f3div(f3sub(f3add(mulv(pose.R, ray), pose.t), origin), vs);
        lo = {smin(lo.x, v.x), smin(lo.y, v.y), smin(lo.z, v.z)};
        hi = {smax(hi.x, v.x), smax(hi.y, v.y), smax(hi.z, v.z)};
    }
    f.ok = f3finite(lo) && f3finite(hi);
    const float top = (float)(p.res - 1);
    auto clampIdx = [top](float v) { return (int)smax(0.0f, smin(top, v)); };
    f.x0 = clampIdx(floorf(lo.x));
    f.x1 = clampIdx(ceilf(hi.x));
    f.y0 = clampIdx(floorf(lo.y));
    f.y1 = clampIdx(ceilf(hi.y));
    f.z0 = clampIdx(floorf(lo.z));
    f.z1 = clampIdx(ceilf(hi.z));
    f.step = {f.R_wc.m[0][0] * vs, f.R_wc.m[1][0] * vs, f.R_wc.m[2][0] * vs};
    return f;
}

// Item pipeline constants and the frame's shared record (see the item pipeline below).
constexpr int kSubRows = 8;      // z rows per sub-tile
constexpr int kSlots = 256;      // work-list slots
constexpr int kGroupItems = 4;   // sub-tiles per wave step (4 x kSubRows = 32 phase-1 rows)
constexpr int kItemBlocks = 2048;  // integrateItems grid: 8192 waves, 32 per slot
static_assert(kGroupItems * kSubRows == 32, "a group fills the 32 phase-1 rows of a wave");
static_assert(kBrick % kSubRows == 0, "a sub-tile lies in one z-brick");

struct IntShared {
    IntFrame f;
    float cbound;           // bound on every camera coordinate's magnitude (tileClass's eps)
    // list lengths per slot, one word each so that a wave appends to both lists with one
    // atomic: exact in the low half, free in the high half
    unsigned long long cnt[kSlots];
};

// The frame's shared record (one lane) and the list lengths (cleared by the block).
__device__ __forceinline__ void itemSetup(IntShared* __restrict__ sh, const VolParams& p, const PoseArg& pa, Intr K,
                                          int W, int H, float max_d) {
    for (int i = threadIdx.x; i < kSlots; i += blockDim.x) sh->cnt[i] = 0ull;
    if (threadIdx.x == 0) {
        const IntFrame f = integrateSetup(p, pa, K, W, H, max_d);
        sh->f = f;
        sh->cbound = fabsf(f.t_wc.x) + fabsf(f.t_wc.y) + fabsf(f.t_wc.z) + fabsf(p.origin[0]) + fabsf(p.origin[1]) +
                     fabsf(p.origin[2]) + 3.0f * (float)p.res * p.vs;
    }
}

// Updated / newly-observed voxel counts. Each wave keeps its counts in registers and
// adds them once, at its end, to one of kCounterSlots slot pairs (chosen by block), so
// no address sees more than a few hundred atomics per frame; readers sum the slots.
constexpr int kCounterSlots = 256;

__device__ __forceinline__ void flushCounts(unsigned long long* counters, unsigned long long upd,
                                            unsigned long long newly) {
    if (!counters || (threadIdx.x & 63) != 0 || (upd | newly) == 0) return;
    unsigned long long* slot = counters + 2 * (blockIdx.x % kCounterSlots);
    if (upd) atomicAdd(&slot[0], upd);
    if (newly) atomicAdd(&slot[1], newly);
}

constexpr int kChunk = 32;  // x columns per tile
constexpr int kSegChunks = 4;  // tiles per wave along a row (integrateExact)
constexpr int kTileRows = 32;  // z rows per tile (integrateExact)
constexpr int kFoldBatch = 8;  // row pairs whose words one phase-2 step loads together
constexpr int kSegCols = kSegChunks * kChunk;
constexpr int kSub = 16;    // columns of a tile half (one phase-1 lane)
constexpr int kGather = 8;  // columns whose depth gathers are issued together
constexpr int kLdsStride = kChunk + 1;

// Depth bounds of the frame: per cell of 2^l x 2^l pixels (l = 1..kBoundLevels), the
// smallest and largest in-band depth {dmin, dmax}; dmin = -1 when any pixel of the cell
// is outside the band (or outside the image), dmax = -inf when none is inside. The
// integration uses them to decide whole tiles without projecting every voxel.
constexpr int kBoundLevels = 6;
struct DepthBounds {
    const float2* lvl;              // all levels, level l at lvl + off[l]
    int off[kBoundLevels + 1];
    int w[kBoundLevels + 1];        // cells per row
    int h[kBoundLevels + 1];
};

__device__ __forceinline__ float2 boundsJoin(float2 a, float2 b) {
    return make_float2((a.x < 0.0f || b.x < 0.0f) ? -1.0f : fminf(a.x, b.x), fmaxf(a.y, b.y));
}

// One block per 64 x 64 pixel region: levels 1..6 of its cells, reduced in LDS. When sh
// is set, one more block (the first of an extra row of the grid) writes the frame's
// shared record of the item pipeline, beside the others rather than ahead of one of them.
__global__ __launch_bounds__(256) void depthBoundsKernel(const float* __restrict__ depth, int W, int H, float min_d,
                                                         float max_d, float2* __restrict__ out, DepthBounds b,
                                                         IntShared* __restrict__ sh, VolParams p, PoseArg pa,
                                                         Intr K) {
    if (sh && blockIdx.y == gridDim.y - 1) {
        if (blockIdx.x == 0) itemSetup(sh, p, pa, K, W, H, max_d);
        return;
    }
    __shared__ float2 cell[32][33];
    const int tid = threadIdx.x;
    const int X0 = blockIdx.x * 64, Y0 = blockIdx.y * 64;
    // level 1: 32 x 32 cells of 2 x 2 pixels, 4 per thread; the 16 reads issued together
    float d[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = tid + 256 * q, cx = k & 31, cy = k >> 5;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = X0 + 2 * cx + (j & 1), y = Y0 + 2 * cy + (j >> 1);
            d[q][j] = (x < W && y < H) ? depth[(size_t)y * W + x] : __builtin_nanf("");
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = tid + 256 * q, cx = k & 31, cy = k >> 5;
        float2 acc = make_float2(__builtin_inff(), -__builtin_inff());
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // outside the image (NaN) or outside the band: {-1, -inf}
            const float dj = d[q][j];
            const float2 pxb = (dj >= min_d && dj <= max_d) ? make_float2(dj, dj) : make_float2(-1.0f, -__builtin_inff());
            acc = j == 0 ? pxb : boundsJoin(acc, pxb);
        }
        cell[cy][cx] = acc;
        const int gx = blockIdx.x * 32 + cx, gy = blockIdx.y * 32 + cy;
        if (gx < b.w[1] && gy < b.h[1]) out[b.off[1] + gy * b.w[1] + gx] = acc;
    }
    __syncthreads();
    for (int l = 2, n = 16; l <= kBoundLevels; ++l, n >>= 1) {
        float2 v = make_float2(0.0f, 0.0f);
        const bool act = tid < n * n;
        const int cx = tid % n, cy = tid / n;
        if (act)
            v = boundsJoin(boundsJoin(cell[2 * cy][2 * cx], cell[2 * cy][2 * cx + 1]),
                           boundsJoin(cell[2 * cy + 1][2 * cx], cell[2 * cy + 1][2 * cx + 1]));
        __syncthreads();
        if (act) {
            cell[cy][cx] = v;
            const int gx = blockIdx.x * n + cx, gy = blockIdx.y * n + cy;
            if (gx < b.w[l] && gy < b.h[l]) out[b.off[l] + gy * b.w[l] + gx] = v;
        }
        __syncthreads();
    }
}

// Bounds over the pixel rectangle [u0, u1] x [v0, v1] (inside the image), from the
// coarsest level needed to cover it with at most 2 x 2 cells; a rectangle wider than a
// level-kBoundLevels cell takes up to 4 x 4 cells of that level; false beyond.
__device__ __forceinline__ bool boundsOver(const DepthBounds& b, int u0, int u1, int v0, int v1, float2& r) {
    const int ext = max(u1 - u0, v1 - v0) + 1;
    int l = 1;
    while ((1 << l) < ext && l < kBoundLevels) ++l;
    const int cx0 = u0 >> l, cx1 = u1 >> l, cy0 = v0 >> l, cy1 = v1 >> l;
    const float2* L = b.lvl + b.off[l];
    if (cx1 - cx0 < 2 && cy1 - cy0 < 2) {
        r = boundsJoin(boundsJoin(L[cy0 * b.w[l] + cx0], L[cy0 * b.w[l] + cx1]),
                       boundsJoin(L[cy1 * b.w[l] + cx0], L[cy1 * b.w[l] + cx1]));
        return true;
    }
    if (cx1 - cx0 > 3 || cy1 - cy0 > 3) return false;
    r = L[cy0 * b.w[l] + cx0];
    for (int cy = cy0; cy <= cy1; ++cy)
        for (int cx = cx0; cx <= cx1; ++cx) r = boundsJoin(r, L[cy * b.w[l] + cx]);
    return true;
}

// The same bounds from a finer level: the finest one on which the rectangle spans at
// most kFineCells cells per axis (at most kFineCells^2 reads instead of 4), so the
// bounds come from little more than the footprint itself.
constexpr int kFineCells = 6;
__device__ __forceinline__ bool boundsOverFine(const DepthBounds& b, int u0, int u1, int v0, int v1, float2& r) {
    int l = 1;
    while (l < kBoundLevels && (max((u1 >> l) - (u0 >> l), (v1 >> l) - (v0 >> l)) + 1) > kFineCells) ++l;
    const int cx0 = u0 >> l, cx1 = u1 >> l, cy0 = v0 >> l, cy1 = v1 >> l;
    if (cx1 - cx0 + 1 > kFineCells || cy1 - cy0 + 1 > kFineCells) return boundsOver(b, u0, u1, v0, v1, r);
    const float2* L = b.lvl + b.off[l];
    const int w = b.w[l];
    // A row of cells at a time, its kFineCells reads issued before the first join (the
    // cells past the rectangle's last column repeat it; the join is idempotent).
    r = L[cy0 * w + cx0];
    for (int cy = cy0; cy <= cy1; ++cy) {
        float2 c[kFineCells];
#pragma unroll
        for (int i = 0; i < kFineCells; ++i) c[i] = L[cy * w + min(cx0 + i, cx1)];
#pragma unroll
        for (int i = 0; i < kFineCells; ++i) r = boundsJoin(r, c[i]);
    }
    return true;
}

// Wave-uniform classification of the tile {xa..xb} x {y} x {za..zb} (a rectangle of
// lattice points at fixed y), decided from its four corners: the rectangle's camera
// coordinates are affine in (x, z), and the perspective image of a rectangle in front of
// the camera is the convex hull of its corners' images. The voxels' exact coordinates
// (the incremental x-step) differ from the affine ones by at most eps (the rounding of
// every step so far, bounded from the coordinates' magnitude), and eps is carried into
// every margin, so the class is the outcome the exact per-voxel path would reach:
//   kTileCull: no voxel can update (behind the camera, beyond max_depth + trunc, wholly
//              outside one image edge, or behind every in-band depth of its footprint by
//              more than trunc);
//   kTileFree: every voxel projects inside the image onto an in-band pixel at least
//              trunc + margin beyond it, so it folds tsdf_new = 1 (sdf >= trunc gives
//              sdf / trunc >= 1);
//   kTileExact: anything else, decided voxel by voxel.
constexpr int kTileCull = 0, kTileExact = 1, kTileFree = 2;

__device__ __forceinline__ int tileClass(const IntFrame& f, const VolParams& p, Intr K, int W, int H, float max_d,
                                         int xa, int xb, int y, int za, int zb, const DepthBounds& bnd,
                                         float cbound, bool fine = false) {
    float zmin = __builtin_inff(), zmax = -__builtin_inff();
    float umin = __builtin_inff(), umax = -__builtin_inff(), vmin = __builtin_inff(), vmax = -__builtin_inff();
    bool left = true, right = true, top = true, bottom = true, front = true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float px = p.origin[0] + (float)((i & 1) ? xb : xa) * p.vs;
        const float py = p.origin[1] + (float)y * p.vs;
        const float pz = p.origin[2] + (float)((i & 2) ? zb : za) * p.vs;
        const F3 c = f3add(mulv(f.R_wc, F3{px, py, pz}), f.t_wc);
        zmin = fminf(zmin, c.z);
        zmax = fmaxf(zmax, c.z);
        if (c.z > 0.1f) {
            const float u = K.fx * c.x / c.z + K.cx + 0.5f;
            const float v = K.fy * c.y / c.z + K.cy + 0.5f;
            left = left && u < -2.0f;
            right = right && u > (float)W + 2.0f;
            top = top && v < -2.0f;
            bottom = bottom && v > (float)H + 2.0f;
            umin = fminf(umin, u);
            umax = fmaxf(umax, u);
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
        } else {
            left = right = top = bottom = front = false;
        }
    }
    if (zmax < -1e-3f) return kTileCull;
    if (zmin > max_d + p.trunc + 1e-3f) return kTileCull;
    if (left || right || top || bottom) return kTileCull;
    if (!front || !bnd.lvl) return kTileExact;
    // eps: a float add rounds by at most 2^-24 of a magnitude <= cbound, once per step
    // of the row so far and a few times in its start point.
    const float eps = 1.2e-7f * cbound * (float)(xb - f.x0 + 16) + 1e-5f;
    if (!(zmin - eps > 0.1f)) return kTileExact;
    const float mpx = 2.0f + eps / (zmin - eps) * (fmaxf(K.fx, K.fy) + (float)max(W, H) + 4.0f);
    const float u0f = floorf(umin - mpx), u1f = floorf(umax + mpx), v0f = floorf(vmin - mpx), v1f = floorf(vmax + mpx);
    const bool inside = u0f >= 0.0f && v0f >= 0.0f && u1f < (float)W && v1f < (float)H;
    const int u0 = (int)fmaxf(u0f, 0.0f), v0 = (int)fmaxf(v0f, 0.0f);
    const int u1 = (int)fminf(u1f, (float)(W - 1)), v1 = (int)fminf(v1f, (float)(H - 1));
    if (u0 > u1 || v0 > v1) return kTileExact;
    float2 r;
    if (!(fine ? boundsOverFine(bnd, u0, u1, v0, v1, r) : boundsOver(bnd, u0, u1, v0, v1, r))) return kTileExact;
    const float gap = p.trunc + 2.0f * eps + 1e-4f;
    if (r.y + gap < zmin) return kTileCull;                      // behind every in-band depth
    if (inside && r.x >= 0.0f && r.x - gap > zmax) return kTileFree;  // in front of every depth, all in band
    return kTileExact;
}

static_assert(kChunk == 2 * kBrick, "a tile spans two bricks in x");

__host__ __device__ inline int tileGrid(int res) { return (res + kChunk - 1) / kChunk; }

// One voxel word through a raw buffer resource (byte offsets: lane + uniform).
template <class S> struct BufWord;
template <> struct BufWord<StoreF16> {
    __device__ static uint32_t load(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0);
    }
    __device__ static void store(uint32_t v, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, off, soff, 0);
    }
};
template <> struct BufWord<StoreF32> {
    __device__ static float2 load(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    }
    __device__ static void store(float2 v, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, off, soff, 0);
    }
};

// Brick-map bytes of one aligned tile half (integrateExact): x columns xc + j, j < 32,
// lie in x-brick bx = xc / 16 + j / 16 (lane bits j and 32 + j); columns 0 and 16 are
// also the first layers of bricks bx - 1 and bx. y = 16 q is also in y-brick q - 1.
// near_h / shared_h as built by the caller for z-bricks bz + h.
__device__ __forceinline__ void markTile(unsigned char* __restrict__ bricks, int nb, int bx, int y, int bz,
                                         unsigned long long near0, unsigned long long near1,
                                         unsigned long long shared0, unsigned long long shared1) {
    const unsigned long long lo16 = 0x0000ffff0000ffffull, hi16 = lo16 << 16;
    const unsigned long long col0 = 0x0000000100000001ull, col16 = col0 << 16;
    const int by = y / kBrick;
    const int ny = (y % kBrick == 0 && by > 0) ? 2 : 1;
    for (int h = 0; h < 2; ++h)
        for (int part = 0; part < 2; ++part) {  // the z-brick, then (shared layer) the one below
            const unsigned long long m =
                part == 0 ? (h == 0 ? near0 : near1) : ((h == 0 ? shared0 : shared1) & 0xffffffffull);
            const int zq = bz + h - part;
            if (m == 0 || zq < 0 || zq >= nb) continue;
            for (int yy = 0; yy < ny; ++yy) {
                unsigned char* row = bricks + ((size_t)zq * nb + (by - yy)) * nb;
                if (m & (lo16 | col16)) row[bx] = 1;
                if ((m & hi16) && bx + 1 < nb) row[bx + 1] = 1;
                if ((m & col0) && bx > 0) row[bx - 1] = 1;
            }
        }
}

template <class S, int FD, bool RC>
__global__ __launch_bounds__(256, 8) void integrateExact(typename S::Word* __restrict__ vox, VolParams p, Intr K,
                                                      int W, int H, const float* __restrict__ depth, float min_d,
                                                      float max_d, PoseArg pa, FixedDiv dtr, DepthBounds bnd,
                                                      unsigned long long* __restrict__ counters,
                                                      unsigned char* __restrict__ bricks, int nb) {
    __shared__ float lds[4][kTileRows * kLdsStride];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const IntFrame f = integrateSetup(p, pa, K, W, H, max_d);
    if (!f.ok) return;
    // Tiles start on multiples of kTileRows rows in z and of kChunk columns in x (rows
    // and columns of a tile then fall into bricks by their index alone, and a tile row of
    // FP16 words is one 128-byte line); rows and columns outside the frustum box stay
    // idle. The aligned groups also measure faster than groups starting at the box (FP16
    // 290 against 347 us at 512^3 on the same frames, with 64-row tiles).
    // A row is cut into segments of kSegChunks tiles, one wave each, so that a frame has
    // several waves per SIMD to hide its gathers behind; a wave carries its rows' camera
    // coordinates to its segment by the same serial x-steps the whole row takes.
    const int ny = f.y1 - f.y0 + 1;
    const int zbase = f.z0 & ~(kTileRows - 1);
    const int nzg = (f.z1 - zbase + kTileRows) / kTileRows;
    const int xstart = f.x0 & ~(kChunk - 1);
    const int nseg = (f.x1 - xstart + kSegCols) / kSegCols;
    const int task = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);  // wave-uniform
    if (task >= ny * nzg * nseg) return;  // whole wave; the kernel uses no block barrier
    const int seg = task / (ny * nzg), rem = task % (ny * nzg);
    const int y = f.y0 + rem % ny;
    const int zb = zbase + (rem / ny) * kTileRows;
    const int xs = xstart + seg * kSegCols, xe = min(xs + kSegCols - 1, f.x1);
    const int za = max(zb, f.z0), ze = min(zb + kTileRows - 1, f.z1);
    // Phase-1 lanes: row r of the tile, column half h (columns 16 h .. 16 h + 15).
    const int r = lane & (kTileRows - 1), h = lane >> 5;
    const int z = zb + r;
    const bool row_ok = z >= f.z0 && z <= f.z1;
    const float vs = p.vs, trunc = p.trunc;
    const F3 pp{p.origin[0] + (float)f.x0 * vs, p.origin[1] + (float)y * vs, p.origin[2] + (float)z * vs};
    F3 c = f3add(mulv(f.R_wc, pp), f.t_wc);
    for (int x = f.x0; x < xs; ++x) c = f3add(c, f.step);
    float* L = lds[wv];
    const size_t res = (size_t)p.res, plane = res * res;
    const float fW = (float)W, fH = (float)H;
    unsigned n_upd = 0, n_new = 0;  // per lane, summed over the wave at the end
    // A bound on every camera coordinate's magnitude over the volume (for tileClass).
    const float cbound = fabsf(f.t_wc.x) + fabsf(f.t_wc.y) + fabsf(f.t_wc.z) +
                         fabsf(p.origin[0]) + fabsf(p.origin[1]) + fabsf(p.origin[2]) + 3.0f * (float)p.res * vs;

    // The classes of the segment's tiles, lane t taking tile t, all tiles at once so that
    // their depth-bounds reads overlap. They come back as two wave masks.
    unsigned long long cls_lo, cls_hi;
    {
        int my = kTileCull;
        const int xt = xs + lane * kChunk;
        if (lane < kSegChunks && xt <= xe) {
            const int js = max(f.x0 - xt, 0), n = min(kChunk, f.x1 - xt + 1);
            my = tileClass(f, p, K, W, H, max_d, xt + js, xt + n - 1, y, za, ze, bnd, cbound);
        }
        cls_lo = __ballot((my & 1) != 0);
        cls_hi = __ballot((my & 2) != 0);
    }
    for (int xc = xs; xc <= xe; xc += kChunk) {
        // columns [js, n) of the tile are in the frustum box
        const int js = max(f.x0 - xc, 0), n = min(kChunk, f.x1 - xc + 1);
        const int ti = (xc - xs) / kChunk;
        const int cls = (int)((cls_lo >> ti) & 1ull) | ((int)((cls_hi >> ti) & 1ull) << 1);
        if (cls != kTileExact) {
            // The row's camera coordinate still advances step by step.
            for (int j = js; j < n; ++j) c = f3add(c, f.step);
            if (cls == kTileCull) continue;
        }
        // Phase 1 (exact tiles): lane = (row, column half), serial over the half's 16
        // columns: the projections first, then all depth gathers at once, then the gates.
        // The second-half lane first carries its row across the first half's columns (the
        // same adds the first-half lane makes while it projects them), and afterwards both
        // lanes of the row continue from its coordinate, which has taken every column.
        if (cls == kTileExact) {
            for (int j = js; j < min(n, kSub); ++j)
                if (h) c = f3add(c, f.step);
#pragma unroll 1
            for (int j0 = h * kSub; j0 < h * kSub + kSub; j0 += kGather) {
                float cz[kGather], dj[kGather];
                int pix[kGather];
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) {
                    pix[jj] = -1;
                    cz[jj] = c.z;
                    if (j0 + jj >= js && j0 + jj < n) {
                        if (row_ok && c.z > 0.0f) {
                            const float inv_z = recipBy<RC>(c.z);
                            const float uf = floorf(K.fx * c.x * inv_z + K.cx + 0.5f);
                            const float vf = floorf(K.fy * c.y * inv_z + K.cy + 0.5f);
                            if (uf >= 0.0f && uf < fW && vf >= 0.0f && vf < fH) pix[jj] = (int)vf * W + (int)uf;
                        }
                        c = f3add(c, f.step);
                    }
                }
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) dj[jj] = depth[pix[jj] < 0 ? 0 : pix[jj]];
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) {
                    float tn = __builtin_nanf("");
                    const float d = dj[jj];
                    if (pix[jj] >= 0 && d >= min_d && d <= max_d) {
                        const float sdf = d - cz[jj];
                        if (!(sdf < -trunc)) tn = smin(1.0f, divBy<FD>(dtr, sdf));
                    }
                    L[r * kLdsStride + j0 + jj] = tn;
                }
            }
            c.x = __shfl(c.x, r + 32);
            c.y = __shfl(c.y, r + 32);
            c.z = __shfl(c.z, r + 32);
        }
        waveSync();
        // Phase 2: lane = column, two rows per instruction; per half of the rows, every
        // load is issued before the first fold.
        const int j = lane & (kChunk - 1), g = lane >> 5;
        const bool col_ok = j >= js && j < n;
        // The tile's words through a buffer resource: base = the tile's first voxel
        // (wave-uniform), lane offset (g plane + j) words, row offset 2 r plane words
        // (uniform). The range is capped at 2^31 bytes (a tile's 32 planes of at most
        // 2048^2 words stay below it), so the lane offset 2^31 reads 0 with no memory access.
        const size_t tile0 = (size_t)zb * plane + (size_t)y * res + (size_t)xc;
        const size_t left = (res * plane - tile0) * sizeof(typename S::Word);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            vox + tile0, (short)0, (int)(left < 0x80000000ull ? left : 0x80000000ull), 0x00020000);
        const unsigned lane_off = (unsigned)(((size_t)g * plane + (size_t)j) * sizeof(typename S::Word));
        const unsigned row_step = (unsigned)(2 * plane * sizeof(typename S::Word));
        // rows r with zb + g + 2 r in [z0, z1] (none for a column outside the box)
        const int r_lo = col_ok ? max(0, (f.z0 - zb - g + 1) >> 1) : kTileRows / 2;
        const int r_hi = min(kTileRows / 2 - 1, (f.z1 - zb - g) >> 1);
        // Brick map: with the tile aligned, row pair kk of either lane half lies in z-brick
        // zb / 16 for kk < 8 and the next one for kk >= 8; near_h are the lanes that left a
        // non-far voxel there, shared_h the lower-half lanes whose voxel at z = 16 q (pair
        // kk = 8 h) was non-far (that layer is also the brick below's).
        unsigned long long near0 = 0, near1 = 0, shared0 = 0, shared1 = 0;
#pragma unroll 1
        for (int k0 = 0; k0 < kTileRows / 2; k0 += kFoldBatch) {
            float tn[kFoldBatch];
            typename S::Word wd[kFoldBatch];
#pragma unroll
            for (int k = 0; k < kFoldBatch; ++k) {
                const int row = g + 2 * (k0 + k);
                const float t = cls == kTileFree ? 1.0f : L[row * kLdsStride + j];
                tn[k] = (k0 + k >= r_lo && k0 + k <= r_hi) ? t : __builtin_nanf("");
            }
            // Lanes with nothing to fold read out of range (no access), so the loads
            // carry no branch and issue back to back.
#pragma unroll
            for (int k = 0; k < kFoldBatch; ++k)
                wd[k] = BufWord<S>::load(rs, __builtin_isnan(tn[k]) ? 0x80000000u : lane_off,
                                         (unsigned)(k0 + k) * row_step);
#pragma unroll
            for (int k = 0; k < kFoldBatch; ++k) {
                const int kk = k0 + k;
                const bool upd = !__builtin_isnan(tn[k]);
                bool newly = false, nf = false;
                if (upd) {
                    const float w_old = S::weight(wd[k]);
                    const float t_old = S::tsdf(wd[k]);
                    newly = (w_old == 0.0f);
                    const float t_new = (t_old * w_old + tn[k]) / (w_old + 1.0f);
                    BufWord<S>::store(S::pack(t_new, smin(w_old + 1.0f, p.max_w)), rs, lane_off,
                                      (unsigned)kk * row_step);
                    nf = !S::farPacked(t_new);  // the new weight is >= 1
                }
                n_upd += upd ? 1u : 0u;
                n_new += newly ? 1u : 0u;
                if (bricks) {
                    const unsigned long long m = __ballot(nf);
                    if (kk < 8) near0 |= m;
                    else near1 |= m;
                    if (kk == 0) shared0 = __ballot(nf && g == 0);
                    if (kk == 8) shared1 = __ballot(nf && g == 0);
                }
            }
        }
        if (bricks && lane == 0 && (near0 | near1) != 0)
            markTile(bricks, nb, xc / kBrick, y, zb / kBrick, near0, near1, shared0, shared1);
        waveSync();
    }
    if (counters) {
        unsigned long long a = n_upd, b = n_new;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        flushCounts(counters, a, b);
    }
}

// ---------------------------------------------------------------------------
// Item pipeline of the exact integration (the default when the volume has its scratch).
// The same per-voxel arithmetic as integrateExact, bit for bit, organised so that only
// the voxels that can update are visited:
//   * integrateSetup runs once per frame (one lane of depthBoundsKernel's extra block) into an
//     IntShared record that every later kernel of the frame reads;
//   * subtileClassKernel classifies sub-tiles of 32 columns x kSubRows rows at one y (a
//     quarter of a tile's rows: a tighter depth range per unit, so more of them are free
//     or culled) and appends every exact and every free one to a work list;
//   * itemCarryKernel walks each row of the frustum box by the serial x-steps and stores
//     the row's camera coordinate at the first column of every exact sub-tile it holds;
//   * integrateItems: a persistent grid whose waves take groups of kGroupItems sub-tiles
//     from the lists; exact groups run integrateExact's phase 1 (projection, gather,
//     gates into LDS) from the stored coordinates, then every group folds its voxels as
//     integrateExact's phase 2 (free sub-tiles with tsdf_new = 1, no projection).
// The lists are split into kSlots slots (a class block appends to slot blockIdx % kSlots),
// so that no list counter takes more than a few dozen atomics a frame; the order of the
// items inside a list is whatever the atomics give, and the result does not depend on it
// (every voxel belongs to one item; counts are sums; brick marks are idempotent).
// ---------------------------------------------------------------------------
__host__ __device__ inline int subGroups(int res) { return (res + kSubRows - 1) / kSubRows; }
__host__ __device__ inline size_t subCount(int res) { return (size_t)res * subGroups(res) * tileGrid(res); }
__host__ __device__ inline unsigned classBlocks(int res) { return (unsigned)((subCount(res) + 255) / 256); }
// Capacity of one slot's list: every sub-tile of the class blocks that append to it.
__host__ __device__ inline unsigned slotCap(int res) { return (classBlocks(res) + kSlots - 1) / kSlots * 256u; }

// Scratch of the item pipeline: the shared record, the exact-item index of every sub-tile
// (-1: not exact), the two lists, and kSubRows row coordinates per exact item.
struct ItemScratch {
    IntShared* sh;
    int* map;
    unsigned* lists;  // [2][kSlots][cap]
    float4* carry;    // [kSlots * cap][kSubRows]
    unsigned cap;
};
static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
static ItemScratch itemScratch(void* base, int res) {
    ItemScratch s;
    unsigned char* b = static_cast<unsigned char*>(base);
    s.cap = slotCap(res);
    const size_t nl = (size_t)kSlots * s.cap;
    s.sh = reinterpret_cast<IntShared*>(b);
    b += align256(sizeof(IntShared));
    s.map = reinterpret_cast<int*>(b);
    b += align256(subCount(res) * sizeof(int));
    s.lists = reinterpret_cast<unsigned*>(b);
    b += align256(2 * nl * sizeof(unsigned));
    s.carry = reinterpret_cast<float4*>(b);
    return s;
}
size_t exactScratchBytes(int res) {
    const size_t nl = (size_t)kSlots * slotCap(res);
    return align256(sizeof(IntShared)) + align256(subCount(res) * sizeof(int)) + align256(2 * nl * sizeof(unsigned)) +
           nl * kSubRows * sizeof(float4);
}

// A list entry: the sub-tile's y, z group and x tile.
__device__ __forceinline__ unsigned itemPack(int y, int zg, int xi) {
    return ((unsigned)y << 14) | ((unsigned)zg << 6) | (unsigned)xi;
}
__device__ __forceinline__ void itemUnpack(unsigned e, int& y, int& zg, int& xt) {
    xt = (int)(e & 63u) * kChunk;
    zg = (int)((e >> 6) & 255u);
    y = (int)(e >> 14);
}

// One lane per sub-tile of the volume (lanes outside the frame's frustum box take none):
// tileClass over the sub-tile's rows inside the box, then the appends.
__global__ __launch_bounds__(256) void subtileClassKernel(VolParams p, Intr K, int W, int H, float max_d,
                                                          DepthBounds bnd, ItemScratch is) {
    const IntFrame f = is.sh->f;
    const float cbound = is.sh->cbound;
    const int lane = threadIdx.x & 63;
    const int nzg = subGroups(p.res), ntx = tileGrid(p.res);
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    int cls = kTileCull, y = 0, zg = 0, xi = 0;
    bool inbox = false;
    if (f.ok && i < (long long)subCount(p.res)) {
        xi = (int)(i % ntx);
        const long long rest = i / ntx;
        zg = (int)(rest % nzg);
        y = (int)(rest / nzg);
        const int xt = xi * kChunk, zb = zg * kSubRows;
        const int za = max(zb, f.z0), ze = min(zb + kSubRows - 1, f.z1);
        const int xa = max(xt, f.x0), xe = min(xt + kChunk - 1, f.x1);
        inbox = y >= f.y0 && y <= f.y1 && za <= ze && xa <= xe;
        if (inbox) cls = tileClass(f, p, K, W, H, max_d, xa, xe, y, za, ze, bnd, cbound, true);
    }
    const int slot = (int)(blockIdx.x % kSlots);
    const unsigned long long me = __ballot(cls == kTileExact), mf = __ballot(cls == kTileFree);
    if ((me | mf) != 0ull) {
        unsigned long long base = 0ull;
        if (lane == 0)
            base = atomicAdd(&is.sh->cnt[slot], (unsigned long long)__popcll(me) |
                                                    ((unsigned long long)__popcll(mf) << 32));
        base = __shfl(base, 0);
        const bool want = cls == kTileExact || cls == kTileFree;
        if (want) {
            const int l = cls == kTileExact ? 0 : 1;
            const unsigned long long m = l == 0 ? me : mf;
            const unsigned pos = (l == 0 ? (unsigned)base : (unsigned)(base >> 32)) +
                                 __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            is.lists[((size_t)l * kSlots + slot) * is.cap + pos] = itemPack(y, zg, xi);
            if (l == 0) is.map[i] = (int)((unsigned)slot * is.cap + pos);
        }
    }
    if (inbox && cls != kTileExact) is.map[i] = -1;
}

// One lane per row (y, z) of the frustum box: the row's camera coordinate from the box's
// first column, stepped serially (integrateExact's own steps, in its order), stored at
// the first in-box column of every exact sub-tile of the row; no step is taken past the
// last of them.
__global__ __launch_bounds__(256) void itemCarryKernel(VolParams p, ItemScratch is) {
    const IntFrame f = is.sh->f;
    if (!f.ok) return;
    const int ny = f.y1 - f.y0 + 1, nz = f.z1 - f.z0 + 1;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)ny * nz) return;
    const int z = f.z0 + (int)(i % nz), y = f.y0 + (int)(i / nz);
    const int ntx = tileGrid(p.res);
    const int t0 = f.x0 / kChunk, t1 = f.x1 / kChunk;  // t1 - t0 < 64 (res <= 2048)
    const int* row = is.map + ((size_t)y * subGroups(p.res) + z / kSubRows) * ntx;
    // The row's tiles 16 at a time, their map entries read together: the coordinate is
    // stepped tile by tile while exact tiles remain ahead (uniform tile loop, no reads in
    // it), and stored at each exact tile's first in-box column.
    const float vs = p.vs;
    const F3 pp{p.origin[0] + (float)f.x0 * vs, p.origin[1] + (float)y * vs, p.origin[2] + (float)z * vs};
    F3 c = f3add(mulv(f.R_wc, pp), f.t_wc);
    int x = f.x0;
    for (int base = t0; base <= t1; base += 16) {
        int v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = row[min(base + k, t1)];
        unsigned emask = 0u;  // bit k: tile base + k is exact
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (base + k <= t1 && v[k] >= 0) emask |= 1u << k;
        // exact tiles in later chunks keep this row stepping through the whole chunk
        bool later = false;
        for (int t = base + 16; t <= t1; ++t) later = later || row[t] >= 0;
        if (emask == 0u && !later) break;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if ((emask >> k) == 0u && !later) break;
            const int xa = max((base + k) * kChunk, f.x0);
            int n = xa - x;
            for (; n >= 8; n -= 8) {
#pragma unroll
                for (int u = 0; u < 8; ++u) c = f3add(c, f.step);
            }
            for (; n > 0; --n) c = f3add(c, f.step);
            x = xa;
            if ((emask >> k) & 1u) is.carry[(size_t)v[k] * kSubRows + (z % kSubRows)] = make_float4(c.x, c.y, c.z, 0.0f);
        }
    }
}

// Brick-map bytes of one sub-tile: mask bit j = column xt + j left a non-far voxel in a
// row of the sub-tile (near) / in its first row (first). Columns 0-16 lie in x-brick
// bx = xt / 16, 16-31 in bx + 1, column 0 also in bx - 1; the rows lie in z-brick
// zb / 16, and row zb = 16 q is also the last layer of z-brick q - 1; y = 16 q is also
// in y-brick q - 1.
__device__ __forceinline__ void markItem(unsigned char* __restrict__ bricks, int nb, int xt, int y, int zb,
                                         unsigned near, unsigned first) {
    const int bx = xt / kBrick, by = y / kBrick, bz = zb / kBrick;
    const int nyb = (y % kBrick == 0 && by > 0) ? 2 : 1;
    for (int part = 0; part < 2; ++part) {
        const unsigned m = part == 0 ? near : ((zb % kBrick == 0) ? first : 0u);
        const int zq = bz - part;
        if (m == 0 || zq < 0) continue;
        for (int yy = 0; yy < nyb; ++yy) {
            unsigned char* row = bricks + ((size_t)zq * nb + (by - yy)) * nb;
            if (m & 0x1ffffu) row[bx] = 1;
            if ((m & 0xffff0000u) && bx + 1 < nb) row[bx + 1] = 1;
            if ((m & 1u) && bx > 0) row[bx - 1] = 1;
        }
    }
}

// ONE: the volume spans at most 2^31 bytes, so one buffer resource over the whole volume
// serves every item (the item's first voxel as the uniform offset).
template <class S, int FD, bool RC, bool ONE>
__global__ __launch_bounds__(256, 8) void integrateItems(typename S::Word* __restrict__ vox, VolParams p, Intr K,
                                                      int W, int H, const float* __restrict__ depth, float min_d,
                                                      float max_d, FixedDiv dtr, ItemScratch is,
                                                      unsigned long long* __restrict__ counters,
                                                      unsigned char* __restrict__ bricks, int nb) {
    __shared__ float lds[4][32 * kLdsStride];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const IntFrame f = is.sh->f;
    if (!f.ok) return;
    // Wave w serves slot w % kSlots and takes every parts-th group of its lists, the
    // exact groups first.
    const int wave = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + wv);
    const int parts = (int)gridDim.x * 4 / kSlots;
    const int slot = wave % kSlots, part = wave / kSlots;
    const unsigned long long nboth = is.sh->cnt[slot];
    const unsigned ne = (unsigned)nboth, nfr = (unsigned)(nboth >> 32);
    const unsigned ge = (ne + kGroupItems - 1) / kGroupItems, gf = (nfr + kGroupItems - 1) / kGroupItems;
    const unsigned* elist = is.lists + (size_t)slot * is.cap;
    const unsigned* flist = is.lists + ((size_t)kSlots + slot) * is.cap;
    float* L = lds[wv];
    const size_t res = (size_t)p.res, plane = res * res;
    const float fW = (float)W, fH = (float)H, trunc = p.trunc;
    // (unused without ONE)
    const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
        vox, (short)0, ONE ? (int)(res * plane * sizeof(typename S::Word)) : 0, 0x00020000);
    unsigned n_upd = 0, n_new = 0;  // per lane, summed over the wave at the end
    // Phase-1 lanes: row r of the group (item r / kSubRows, row r % kSubRows), column half
    // h; phase-2 lanes: column j, row parity g.
    const int r = lane & 31, h = lane >> 5, j = lane & 31, g = lane >> 5;
    for (unsigned grp = (unsigned)part; grp < ge + gf; grp += (unsigned)parts) {
        const bool exact = grp < ge;  // wave-uniform
        const unsigned first = (exact ? grp : grp - ge) * kGroupItems;
        const unsigned nitems = min((unsigned)kGroupItems, (exact ? ne : nfr) - first);
        const unsigned* list = exact ? elist : flist;
        if (exact) {
            // Phase 1: projection, depth gather and gates of the exact items' rows (the
            // code of integrateExact's phase 1), tsdf_new (NaN = no update) into LDS.
            const int it = r / kSubRows, row = r % kSubRows;
            const bool act = (unsigned)it < nitems;
            int y, zg, xt;
            itemUnpack(act ? list[first + it] : 0u, y, zg, xt);
            const int z = zg * kSubRows + row;
            const bool row_ok = act && z >= f.z0 && z <= f.z1;
            const int js = max(f.x0 - xt, 0), n = min(kChunk, f.x1 - xt + 1);
            F3 c{0.0f, 0.0f, 0.0f};
            if (row_ok) {
                const float4 q = is.carry[((size_t)slot * is.cap + first + it) * kSubRows + row];
                c = F3{q.x, q.y, q.z};
            }
            for (int jj = js; jj < min(n, kSub); ++jj)
                if (h) c = f3add(c, f.step);
#pragma unroll 1
            for (int j0 = h * kSub; j0 < h * kSub + kSub; j0 += kGather) {
                float cz[kGather], dj[kGather];
                int pix[kGather];
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) {
                    pix[jj] = -1;
                    cz[jj] = c.z;
                    if (j0 + jj >= js && j0 + jj < n) {
                        if (row_ok && c.z > 0.0f) {
                            const float inv_z = recipBy<RC>(c.z);
                            const float uf = floorf(K.fx * c.x * inv_z + K.cx + 0.5f);
                            const float vf = floorf(K.fy * c.y * inv_z + K.cy + 0.5f);
                            if (uf >= 0.0f && uf < fW && vf >= 0.0f && vf < fH) pix[jj] = (int)vf * W + (int)uf;
                        }
                        c = f3add(c, f.step);
                    }
                }
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) dj[jj] = depth[pix[jj] < 0 ? 0 : pix[jj]];
#pragma unroll
                for (int jj = 0; jj < kGather; ++jj) {
                    float tn = __builtin_nanf("");
                    const float d = dj[jj];
                    if (pix[jj] >= 0 && d >= min_d && d <= max_d) {
                        const float sdf = d - cz[jj];
                        if (!(sdf < -trunc)) tn = smin(1.0f, divBy<FD>(dtr, sdf));
                    }
                    L[r * kLdsStride + j0 + jj] = tn;
                }
            }
            waveSync();
        }
        // Phase 2, two items at a time: lane = column, rows g, g + 2, g + 4, g + 6 of each
        // item; every load of the pair is issued before the first fold. Voxels are read
        // and written through a buffer resource based at the item's first voxel; a lane
        // with nothing to fold reads at the out-of-range offset 2^31 (no access).
        constexpr int kPairRows = kSubRows / 2;
        const unsigned lane_off = (unsigned)(((size_t)g * plane + (size_t)j) * sizeof(typename S::Word));
        const unsigned row_step = (unsigned)(2 * plane * sizeof(typename S::Word));
#pragma unroll 1
        for (unsigned i0 = 0; i0 < nitems; i0 += 2) {
            __amdgpu_buffer_rsrc_t rs[2];
            unsigned ibase[2];  // ONE: the item's first voxel's byte offset
            int iy[2], izb[2], ixt[2];
            float tn[2 * kPairRows];
            typename S::Word wd[2 * kPairRows];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const bool has = i0 + q < nitems;
                int y, zg, xt;
                itemUnpack(has ? __builtin_amdgcn_readfirstlane(list[first + i0 + q]) : 0u, y, zg, xt);
                const int zb = zg * kSubRows;
                iy[q] = y;
                izb[q] = zb;
                ixt[q] = xt;
                const size_t item0 = (size_t)zb * plane + (size_t)y * res + (size_t)xt;
                if (ONE) {
                    rs[q] = vrs;
                    ibase[q] = (unsigned)(item0 * sizeof(typename S::Word));
                } else {
                    const size_t left = (res * plane - item0) * sizeof(typename S::Word);
                    rs[q] = __builtin_amdgcn_make_buffer_rsrc(vox + item0, (short)0,
                                                              (int)(left < 0x80000000ull ? left : 0x80000000ull),
                                                              0x00020000);
                    ibase[q] = 0u;
                }
                const int js = max(f.x0 - xt, 0), n = min(kChunk, f.x1 - xt + 1);
                const bool col_ok = has && j >= js && j < n;
#pragma unroll
                for (int k = 0; k < kPairRows; ++k) {
                    const int row = g + 2 * k, z = zb + row;
                    const float t = exact ? L[((int)(i0 + q) * kSubRows + row) * kLdsStride + j] : 1.0f;
                    tn[q * kPairRows + k] = (col_ok && z >= f.z0 && z <= f.z1) ? t : __builtin_nanf("");
                }
            }
#pragma unroll
            for (int k = 0; k < 2 * kPairRows; ++k)
                wd[k] = BufWord<S>::load(rs[k / kPairRows], __builtin_isnan(tn[k]) ? 0x80000000u : lane_off,
                                         ibase[k / kPairRows] + (unsigned)(k % kPairRows) * row_step);
            // Branch-free folds: every lane computes one, and lanes without an update
            // store at the out-of-range offset (no access). nfb bit k: the fold left a
            // non-far voxel.
            unsigned nfb = 0u;
#pragma unroll
            for (int k = 0; k < 2 * kPairRows; ++k) {
                const bool upd = !__builtin_isnan(tn[k]);
                const float w_old = S::weight(wd[k]);
                const float t_old = S::tsdf(wd[k]);
                const float t_new = (t_old * w_old + tn[k]) / (w_old + 1.0f);
                BufWord<S>::store(S::pack(t_new, smin(w_old + 1.0f, p.max_w)), rs[k / kPairRows],
                                  upd ? lane_off : 0x80000000u, ibase[k / kPairRows] + (unsigned)(k % kPairRows) * row_step);
                n_upd += upd ? 1u : 0u;
                n_new += (upd && w_old == 0.0f) ? 1u : 0u;
                nfb |= (upd && !S::farPacked(t_new)) ? (1u << k) : 0u;  // the new weight is >= 1
            }
            if (bricks) {
                constexpr unsigned kItemBits = (1u << kPairRows) - 1u;
                const unsigned long long near0 = __ballot((nfb & kItemBits) != 0u);
                const unsigned long long near1 = __ballot((nfb >> kPairRows) != 0u);
                const unsigned long long first0 = __ballot(g == 0 && (nfb & 1u));
                const unsigned long long first1 = __ballot(g == 0 && ((nfb >> kPairRows) & 1u));
                if (lane == 0) {
                    const unsigned c0 = (unsigned)(near0 | (near0 >> 32)), c1 = (unsigned)(near1 | (near1 >> 32));
                    if (c0) markItem(bricks, nb, ixt[0], iy[0], izb[0], c0, (unsigned)first0);
                    if (c1) markItem(bricks, nb, ixt[1], iy[1], izb[1], c1, (unsigned)first1);
                }
            }
        }
        waveSync();
    }
    if (counters) {
        unsigned long long a = n_upd, b = n_new;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        flushCounts(counters, a, b);
    }
}

// ---------------------------------------------------------------------------
// Fast kernel (the opt-in FAST mode, either storage form). The same
// tiles and classes as integrateExact (cull / free / per-voxel), but a voxel's camera
// coordinate is the affine c0 + x ax + z az evaluated with fmaf and the reciprocals
// are the hardware's, so the lanes of a wave are independent voxels: lane = x column
// of a 64-column tile, the wave walks kFastRows z rows of it at one y. Free tiles fold
// tsdf_new = 1 with no projection. tileClass's margins cover the serial-add rounding
// of the exact kernel, which bounds this kernel's few-ulp affine rounding as well, so
// every class is the outcome the per-voxel fast path would reach.
// ---------------------------------------------------------------------------
constexpr int kFastRows = 16;  // z rows per tile
constexpr int kFastBatch = 8;  // rows whose loads are issued together

// Brick-map bytes for one row of a 64-column tile at (y, z): mask bit j = column xc + j
// left a non-far voxel. Column 16 k is also the last layer of x-brick xc / 16 + k - 1;
// y = 16 q and z = 16 q are also in the y- and z-bricks below.
__device__ __forceinline__ void markRow(unsigned char* __restrict__ bricks, int nb, int xc, int y, int z,
                                        unsigned long long mask) {
    const int by = y / kBrick, bz = z / kBrick, bx0 = xc / kBrick;
    const int ny = (y % kBrick == 0 && by > 0) ? 2 : 1, nz = (z % kBrick == 0 && bz > 0) ? 2 : 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned sub = (unsigned)(mask >> (16 * k)) & 0xffffu;
        if (!sub) continue;
        const int bx = bx0 + k;
        const int nx = ((sub & 1u) && bx > 0) ? 2 : 1;
        for (int kz = 0; kz < nz; ++kz)
            for (int ky = 0; ky < ny; ++ky)
                for (int kx = 0; kx < nx; ++kx)
                    if (bx - kx < nb) bricks[((size_t)(bz - kz) * nb + (by - ky)) * nb + (bx - kx)] = 1;
    }
}

template <class S>
__global__ __launch_bounds__(256) void integrateFast(typename S::Word* __restrict__ vox, VolParams p, Intr K, int W,
                                                     int H, const float* __restrict__ depth, float min_d, float max_d,
                                                     PoseArg pa, DepthBounds bnd,
                                                     unsigned long long* __restrict__ counters,
                                                     unsigned char* __restrict__ bricks, int nb) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const IntFrame f = integrateSetup(p, pa, K, W, H, max_d);
    if (!f.ok) return;
    const int ny = f.y1 - f.y0 + 1;
    const int zbase = f.z0 & ~(kFastRows - 1);
    const int nzg = (f.z1 - zbase + kFastRows) / kFastRows;
    const int task = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (task >= ny * nzg) return;  // whole wave; no block barrier below
    const int y = f.y0 + task % ny;
    const int zb = zbase + (task / ny) * kFastRows;
    const int za = max(zb, f.z0), ze = min(zb + kFastRows - 1, f.z1);
    const float vs = p.vs, trunc = p.trunc;
    const F3 ax = f.step;
    const F3 az{f.R_wc.m[0][2] * vs, f.R_wc.m[1][2] * vs, f.R_wc.m[2][2] * vs};
    const F3 c0 = f3add(mulv(f.R_wc, F3{p.origin[0], p.origin[1] + (float)y * vs, p.origin[2]}), f.t_wc);
    const float cu = K.cx + 0.5f, cv = K.cy + 0.5f;
    // u = (fx c.x + cu c.z) / c.z, affine numerators in x
    const float su = fmaf(K.fx, ax.x, cu * ax.z), sv = fmaf(K.fy, ax.y, cv * ax.z);
    const float inv_trunc = __builtin_amdgcn_rcpf(trunc);
    const float fW = (float)W, fH = (float)H;
    const size_t res = (size_t)p.res, plane = res * res;
    const float cbound = fabsf(f.t_wc.x) + fabsf(f.t_wc.y) + fabsf(f.t_wc.z) + fabsf(p.origin[0]) +
                         fabsf(p.origin[1]) + fabsf(p.origin[2]) + 3.0f * (float)p.res * vs;
    unsigned long long n_upd = 0, n_new = 0;  // wave-uniform (ballot counts)

    const int xstart = f.x0 & ~63;
    unsigned long long cls_lo, cls_hi;
    {
        int my = kTileCull;
        const int xt = xstart + lane * 64;
        if (xt <= f.x1) {
            const int js = max(f.x0 - xt, 0), n = min(64, f.x1 - xt + 1);
            my = tileClass(f, p, K, W, H, max_d, xt + js, xt + n - 1, y, za, ze, bnd, cbound);
        }
        cls_lo = __ballot((my & 1) != 0);
        cls_hi = __ballot((my & 2) != 0);
    }
    for (int xc = xstart, ti = 0; xc <= f.x1; xc += 64, ++ti) {
        const int cls = (int)((cls_lo >> ti) & 1ull) | ((int)((cls_hi >> ti) & 1ull) << 1);
        if (cls == kTileCull) continue;
        const int x = xc + lane;
        const bool col_ok = x >= f.x0 && x <= f.x1;
        const float xf = (float)x;
        const size_t tile0 = (size_t)za * plane + (size_t)y * res + (size_t)xc;
        const size_t left = (res * plane - tile0) * sizeof(typename S::Word);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            vox + tile0, (short)0, (int)(left < 0x80000000ull ? left : 0x80000000ull), 0x00020000);
        const unsigned lane_off = (unsigned)lane * (unsigned)sizeof(typename S::Word);
        const unsigned row_step = (unsigned)(plane * sizeof(typename S::Word));
#pragma unroll 1
        for (int z0 = za; z0 <= ze; z0 += kFastBatch) {
            float tn[kFastBatch];
            if (cls == kTileFree) {
#pragma unroll
                for (int k = 0; k < kFastBatch; ++k) tn[k] = (col_ok && z0 + k <= ze) ? 1.0f : __builtin_nanf("");
            } else {
                float cz[kFastBatch], dj[kFastBatch];
                int pix[kFastBatch];
#pragma unroll
                for (int k = 0; k < kFastBatch; ++k) {
                    pix[k] = -1;
                    const float zf = (float)(z0 + k);
                    const F3 cr{fmaf(zf, az.x, c0.x), fmaf(zf, az.y, c0.y), fmaf(zf, az.z, c0.z)};
                    cz[k] = fmaf(xf, ax.z, cr.z);
                    if (col_ok && z0 + k <= ze && cz[k] > 0.0f) {
                        const float bu = fmaf(K.fx, cr.x, cu * cr.z), bv = fmaf(K.fy, cr.y, cv * cr.z);
                        const float rz = __builtin_amdgcn_rcpf(cz[k]);
                        const float uf = floorf(fmaf(xf, su, bu) * rz);
                        const float vf = floorf(fmaf(xf, sv, bv) * rz);
                        if (uf >= 0.0f && uf < fW && vf >= 0.0f && vf < fH) pix[k] = (int)vf * W + (int)uf;
                    }
                }
#pragma unroll
                for (int k = 0; k < kFastBatch; ++k) dj[k] = depth[pix[k] < 0 ? 0 : pix[k]];
#pragma unroll
                for (int k = 0; k < kFastBatch; ++k) {
                    tn[k] = __builtin_nanf("");
                    const float d = dj[k];
                    if (pix[k] >= 0 && d >= min_d && d <= max_d) {
                        const float sdf = d - cz[k];
                        if (!(sdf < -trunc)) tn[k] = fminf(1.0f, sdf * inv_trunc);
                    }
                }
            }
            typename S::Word wd[kFastBatch];
            const unsigned soff0 = (unsigned)(z0 - za) * row_step;
#pragma unroll
            for (int k = 0; k < kFastBatch; ++k)
                wd[k] = BufWord<S>::load(rs, __builtin_isnan(tn[k]) ? 0x80000000u : lane_off, soff0 + k * row_step);
#pragma unroll
            for (int k = 0; k < kFastBatch; ++k) {
                const bool upd = !__builtin_isnan(tn[k]);
                bool nf = false, newly = false;
                if (upd) {
                    const float w_old = S::weight(wd[k]);
                    const float t_old = S::tsdf(wd[k]);
                    newly = w_old == 0.0f;
                    const float t_new = fmaf(t_old, w_old, tn[k]) * __builtin_amdgcn_rcpf(w_old + 1.0f);
                    const typename S::Word nw = S::pack(t_new, fminf(w_old + 1.0f, p.max_w));
                    BufWord<S>::store(nw, rs, lane_off, soff0 + k * row_step);
                    nf = !S::far(nw);
                }
                n_upd += __builtin_popcountll(__ballot(upd));
                n_new += __builtin_popcountll(__ballot(newly));
                if (bricks) {
                    const unsigned long long m = __ballot(nf);
                    if (m && lane == 0) markRow(bricks, nb, xc, y, z0 + k, m);
                }
            }
        }
    }
    if (counters) flushCounts(counters, n_upd, n_new);
}

// Exhaustive comparison of a fast form with the IEEE result over every float whose
// magnitude's bit pattern lies in [lo, hi), both signs; counts the mismatches (one
// atomic per wave that found any). mode 0: divFixed(x) vs x / b; mode 1: recipNR(x) vs
// 1.0f / x.
__global__ void verifyKernel(int mode, float b, float r, unsigned lo, unsigned hi, unsigned* __restrict__ bad) {
    const FixedDiv d{b, r, 1};
    const unsigned long long n = 2ull * (unsigned long long)(hi - lo);
    unsigned miss = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned bits = (lo + (unsigned)(i >> 1)) | ((unsigned)(i & 1) << 31);
        const float x = __uint_as_float(bits);
        const float fast = mode == 0 ? divFixed(d, x) : recipNR(x);
        const float ieee = mode == 0 ? x / b : 1.0f / x;
        miss += !(fast == ieee);
    }
    const unsigned long long any = __ballot(miss != 0);
    if (any) {
        unsigned total = miss;
        for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(bad, total);
    }
}

static unsigned runCheck(int mode, float b, float r, unsigned lo, unsigned hi) {
    unsigned* bad = nullptr;
    unsigned h = 1;
    if (hipMalloc(&bad, sizeof(unsigned)) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    bool ok = hipMemset(bad, 0, sizeof(unsigned)) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(verifyKernel, dim3(8192), dim3(256), 0, nullptr, mode, b, r, lo, hi, bad);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(&h, bad, sizeof(unsigned), hipMemcpyDeviceToHost) == hipSuccess;
    }
    (void)hipFree(bad);
    (void)hipGetLastError();
    return ok ? h : 1;
}

bool recipVerified() {
    static std::mutex mu;
    static std::map<int, bool> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) return it->second;
    // 2^-126 <= |x| < 2^64: bit patterns [0x00800000, 0x5F800000)
    const bool ok = runCheck(1, 0.0f, 0.0f, 0x00800000u, 0x5F800000u) == 0;
    known[dev] = ok;
    return ok;
}

FixedDiv makeFixedDiv(float b) {
    FixedDiv d{b, 1.0f / b, 0};
    if (!(b > 0.0f) || !std::isfinite(d.r)) return d;
    // 2^-100 <= |x| < 2^32: bit patterns [0x0D800000, 0x4F800000)
    d.fast = runCheck(0, b, d.r, 0x0D800000u, 0x4F800000u) == 0;
    return d;
}

// Reset to the empty state.
template <class S>
__global__ void fillEmpty(typename S::Word* __restrict__ vox, size_t n) {
    const typename S::Word e = S::empty();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        vox[i] = e;
}

__global__ void clearBytes(unsigned char* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// The brick map of the whole volume from its current contents (after clearBytes).
template <class S>
__global__ void brickScan(const typename S::Word* __restrict__ vox, int res, unsigned char* __restrict__ bricks,
                          int nb) {
    const size_t n = (size_t)res * res * res;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (S::far(vox[i])) continue;
        const int x = (int)(i % (size_t)res), y = (int)((i / (size_t)res) % (size_t)res),
                  z = (int)(i / ((size_t)res * res));
        markBricks(bricks, nb, x, y, z);
    }
}

// ---------------------------------------------------------------------------
// Host launchers (called from capi.hip).
// ---------------------------------------------------------------------------
// Bytes of the depth bounds of a W x H frame.
size_t depthBoundsBytes(int W, int H) {
    size_t n = 0;
    for (int l = 1; l <= kBoundLevels; ++l)
        n += (size_t)((W + (1 << l) - 1) >> l) * (size_t)((H + (1 << l) - 1) >> l);
    return n * sizeof(float2);
}

// dtr: division by the truncation, verified exact (dtr.fast) or IEEE; recip: 1 / c.z by
// the verified Newton-Raphson reciprocal; bounds: scratch for the depth bounds
// (depthBoundsBytes), or null to decide every tile voxel by voxel.
// The depth-bounds pyramid of the frame into `bounds` (null: no bounds, every
// in-frustum tile is decided voxel by voxel).
static DepthBounds boundsFor(void* bounds, const float* depth, int W, int H, float min_d, float max_d,
                             hipStream_t st, IntShared* sh = nullptr, const VolParams& p = VolParams{},
                             const PoseArg& pa = PoseArg{}, Intr K = Intr{}) {
    DepthBounds bnd{};
    if (!bounds) return bnd;
    int off = 0;
    for (int l = 1; l <= kBoundLevels; ++l) {
        bnd.off[l] = off;
        bnd.w[l] = (W + (1 << l) - 1) >> l;
        bnd.h[l] = (H + (1 << l) - 1) >> l;
        off += bnd.w[l] * bnd.h[l];
    }
    bnd.lvl = static_cast<const float2*>(bounds);
    // (with sh: one extra row of blocks, whose first block writes the frame record)
    hipLaunchKernelGGL(depthBoundsKernel, dim3((W + 63) / 64, (H + 63) / 64 + (sh ? 1 : 0)), dim3(256), 0, st, depth, W,
                       H, min_d, max_d, static_cast<float2*>(bounds), bnd, sh, p, pa, K);
    return bnd;
}

// bricks: the brick map to keep (brickCount(res)^3 bytes), or null. classes: the item
// pipeline's per-frame scratch (exactScratchBytes(res)), or null to run integrateExact,
// which classifies and carries inside the kernel (needs no scratch besides the bounds).
hipError_t launchIntegrate(int format, bool exact, void* vox, const VolParams& p, Intr K, int W, int H,
                           const float* depth, float min_d, float max_d, const PoseArg& pa, const FixedDiv& dtr,
                           bool recip, void* bounds, unsigned long long* counters, unsigned char* bricks,
                           unsigned char* classes, hipStream_t st) {
    const int nb = brickCount(p.res);
    if (exact) {
        // sdf = d - c.z with d >= min_d: when min_d >= 2^-60 a nonzero sdf cannot fall below
        // 2^-100 and the unguarded fixed-divisor form runs.
        const bool fd = dtr.fast != 0, rc = recip, far = min_d >= 0x1p-60f;
        const int W_ = W;
        if (classes && bounds) {
            const ItemScratch is = itemScratch(classes, p.res);
            const DepthBounds bnd = boundsFor(bounds, depth, W, H, min_d, max_d, st, is.sh, p, pa, K);
            hipLaunchKernelGGL(subtileClassKernel, dim3(classBlocks(p.res)), dim3(256), 0, st, p, K, W, H, max_d, bnd,
                               is);
            // every row of the volume (rows outside the frame's box exit at once)
            const long long rows = (long long)p.res * p.res;
            hipLaunchKernelGGL(itemCarryKernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, p, is);
            const bool one = (size_t)p.res * p.res * p.res * (format == kFmtF32 ? 8 : 4) <= ((size_t)1 << 31);
#define AZU_INT_ITEMS(S, W, FD, RC)                                                                                   \
    do {                                                                                                              \
        if (one)                                                                                                      \
            hipLaunchKernelGGL((integrateItems<S, FD, RC, true>), dim3(kItemBlocks), dim3(256), 0, st, (W*)vox, p, K, \
                               W_, H, depth, min_d, max_d, dtr, is, counters, bricks, nb);                            \
        else                                                                                                          \
            hipLaunchKernelGGL((integrateItems<S, FD, RC, false>), dim3(kItemBlocks), dim3(256), 0, st, (W*)vox, p,   \
                               K, W_, H, depth, min_d, max_d, dtr, is, counters, bricks, nb);                         \
    } while (0)
            if (format == kFmtF32) {
                if (fd && far && rc) AZU_INT_ITEMS(StoreF32, float2, 2, true);
                else if (fd && rc) AZU_INT_ITEMS(StoreF32, float2, 1, true);
                else if (rc) AZU_INT_ITEMS(StoreF32, float2, 0, true);
                else AZU_INT_ITEMS(StoreF32, float2, 0, false);
            } else {
                if (fd && far && rc) AZU_INT_ITEMS(StoreF16, uint32_t, 2, true);
                else if (fd && rc) AZU_INT_ITEMS(StoreF16, uint32_t, 1, true);
                else if (rc) AZU_INT_ITEMS(StoreF16, uint32_t, 0, true);
                else AZU_INT_ITEMS(StoreF16, uint32_t, 0, false);
            }
#undef AZU_INT_ITEMS
            return hipGetLastError();
        }
        const DepthBounds bnd = boundsFor(bounds, depth, W, H, min_d, max_d, st);
        // One wave per (y, kTileRows-row z group, row segment) of the whole volume; waves
        // outside the frustum box exit at once (the box is only known on the device).
        const long long tasks =
            (long long)p.res * ((p.res + kTileRows - 1) / kTileRows + 1) * ((p.res + kSegCols - 1) / kSegCols + 1);
        const dim3 grid((unsigned)((tasks + 3) / 4)), block(256);
#define AZU_INT_EXACT(S, W, FD, RC)                                                                        \
    hipLaunchKernelGGL((integrateExact<S, FD, RC>), grid, block, 0, st, (W*)vox, p, K, W_, H, depth, min_d, \
                       max_d, pa, dtr, bnd, counters, bricks, nb)
        if (format == kFmtF32) {
            if (fd && far && rc) AZU_INT_EXACT(StoreF32, float2, 2, true);
            else if (fd && rc) AZU_INT_EXACT(StoreF32, float2, 1, true);
            else if (rc) AZU_INT_EXACT(StoreF32, float2, 0, true);
            else AZU_INT_EXACT(StoreF32, float2, 0, false);
        } else {
            if (fd && far && rc) AZU_INT_EXACT(StoreF16, uint32_t, 2, true);
            else if (fd && rc) AZU_INT_EXACT(StoreF16, uint32_t, 1, true);
            else if (rc) AZU_INT_EXACT(StoreF16, uint32_t, 0, true);
            else AZU_INT_EXACT(StoreF16, uint32_t, 0, false);
        }
#undef AZU_INT_EXACT
    } else {
        const DepthBounds bnd = boundsFor(bounds, depth, W, H, min_d, max_d, st);
        // One wave per (y, kFastRows-row z group) of the whole volume; waves outside the
        // frustum box exit at once.
        const long long tasks = (long long)p.res * ((p.res + kFastRows - 1) / kFastRows + 1);
        const dim3 grid((unsigned)((tasks + 3) / 4));
        if (format == kFmtF32)
            hipLaunchKernelGGL(integrateFast<StoreF32>, grid, dim3(256), 0, st, (float2*)vox, p, K, W, H, depth,
                               min_d, max_d, pa, bnd, counters, bricks, nb);
        else
            hipLaunchKernelGGL(integrateFast<StoreF16>, grid, dim3(256), 0, st, (uint32_t*)vox, p, K, W, H, depth,
                               min_d, max_d, pa, bnd, counters, bricks, nb);
    }
    return hipGetLastError();
}

// Brick map of res^3 voxels: cleared (all far, the empty volume's map), then, when
// scan is set, marked from the voxels' current contents.
hipError_t launchBrickMap(int format, const void* vox, int res, unsigned char* bricks, bool scan, hipStream_t st) {
    const int nb = brickCount(res);
    const size_t nbytes = (size_t)nb * nb * nb;
    hipLaunchKernelGGL(clearBytes, dim3((unsigned)std::min<size_t>((nbytes + 255) / 256, 4096)), dim3(256), 0, st,
                       bricks, nbytes);
    if (scan) {
        const size_t n = (size_t)res * res * res;
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
        if (format == kFmtF32)
            hipLaunchKernelGGL(brickScan<StoreF32>, dim3(blocks), dim3(256), 0, st, (const float2*)vox, res, bricks,
                               nb);
        else
            hipLaunchKernelGGL(brickScan<StoreF16>, dim3(blocks), dim3(256), 0, st, (const uint32_t*)vox, res,
                               bricks, nb);
    }
    return hipGetLastError();
}

hipError_t launchFillEmpty(int format, void* vox, size_t n, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    if (format == kFmtF32)
        hipLaunchKernelGGL(fillEmpty<StoreF32>, dim3(blocks), dim3(256), 0, st, (float2*)vox, n);
    else
        hipLaunchKernelGGL(fillEmpty<StoreF16>, dim3(blocks), dim3(256), 0, st, (uint32_t*)vox, n);
    return hipGetLastError();
}

}  // namespace azu

namespace azu {

// Whole-volume conversion between the storage words and float tsdf / weight arrays
// (download / upload).
template <class S>
__global__ void volConvert(typename S::Word* __restrict__ vox, float* __restrict__ tsdf, float* __restrict__ weight,
                           size_t n, int to_float) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (to_float) {
            const typename S::Word w = vox[i];
            tsdf[i] = S::tsdf(w);
            weight[i] = S::weight(w);
        } else {
            vox[i] = S::pack(tsdf[i], weight[i]);
        }
    }
}

// Voxels with weight > 0 (TSDFVolume::usageFraction's count), one wave sum per 64
// voxels and one atomic per block into *out (cleared by the caller).
template <class S>
__global__ __launch_bounds__(256) void countObserved(const typename S::Word* __restrict__ vox, size_t n,
                                                     unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c += S::weight(vox[i]) > 0.0f ? 1u : 0u;
    __shared__ unsigned long long part[4];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

hipError_t launchCountObserved(int format, const void* vox, size_t n, unsigned long long* out, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (format == kFmtF32)
        hipLaunchKernelGGL(countObserved<StoreF32>, dim3(blocks), dim3(256), 0, st, (const float2*)vox, n, out);
    else
        hipLaunchKernelGGL(countObserved<StoreF16>, dim3(blocks), dim3(256), 0, st, (const uint32_t*)vox, n, out);
    return hipGetLastError();
}

hipError_t launchVolConvert(int format, bool to_float, void* vox, float* tsdf, float* weight, size_t n,
                            hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    if (format == kFmtF32)
        hipLaunchKernelGGL(volConvert<StoreF32>, dim3(blocks), dim3(256), 0, st, (float2*)vox, tsdf, weight, n,
                           to_float ? 1 : 0);
    else
        hipLaunchKernelGGL(volConvert<StoreF16>, dim3(blocks), dim3(256), 0, st, (uint32_t*)vox, tsdf, weight, n,
                           to_float ? 1 : 0);
    return hipGetLastError();
}

}  // namespace azu
