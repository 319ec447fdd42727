// Shared device-side definitions for the gfx950 KinectFusion kernels.
//
// Numeric policy. The whole library is compiled with -ffp-contract=off and IEEE-exact
// fp32 division and square root, so the "exact" kernels (the FP32 parity path) run the
// same float operation sequence as the CPU restatement of the reference: three-term
// sums as a0 + (a1 + a2), products and quotients exactly where the reference has them.
// The "fast" kernels (the FP16 + UInt16 storage form) use hardware reciprocals and
// explicit fmaf; their accuracy is stated against the FP64 oracle instead.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace azu {

// ---------------------------------------------------------------------------
// Parameters (layout mirrors include/azu_hip.h's azu_tsdf_params).
// ---------------------------------------------------------------------------
struct VolParams {
    int res;
    float vs, trunc, max_w;
    float origin[3];
    float min_depth, max_depth;
};

struct Intr {
    float fx, fy, cx, cy;
};

// A pose handed to a kernel either by value (host caller) or by device pointer
// (chained on the device, e.g. the ICP output). Column-major 4x4, world-from-camera.
struct PoseArg {
    float m[16];
    const float* dptr;
};

__device__ __forceinline__ float poseEl(const PoseArg& p, int i) { return p.dptr ? p.dptr[i] : p.m[i]; }

// ---------------------------------------------------------------------------
// Voxel storage forms.
//   F32:     float2 {tsdf, weight}, 8 B per voxel.
//   F16U16:  uint32 {half tsdf (bits 0-15), uint16 weight (bits 16-31)}, 4 B per voxel.
// Both hold the empty state tsdf = 1, weight = 0.
// ---------------------------------------------------------------------------
enum Format : int { kFmtF32 = 0, kFmtF16U16 = 1 };

// ---------------------------------------------------------------------------
// Brick map (the raycast's empty-space skipping). The volume is cut into bricks of
// kBrick^3 voxels; brick b (per axis) spans the voxels [kBrick b, kBrick b + kBrick],
// one layer shared with the next brick, so every trilinear cell with its lower corner
// in [kBrick b, kBrick b + kBrick - 1] has all eight corners inside one brick's span.
// A brick's byte is 0 only while every voxel of its span is "far": unobserved, or
// observed with tsdf in [kFarTsdf, 1.5]. A sample whose corners are all far is either
// unobserved (invalid) or finite with value >= kFarTsdf (1 - 8 2^-24) > 0.999f (three
// levels of linear interpolation, each losing at most 2.5 float roundings of a convex
// combination of values >= kFarTsdf); either way the march takes its far step and
// neither ends nor hits there. Writers mark bricks (byte = 1) for every voxel they leave
// non-far, so the map only ever over-approximates the non-far voxels.
// ---------------------------------------------------------------------------
constexpr int kBrick = 16;
constexpr float kFarTsdf = 0.999005f;  // <= half(0.99902344), > 0.999f (1 + 5e-6)

__host__ __device__ inline int brickCount(int res) { return (res + kBrick - 1) / kBrick; }

// Marks, in z-brick bz, every brick whose x and y spans hold voxel column (x, y): brick
// v / kBrick on each axis, and the one below it when v is that brick's shared first layer.
__device__ __forceinline__ void markBricksXY(unsigned char* __restrict__ bricks, int nb, int x, int y, int bz) {
    const int bx = x / kBrick, by = y / kBrick;
    const int nx = (x % kBrick == 0 && bx > 0) ? 2 : 1, ny = (y % kBrick == 0 && by > 0) ? 2 : 1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) bricks[((size_t)bz * nb + (by - j)) * nb + (bx - i)] = 1;
}

// Marks every brick whose span holds voxel (x, y, z).
__device__ __forceinline__ void markBricks(unsigned char* __restrict__ bricks, int nb, int x, int y, int z) {
    const int bz = z / kBrick;
    markBricksXY(bricks, nb, x, y, bz);
    if (z % kBrick == 0 && bz > 0) markBricksXY(bricks, nb, x, y, bz - 1);
}

struct StoreF32 {
    using Word = float2;
    static constexpr int kBytes = 8;
    __device__ __forceinline__ static float tsdf(Word w) { return w.x; }
    __device__ __forceinline__ static float weight(Word w) { return w.y; }
    // every weight > 0 (a NaN weight fails, as in the reference's !(w > 0) test)
    __device__ __forceinline__ static bool observed8(Word a, Word b, Word c, Word d, Word e, Word f, Word g,
                                                     Word h) {
        return (a.y > 0.0f) & (b.y > 0.0f) & (c.y > 0.0f) & (d.y > 0.0f) & (e.y > 0.0f) & (f.y > 0.0f) &
               (g.y > 0.0f) & (h.y > 0.0f);
    }
    // far for the brick map: unobserved (as in observed8), or tsdf in [kFarTsdf, 1.5]
    __device__ __forceinline__ static bool far(Word w) {
        return !(w.y > 0.0f) || (w.x >= kFarTsdf && w.x <= 1.5f);
    }
    // far for a word packed from tsdf t with a weight >= 1 (what integrate writes)
    __device__ __forceinline__ static bool farPacked(float t) { return t >= kFarTsdf && t <= 1.5f; }
    __device__ __forceinline__ static Word pack(float t, float wt) { return make_float2(t, wt); }
    __host__ __device__ static Word empty() { return float2{1.0f, 0.0f}; }
};

struct StoreF16 {
    using Word = uint32_t;
    static constexpr int kBytes = 4;
    __device__ __forceinline__ static float tsdf(Word w) {
        return __half2float(__ushort_as_half(static_cast<unsigned short>(w & 0xffffu)));
    }
    __device__ __forceinline__ static float weight(Word w) { return static_cast<float>(w >> 16); }
    // every weight > 0: the weight is the high half, so the smallest word is >= 2^16
    __device__ __forceinline__ static bool observed8(Word a, Word b, Word c, Word d, Word e, Word f, Word g,
                                                     Word h) {
        return min(min(min(a, b), min(c, d)), min(min(e, f), min(g, h))) >= 0x10000u;
    }
    __device__ __forceinline__ static bool far(Word w) {
        const float t = tsdf(w);
        return w < 0x10000u || (t >= kFarTsdf && t <= 1.5f);
    }
    // far for a word packed from tsdf t with a weight >= 1: binary16 rounding (to
    // nearest, ties to even) lands in [0.99902344, 1.5] exactly for t in
    // [0.99877930, 1.50048828] (the midpoints, whose ties go to the even ends inside)
    __device__ __forceinline__ static bool farPacked(float t) { return t >= 0.99877930f && t <= 1.50048828f; }
    __device__ __forceinline__ static Word pack(float t, float wt) {
        const uint32_t h = __half_as_ushort(__float2half_rn(t));
        return h | (static_cast<uint32_t>(wt) << 16);
    }
    __host__ __device__ static Word empty() { return 0x3C00u; }  // half(1.0), weight 0
};

// ---------------------------------------------------------------------------
// Division by a divisor fixed for a whole launch (the voxel size, the truncation).
// divFixed computes q = x r (r = 1/b rounded to float) and corrects it once with the
// exact residual x - q b (Markstein's form; whether it gives the correctly rounded
// quotient for every x depends on b, and it fails where the quotient underflows). It
// replaces the IEEE quotient x / b only after verifyKernel has compared the two on the
// device for every float with 2^-100 <= |x| < 2^32 and found them equal (for the voxel
// sizes and truncations 1, 2, 3, 4, 6 and 12 mm it holds). Smaller |x| takes the IEEE division; a
// larger |x| only arises for a sample or distance that both forms reject or clamp (far
// outside the volume, or sdf / trunc > 1 folded as 1).
// ---------------------------------------------------------------------------
struct FixedDiv {
    float b, r;
    int fast;  // verified on this device
};

__device__ __forceinline__ float divFixed(const FixedDiv& d, float x) {
    const float q0 = x * d.r;
    return fmaf(fmaf(-q0, d.b, x), d.r, q0);
}

// FD = 0: IEEE division. FD = 1: divFixed where verified (|x| >= 2^-100), IEEE below.
// FD = 2: divFixed for every x, for callers whose nonzero x cannot fall below 2^-100:
// for x = a - b of two floats with |b| >= 2^-60, |x| < 2^-100 needs |a| > 2^-61, so both
// are multiples of 2^-84 and so is the (exact, by Sterbenz) difference: x = +0 or
// |x| >= 2^-84. divFixed(+0) = +0 = +0 / b.
template <int FD>
__device__ __forceinline__ float divBy(const FixedDiv& d, float x) {
    if (FD == 0) return x / d.b;
    if (FD == 2) return divFixed(d, x);
    return fabsf(x) >= 0x1p-100f ? divFixed(d, x) : x / d.b;
}

// Host: the FixedDiv for divisor b, verified on the current device (fast = 0 when the
// check fails or cannot run).
FixedDiv makeFixedDiv(float b);

// 1 / x from the hardware reciprocal estimate and two Newton-Raphson steps with fused
// multiply-adds. For 2^-126 <= |x| < 2^64 it replaces the IEEE 1.0f / x once
// verifyRecipKernel has found the two equal over that whole range on this device
// (recipVerified); outside it (subnormal x, or results that underflow) the IEEE
// division is used.
__device__ __forceinline__ float recipNR(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    r = fmaf(fmaf(-x, r, 1.0f), r, r);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}

template <bool RC>
__device__ __forceinline__ float recipBy(float x) {
    if (!RC) return 1.0f / x;
    const float a = fabsf(x);
    return (a >= 0x1p-126f && a < 0x1p64f) ? recipNR(x) : 1.0f / x;
}

// Host: whether recipNR equals 1.0f / x for every float 2^-126 <= |x| < 2^64 on the
// current device (checked once per device and remembered).
bool recipVerified();

// ---------------------------------------------------------------------------
// Small math in the fixed evaluation order.
// ---------------------------------------------------------------------------
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ float sum3(float a0, float a1, float a2) { return a0 + (a1 + a2); }
__device__ __forceinline__ F3 f3add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 f3mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 f3div(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float f3dot(F3 a, F3 b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float f3norm(F3 a) { return sqrtf(f3dot(a, a)); }
__device__ __forceinline__ F3 f3cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool f3finite(F3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
// std::min / std::max semantics
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

struct M33 {
    float m[3][3];
};
__device__ __forceinline__ F3 mulv(const M33& R, F3 v) {
    return {sum3(R.m[0][0] * v.x, R.m[0][1] * v.y, R.m[0][2] * v.z),
            sum3(R.m[1][0] * v.x, R.m[1][1] * v.y, R.m[1][2] * v.z),
            sum3(R.m[2][0] * v.x, R.m[2][1] * v.y, R.m[2][2] * v.z)};
}
__device__ __forceinline__ F3 mulvT(const M33& R, F3 v) {
    return {sum3(R.m[0][0] * v.x, R.m[1][0] * v.y, R.m[2][0] * v.z),
            sum3(R.m[0][1] * v.x, R.m[1][1] * v.y, R.m[2][1] * v.z),
            sum3(R.m[0][2] * v.x, R.m[1][2] * v.y, R.m[2][2] * v.z)};
}
__device__ __forceinline__ M33 mulm(const M33& A, const M33& B) {
    M33 C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C.m[i][j] = sum3(A.m[i][0] * B.m[0][j], A.m[i][1] * B.m[1][j], A.m[i][2] * B.m[2][j]);
    return C;
}

struct Rigid {
    M33 R;
    F3 t;
};
__device__ __forceinline__ Rigid rigidFrom(const PoseArg& p) {
    Rigid P;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) P.R.m[r][c] = poseEl(p, c * 4 + r);
    P.t = {poseEl(p, 12), poseEl(p, 13), poseEl(p, 14)};
    return P;
}
__device__ __forceinline__ Rigid rigidFromPtr(const float* m) {
    Rigid P;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) P.R.m[r][c] = m[c * 4 + r];
    P.t = {m[12], m[13], m[14]};
    return P;
}
// Rigid inverse R^T, -(R^T t) (the CPU path's pose inverse for rigid poses).
__device__ __forceinline__ Rigid rigidInverse(const Rigid& P) {
    Rigid I;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) I.R.m[r][c] = P.R.m[c][r];
    const F3 rt = mulv(I.R, P.t);
    I.t = {-rt.x, -rt.y, -rt.z};
    return I;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace azu
