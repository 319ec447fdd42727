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

struct StoreF32 {
    using Word = float2;
    static constexpr int kBytes = 8;
    __device__ __forceinline__ static float tsdf(Word w) { return w.x; }
    __device__ __forceinline__ static float weight(Word w) { return w.y; }
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
    __device__ __forceinline__ static Word pack(float t, float wt) {
        const uint32_t h = __half_as_ushort(__float2half_rn(t));
        return h | (static_cast<uint32_t>(wt) << 16);
    }
    __host__ __device__ static Word empty() { return 0x3C00u; }  // half(1.0), weight 0
};

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
