// Depth conditioning of raw Kinect v1 frames on gfx950: the depth half of the
// reference's SignalConditioner (src/sensor/SignalConditioner_omp.cpp:337-716):
//
//   condMedian    denoiseDepthSpatial (:442-478)   3x3 median of the usable samples
//   condFillHoles fillDepthHoles      (:578-635)   raw-0 pixels <- inverse-distance mean
//   condGuided    guidedDepthFilter   (:637-716)   9x9 luma/depth/spatial-weighted mean
//   condEma       applyDepthEma       (:494-576)   per-pixel temporal blend with resets
//
// Every stage reads a source frame and writes a separate destination (the reference's
// scratch-buffer split), one lane per pixel, borders reflected as cpuReflectCoord
// (include/sensor/BorderMode.h:42-46). All arithmetic is the reference's float sequence
// (compiled without contraction, IEEE division and square root), so the codes are the
// reference's bit for bit, with one stated exception: the guided filter's two
// data-dependent Gaussian weights are the double exponential rounded to float (the
// correctly rounded expf, not one library's approximation of it). The spatial Gaussian
// is evaluated the same way; for its 33 arguments -d^2 / 32 every faithful expf agrees.
//
// The guided filter is the only stage with real work (81 taps, two exponentials a tap):
// its blocks stage the usable metres, the codes and the luma of a (32 + 8) x (8 + 8)
// tile in LDS, so each sample is decoded once per block instead of 81 times.
#include "common.hpp"

namespace azu {

namespace {

constexpr float kCurveA = -0.0030711016f;  // DepthValidity.h:53-54
constexpr float kCurveB = 3.3309495161f;
constexpr int kGuidedR = 4;    // SignalConditioner_omp.cpp:33
constexpr int kHoleR = 2;      // :30
constexpr float kEmaJump = 0.05f;  // :41
constexpr float kSigmaLuma = 0.10f;
constexpr float kSigmaDepth = 0.04f;

constexpr int kTx = 32, kTy = 8;  // guided-filter block (256 lanes)
constexpr int kTileW = kTx + 2 * kGuidedR, kTileH = kTy + 2 * kGuidedR;

// cpuDepthMeters (DepthValidity.h:87-93): 0 for raw 0, raw >= 2047, or metres outside
// [lo, hi] (including the negative metres past the curve's pole).
__device__ __forceinline__ float usableM(unsigned r, float lo, float hi) {
    if (r == 0u || r >= 2047u) return 0.0f;
    const float m = 1.0f / (static_cast<float>(r) * kCurveA + kCurveB);
    return (isfinite(m) && m >= lo && m <= hi) ? m : 0.0f;
}

// cpuDepthMetersToRaw (DepthValidity.h:100-113). lround rounds half away from zero, as
// roundf does; a finite in-band value is far inside the integer range.
__device__ __forceinline__ uint16_t toRaw(float m, float lo, float hi) {
    if (!(isfinite(m) && m >= lo && m <= hi)) return 0;
    const float raw = (1.0f / m - kCurveB) / kCurveA;
    if (!isfinite(raw)) return 0;
    const float c = roundf(raw);
    if (!(c >= 1.0f && c <= 2046.0f)) return 0;
    return static_cast<uint16_t>(c);
}

__device__ __forceinline__ int reflect(int c, int n) { return c < 0 ? -c - 1 : (c >= n ? 2 * n - c - 1 : c); }

// exp(x) rounded to float from the double exponential. exp(-0) is exactly 1 either way;
// the branch only saves the double evaluation for the (common) equal-value taps.
__device__ __forceinline__ float expWeight(float x) {
    return x == 0.0f ? 1.0f : static_cast<float>(exp(static_cast<double>(x)));
}

// ---------------------------------------------------------------------------------
// denoiseDepthSpatial: for a usable centre, the element of rank count / 2 among the
// usable window samples (std::nth_element's value; the rank is a property of the
// multiset, not of the selection algorithm), when at least 3 are usable.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void condMedian(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int W,
                                                  int H, float lo, float hi) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const int i = y * W + x;
    const uint16_t c = in[i];
    if (usableM(c, lo, hi) == 0.0f) {
        out[i] = c;
        return;
    }
    float v[9];
    int n = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int sy = reflect(y + k / 3 - 1, H), sx = reflect(x + k % 3 - 1, W);
        v[k] = usableM(in[sy * W + sx], lo, hi);
        n += v[k] != 0.0f;
    }
    uint16_t r = c;
    if (n >= 3) {
        const int rank = n / 2;
        float med = 0.0f;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            int less = 0, le = 0;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                less += (v[k] != 0.0f) & (v[k] < v[j]);
                le += (v[k] != 0.0f) & (v[k] <= v[j]);
            }
            if (v[j] != 0.0f && less <= rank && rank < le) med = v[j];
        }
        r = toRaw(med, lo, hi);
    }
    out[i] = r;
}

// ---------------------------------------------------------------------------------
// fillDepthHoles: raw 0 only (raw 2047 is a measurement, not a hole); weights
// 1 / (dx^2 + dy^2 + 1) over the usable 5 x 5 neighbours, summed row by row.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void condFillHoles(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                                     int W, int H, float lo, float hi) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const int i = y * W + x;
    const uint16_t c = in[i];
    if (c != 0) {
        out[i] = c;
        return;
    }
    float ws = 0.0f, wsum = 0.0f;
    int nv = 0;
#pragma unroll
    for (int dy = -kHoleR; dy <= kHoleR; ++dy) {
        const int sy = reflect(y + dy, H);
#pragma unroll
        for (int dx = -kHoleR; dx <= kHoleR; ++dx) {
            if (dx == 0 && dy == 0) continue;
            const float m = usableM(in[sy * W + reflect(x + dx, W)], lo, hi);
            if (m == 0.0f) continue;
            const float w = 1.0f / (static_cast<float>(dx * dx + dy * dy) + 1.0f);
            ws += w * m;
            wsum += w;
            nv++;
        }
    }
    out[i] = (nv > 0 && wsum > 1e-6f) ? toRaw(ws / wsum, lo, hi) : c;
}

// ---------------------------------------------------------------------------------
// guidedDepthFilter. Weight of a usable tap = spatial * luma * depth * edge factor with
//   spatial = exp(-(dx^2 + dy^2) / 32), luma = exp(-(dl^2) / (2 0.1^2 + 0.01)),
//   depth = exp(-(dd^2) / (2 0.04^2 + 0.01)), edge factor = 1 - g / 2,
// g = min(|(r - l, d - u)| / 100, 1) from the raw codes of the 4-neighbourhood (1 at the
// image border or next to an unusable code). LUMA = false reads an all-zero guidance
// image (every luma weight is exp(-0) = 1).
// ---------------------------------------------------------------------------------
template <bool LUMA>
__global__ __launch_bounds__(256) void condGuided(const uint16_t* __restrict__ in, const float* __restrict__ luma,
                                                  uint16_t* __restrict__ out, int W, int H, float lo, float hi) {
    __shared__ float sm[kTileH][kTileW];
    __shared__ float sl[LUMA ? kTileH : 1][LUMA ? kTileW : 1];
    __shared__ uint16_t sc[kTileH][kTileW];
    __shared__ float sw_tab[2 * kGuidedR * kGuidedR + 1];

    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kTx - kGuidedR, y0 = blockIdx.y * kTy - kGuidedR;
    for (int k = tid; k < kTileW * kTileH; k += 256) {
        const int ty = k / kTileW, tx = k % kTileW;
        const int gx = x0 + tx, gy = y0 + ty;
        uint16_t c = 0;
        float l = 0.0f;
        // Positions more than the radius past the image are read only by lanes with no
        // pixel; leave them empty.
        if (gx >= -kGuidedR && gx < W + kGuidedR && gy >= -kGuidedR && gy < H + kGuidedR) {
            const int j = reflect(gy, H) * W + reflect(gx, W);
            c = in[j];
            if (LUMA) l = luma[j];
        }
        sc[ty][tx] = c;
        sm[ty][tx] = usableM(c, lo, hi);
        if (LUMA) sl[ty][tx] = l;
    }
    if (tid <= 2 * kGuidedR * kGuidedR)
        sw_tab[tid] = expWeight(-static_cast<float>(tid) / (2.0f * kGuidedR * kGuidedR));
    __syncthreads();

    const int lx = tid % kTx, ly = tid / kTx;
    const int x = blockIdx.x * kTx + lx, y = blockIdx.y * kTy + ly;
    if (x >= W || y >= H) return;
    const int cx = lx + kGuidedR, cy = ly + kGuidedR;
    const float cd = sm[cy][cx];
    const uint16_t code = sc[cy][cx];
    if (cd == 0.0f) {
        out[y * W + x] = code;
        return;
    }
    // computeDepthGradient (SignalConditioner_omp.cpp:118-142) on the stage's input codes.
    float edge = 1.0f;
    if (x > 0 && x < W - 1 && y > 0 && y < H - 1 && sm[cy][cx + 1] != 0.0f && sm[cy][cx - 1] != 0.0f &&
        sm[cy + 1][cx] != 0.0f && sm[cy - 1][cx] != 0.0f) {
        const float gx = fabsf(static_cast<float>(sc[cy][cx + 1]) - static_cast<float>(sc[cy][cx - 1]));
        const float gy = fabsf(static_cast<float>(sc[cy + 1][cx]) - static_cast<float>(sc[cy - 1][cx]));
        edge = fminf(sqrtf(gx * gx + gy * gy) / 100.0f, 1.0f);
    }
    const float ef = 1.0f - edge * 0.5f;
    const float cl = LUMA ? sl[cy][cx] : 0.0f;
    const float lumaDen = 2.0f * kSigmaLuma * kSigmaLuma + 0.01f;
    const float depthDen = 2.0f * kSigmaDepth * kSigmaDepth + 0.01f;

    float sw = 0.0f, sd = 0.0f;
    for (int dy = -kGuidedR; dy <= kGuidedR; ++dy) {
#pragma unroll
        for (int dx = -kGuidedR; dx <= kGuidedR; ++dx) {
            const float nd = sm[cy + dy][cx + dx];
            if (nd == 0.0f) continue;
            const float ws = sw_tab[dx * dx + dy * dy];
            float wl = 1.0f;
            if (LUMA) {
                const float gdl = sl[cy + dy][cx + dx] - cl;
                wl = expWeight(-(gdl * gdl) / lumaDen);
            }
            const float ddl = nd - cd;
            const float wd = expWeight(-(ddl * ddl) / depthDen);
            const float w = ws * wl * wd * ef;
            sw += w;
            sd += w * nd;
        }
    }
    out[y * W + x] = (sw > 1e-6f) ? toRaw(sd / sw, lo, hi) : code;
}

// ---------------------------------------------------------------------------------
// applyDepthEma: state[i] is the pixel's history in metres (0 = none). The reset test's
// neighbours are read from the input frame only (the reference's source snapshot), so
// the result does not depend on lane order.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void condEma(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                               float* __restrict__ state, int W, int H, float lo, float hi) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const int i = y * W + x;
    const uint16_t c = in[i];
    const float m = usableM(c, lo, hi);
    if (m == 0.0f) {
        state[i] = 0.0f;
        out[i] = c;
        return;
    }
    const float prev = state[i];
    const float delta = fabsf(m - prev);
    bool reset = isnan(prev) || isinf(prev) || prev <= 0.0f || delta > kEmaJump;
    if (!reset && delta > kEmaJump * 0.5f) {
        float sum = 0.0f;
        int n = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const float nd = usableM(in[reflect(y + dy, H) * W + reflect(x + dx, W)], lo, hi);
                if (nd != 0.0f) {
                    sum += nd;
                    n++;
                }
            }
        if (n > 0 && fabsf(m - sum / static_cast<float>(n)) > kEmaJump) reset = true;
    }
    const float filtered = reset ? m : (0.7f * m + 0.3f * prev);
    const uint16_t enc = toRaw(filtered, lo, hi);
    state[i] = enc != 0 ? filtered : 0.0f;
    out[i] = enc;
}

__global__ __launch_bounds__(256) void condZero(float* __restrict__ p, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace

// Stage launchers. Sizes are checked by the caller (W, H >= kCondMinSide).
hipError_t launchCondStage(int stage, const uint16_t* in, const float* luma, uint16_t* out, float* state, int W,
                           int H, float lo, float hi, hipStream_t st) {
    const dim3 g64((W + 63) / 64, (H + 3) / 4);
    switch (stage) {
        case 1:
            hipLaunchKernelGGL(condMedian, g64, dim3(256), 0, st, in, out, W, H, lo, hi);
            break;
        case 2:
            hipLaunchKernelGGL(condFillHoles, g64, dim3(256), 0, st, in, out, W, H, lo, hi);
            break;
        case 4: {
            const dim3 g((W + kTx - 1) / kTx, (H + kTy - 1) / kTy);
            if (luma)
                hipLaunchKernelGGL(condGuided<true>, g, dim3(256), 0, st, in, luma, out, W, H, lo, hi);
            else
                hipLaunchKernelGGL(condGuided<false>, g, dim3(256), 0, st, in, luma, out, W, H, lo, hi);
            break;
        }
        case 8:
            hipLaunchKernelGGL(condEma, g64, dim3(256), 0, st, in, out, state, W, H, lo, hi);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchCondZero(float* p, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(condZero, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace azu
