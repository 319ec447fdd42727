// TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's depth conditioning stages
// (src/sensor/SignalConditioner_omp.cpp:442-716 of iphands/azu) over raw Kinect v1 codes,
// with the raw <-> metres conversions of include/sensor/DepthValidity.h:55-119 and the
// reflect border of include/sensor/BorderMode.h:42-46. Single-threaded, float arithmetic
// in the reference's order. The product never links it.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc {
namespace cond {

constexpr float kCurveA = -0.0030711016f;
constexpr float kCurveB = 3.3309495161f;
constexpr int kMedianRadius = 1;        // SignalConditioner_omp.cpp:35
constexpr int kHoleFillRadius = 2;      // :30
constexpr int kGuidedRadius = 4;        // :33
constexpr float kEmaJumpReset = 0.05f;  // :41
constexpr float kSigmaLuma = 0.10f;     // :42
constexpr float kSigmaDepth = 0.04f;    // :43

// isInvalidRawDepth / rawDepthToMeters / isDepthMetersInBand / cpuDepthMeters
// (DepthValidity.h:55-97)
inline bool invalidRaw(uint16_t r) { return r == 0 || r >= 2047; }
inline float rawToMeters(uint16_t r) {
    if (invalidRaw(r)) return 0.0f;
    return 1.0f / (static_cast<float>(r) * kCurveA + kCurveB);
}
inline bool inBand(float m, float lo, float hi) { return std::isfinite(m) && m >= lo && m <= hi; }
inline float usableMeters(uint16_t r, float lo, float hi) {
    if (invalidRaw(r)) return 0.0f;
    const float m = rawToMeters(r);
    return inBand(m, lo, hi) ? m : 0.0f;
}
// cpuDepthMetersToRaw (DepthValidity.h:100-113): 0 when out of band or off the code range
inline uint16_t metersToRaw(float m, float lo, float hi) {
    if (!inBand(m, lo, hi)) return 0;
    const float raw = (1.0f / m - kCurveB) / kCurveA;
    if (!std::isfinite(raw)) return 0;
    const long code = std::lround(raw);
    if (code < 1 || code > 2046) return 0;
    return static_cast<uint16_t>(code);
}
// cpuReflectCoord (BorderMode.h:42-46)
inline int reflect(int c, int n) {
    if (c < 0) return -c - 1;
    if (c >= n) return 2 * n - c - 1;
    return c;
}

// denoiseDepthSpatial (SignalConditioner_omp.cpp:442-478): the median of the usable
// samples of the reflected 3x3 window (element count / 2 of their sorted order) for a
// usable centre with at least 3 of them; every other pixel keeps its code.
inline void median(const uint16_t* in, uint16_t* out, int W, int H, float lo, float hi) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int i = y * W + x;
            out[i] = in[i];
            if (usableMeters(in[i], lo, hi) == 0.0f) continue;
            float win[9];
            int n = 0;
            for (int dy = -kMedianRadius; dy <= kMedianRadius; ++dy) {
                const int sy = reflect(y + dy, H);
                for (int dx = -kMedianRadius; dx <= kMedianRadius; ++dx) {
                    const float m = usableMeters(in[sy * W + reflect(x + dx, W)], lo, hi);
                    if (m != 0.0f) win[n++] = m;
                }
            }
            if (n >= 3) {
                std::nth_element(win, win + n / 2, win + n);
                out[i] = metersToRaw(win[n / 2], lo, hi);
            }
        }
}

// fillDepthHoles (SignalConditioner_omp.cpp:578-635): raw 0 only; the inverse-distance
// weighted mean (weight 1 / (dx^2 + dy^2 + 1)) of the usable samples of the reflected
// 5 x 5 window, centre excluded, summed row by row.
inline void fillHoles(const uint16_t* in, uint16_t* out, int W, int H, float lo, float hi) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int i = y * W + x;
            out[i] = in[i];
            if (in[i] != 0) continue;
            float ws = 0.0f, wsum = 0.0f;
            int nv = 0;
            for (int dy = -kHoleFillRadius; dy <= kHoleFillRadius; ++dy) {
                const int sy = reflect(y + dy, H);
                for (int dx = -kHoleFillRadius; dx <= kHoleFillRadius; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    const float m = usableMeters(in[sy * W + reflect(x + dx, W)], lo, hi);
                    if (m == 0.0f) continue;
                    const int d2 = dx * dx + dy * dy;
                    const float w = 1.0f / (d2 + 1.0f);
                    ws += w * m;
                    wsum += w;
                    nv++;
                }
            }
            if (nv > 0 && wsum > 1e-6f) out[i] = metersToRaw(ws / wsum, lo, hi);
        }
}

// computeDepthGradient (SignalConditioner_omp.cpp:118-142)
inline float depthGradient(const uint16_t* d, int x, int y, int W, int H, float lo, float hi) {
    if (x <= 0 || x >= W - 1 || y <= 0 || y >= H - 1) return 1.0f;
    const uint16_t c = d[y * W + x], r = d[y * W + x + 1], l = d[y * W + x - 1], dn = d[(y + 1) * W + x],
                   up = d[(y - 1) * W + x];
    if (usableMeters(c, lo, hi) == 0.0f || usableMeters(r, lo, hi) == 0.0f || usableMeters(l, lo, hi) == 0.0f ||
        usableMeters(dn, lo, hi) == 0.0f || usableMeters(up, lo, hi) == 0.0f)
        return 1.0f;
    const float gx = std::abs(static_cast<float>(r) - static_cast<float>(l));
    const float gy = std::abs(static_cast<float>(dn) - static_cast<float>(up));
    const float g = std::sqrt(gx * gx + gy * gy);
    return std::min(g / 100.0f, 1.0f);
}

// The reference evaluates its Gaussian weights with std::exp on float (the C library's
// expf, which is faithfully but not always correctly rounded). Mode 1 rounds the double
// exponential to float instead: correctly rounded except within ~1e-16 relative of a
// float rounding midpoint, i.e. independent of whose expf is linked.
inline float expWeight(float x, int mode) {
    return mode == 1 ? static_cast<float>(std::exp(static_cast<double>(x))) : std::exp(x);
}

// guidedDepthFilter (SignalConditioner_omp.cpp:637-716): for a usable centre, the
// weighted mean of the usable samples of the reflected 9 x 9 window, weight =
// spatial * luma * depth * edge factor, each Gaussian term as the reference writes it.
inline void guided(const uint16_t* in, const float* luma, uint16_t* out, int W, int H, float lo, float hi,
                   int exp_mode = 0) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int i = y * W + x;
            const float cd = usableMeters(in[i], lo, hi);
            if (cd == 0.0f) {
                out[i] = in[i];
                continue;
            }
            const float cl = luma[i];
            const float edge = depthGradient(in, x, y, W, H, lo, hi);
            float sw = 0.0f, sd = 0.0f;
            for (int dy = -kGuidedRadius; dy <= kGuidedRadius; ++dy) {
                const int sy = reflect(y + dy, H);
                for (int dx = -kGuidedRadius; dx <= kGuidedRadius; ++dx) {
                    const int n = sy * W + reflect(x + dx, W);
                    const float nd = usableMeters(in[n], lo, hi);
                    if (nd == 0.0f) continue;
                    const float sp = static_cast<float>(dx * dx + dy * dy);
                    const float gdl = luma[n] - cl;
                    const float ddl = (cd > 0.0f) ? (nd - cd) : 0.0f;
                    const float ws = expWeight(-sp / (2.0f * kGuidedRadius * kGuidedRadius), exp_mode);
                    const float wl = expWeight(-(gdl * gdl) / (2.0f * kSigmaLuma * kSigmaLuma + 0.01f), exp_mode);
                    const float wd = expWeight(-(ddl * ddl) / (2.0f * kSigmaDepth * kSigmaDepth + 0.01f), exp_mode);
                    const float ef = 1.0f - edge * 0.5f;
                    const float w = ws * wl * wd * ef;
                    sw += w;
                    sd += w * nd;
                }
            }
            out[i] = (sw > 1e-6f) ? metersToRaw(sd / sw, lo, hi) : in[i];
        }
}

// applyDepthEma (SignalConditioner_omp.cpp:494-576): per-pixel temporal state (metres,
// 0 = no history) read against the unchanged input frame.
inline void ema(const uint16_t* in, uint16_t* out, float* state, int W, int H, float lo, float hi) {
    for (int i = 0; i < W * H; ++i) {
        const float m = usableMeters(in[i], lo, hi);
        if (m == 0.0f) {
            state[i] = 0.0f;
            out[i] = in[i];
            continue;
        }
        const float prev = state[i];
        const float delta = std::abs(m - prev);
        bool reset = std::isnan(prev) || std::isinf(prev) || prev <= 0.0f || delta > kEmaJumpReset;
        if (!reset && delta > kEmaJumpReset * 0.5f) {
            const int x = i % W, y = i / W;
            float sum = 0.0f;
            int n = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    const float nd = usableMeters(in[reflect(y + dy, H) * W + reflect(x + dx, W)], lo, hi);
                    if (nd != 0.0f) {
                        sum += nd;
                        n++;
                    }
                }
            if (n > 0) {
                const float avg = sum / n;
                if (std::abs(m - avg) > kEmaJumpReset) reset = true;
            }
        }
        const float filtered = reset ? m : (0.7f * m + 0.3f * prev);
        const uint16_t code = metersToRaw(filtered, lo, hi);
        state[i] = (code != 0) ? filtered : 0.0f;
        out[i] = code;
    }
}

// processCpu's depth half (SignalConditioner_omp.cpp:337-349): median, then (unless
// the conditioner is in minimal mode) hole fill, guided filter, temporal EMA.
// stages: bit 0 median, bit 1 holes, bit 2 guided, bit 3 EMA, applied in that order.
inline void process(const uint16_t* in, const float* luma, uint16_t* out, float* state, int W, int H, float lo,
                    float hi, int stages, int exp_mode = 0) {
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<uint16_t> a(in, in + n), b(n);
    if (stages & 1) median(a.data(), b.data(), W, H, lo, hi), a.swap(b);
    if (stages & 2) fillHoles(a.data(), b.data(), W, H, lo, hi), a.swap(b);
    if (stages & 4) guided(a.data(), luma, b.data(), W, H, lo, hi, exp_mode), a.swap(b);
    if (stages & 8) ema(a.data(), b.data(), state, W, H, lo, hi), a.swap(b);
    std::copy(a.begin(), a.end(), out);
}

}  // namespace cond
}  // namespace orc
