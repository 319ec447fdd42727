// Adjoint of the raycast-depth loss on gfx950 (no counterpart in the reference; the
// march and the secant follow TSDFVolume::raycast, src/tsdf/TSDFVolume.cpp:384-445,
// the increment convention follows ICPTracker.cpp:100-121).
//
//   L = sum over pixels with a hit and an observed depth of (d_ray - d_obs)^2,
//   d_ray = t* / |ray|   (camera Z-depth of the hit),
//   t*    = t_a + (t_b - t_a) f_a / (f_a - f_b)   (secant on the first front-face bracket)
//
// The march uses a uniform step (a fixed fraction of a voxel) so that, with the bracket
// fixed, t* is a smooth function of the two bracketing samples f_a = f(p(t_a)),
// f_b = f(p(t_b)). Each f is the trilinear interpolant of 8 voxel values, so
//   dL/dV_i  = dL/dt* (dt*/df_a w_i(p_a) + dt*/df_b w_i(p_b))     (scattered with atomics)
//   dL/dxi   = dL/dt* sum_s dt*/df_s grad f(p_s)^T R [I | -[q_s]x]
// for the pose perturbation pose <- pose * [R(w) | dt], xi = [dt; w], q_s = t_s r_hat the
// sample point in camera coordinates. The backward kernel recomputes the march instead
// of storing per-step samples. Pose gradients are accumulated per pixel in FP64, folded
// per block in a fixed tree, and summed over blocks in block order by a second
// one-block kernel, so the pose gradient is deterministic; the TSDF scatter is not
// (float atomics), within the stated tolerance.
#include "common.hpp"

namespace azu {

template <class S> struct AdjVol {
    const typename S::Word* vox;
    int res;
    float vs, inv_vs;
    float ox, oy, oz;
};

struct TriSample {
    bool valid;
    float f;
    float g[3];        // df / d(world position)
    size_t base;       // index of corner (x0, y0, z0)
    float tx, ty, tz;  // fractional position in the cell
};

template <class S>
__device__ __forceinline__ TriSample sampleGrad(const AdjVol<S>& V, F3 p) {
    TriSample s;
    s.valid = false;
    const float gx = (p.x - V.ox) * V.inv_vs, gy = (p.y - V.oy) * V.inv_vs, gz = (p.z - V.oz) * V.inv_vs;
    const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
    const float top = (float)(V.res - 1);
    if (!(fx0 >= 0.0f && fy0 >= 0.0f && fz0 >= 0.0f && fx0 < top && fy0 < top && fz0 < top)) return s;
    const size_t sy = (size_t)V.res, sz = sy * sy;
    const size_t b = (size_t)fz0 * sz + (size_t)fy0 * sy + (size_t)fx0;
    const size_t off[8] = {0, 1, sy, sy + 1, sz, sz + 1, sz + sy, sz + sy + 1};
    float c[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const typename S::Word w = V.vox[b + off[i]];
        if (!(S::weight(w) > 0.0f)) return s;
        c[i] = S::tsdf(w);
    }
    const float tx = gx - fx0, ty = gy - fy0, tz = gz - fz0;
    const float ux = 1.0f - tx, uy = 1.0f - ty, uz = 1.0f - tz;
    const float v00 = c[0] * ux + c[1] * tx, v10 = c[2] * ux + c[3] * tx;
    const float v01 = c[4] * ux + c[5] * tx, v11 = c[6] * ux + c[7] * tx;
    const float v0 = v00 * uy + v10 * ty, v1 = v01 * uy + v11 * ty;
    s.f = v0 * uz + v1 * tz;
    // partial derivatives of the interpolant in voxel units, then per metre
    const float dx = ((c[1] - c[0]) * uy + (c[3] - c[2]) * ty) * uz + ((c[5] - c[4]) * uy + (c[7] - c[6]) * ty) * tz;
    const float dy = (v10 - v00) * uz + (v11 - v01) * tz;
    const float dz = v1 - v0;
    s.g[0] = dx * V.inv_vs;
    s.g[1] = dy * V.inv_vs;
    s.g[2] = dz * V.inv_vs;
    s.base = b;
    s.tx = tx;
    s.ty = ty;
    s.tz = tz;
    s.valid = true;
    return s;
}

// The value alone (the march): the same cell, gate and interpolation as sampleGrad, its
// eight corners read as four x-adjacent pairs.
template <class S>
__device__ __forceinline__ bool sampleValue(const AdjVol<S>& V, F3 p, float& f) {
    const float gx = (p.x - V.ox) * V.inv_vs, gy = (p.y - V.oy) * V.inv_vs, gz = (p.z - V.oz) * V.inv_vs;
    const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
    const float top = (float)(V.res - 1);
    if (!(fx0 >= 0.0f && fy0 >= 0.0f && fz0 >= 0.0f && fx0 < top && fy0 < top && fz0 < top)) return false;
    const size_t sy = (size_t)V.res, sz = sy * sy;
    const size_t b = (size_t)fz0 * sz + (size_t)fy0 * sy + (size_t)fx0;
    typename S::Word w[8];
    __builtin_memcpy(&w[0], V.vox + b, 2 * sizeof(typename S::Word));
    __builtin_memcpy(&w[2], V.vox + b + sy, 2 * sizeof(typename S::Word));
    __builtin_memcpy(&w[4], V.vox + b + sz, 2 * sizeof(typename S::Word));
    __builtin_memcpy(&w[6], V.vox + b + sz + sy, 2 * sizeof(typename S::Word));
    bool ok = true;
    float c[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        ok = ok && S::weight(w[i]) > 0.0f;
        c[i] = S::tsdf(w[i]);
    }
    if (!ok) return false;
    const float tx = gx - fx0, ty = gy - fy0, tz = gz - fz0;
    const float ux = 1.0f - tx, uy = 1.0f - ty, uz = 1.0f - tz;
    const float v00 = c[0] * ux + c[1] * tx, v10 = c[2] * ux + c[3] * tx;
    const float v01 = c[4] * ux + c[5] * tx, v11 = c[6] * ux + c[7] * tx;
    const float v0 = v00 * uy + v10 * ty, v1 = v01 * uy + v11 * ty;
    f = v0 * uz + v1 * tz;
    return true;
}

__device__ __forceinline__ void scatterTrilinear(float* grad, size_t b, size_t sy, size_t sz, float tx, float ty,
                                                 float tz, float g) {
    const float ux = 1.0f - tx, uy = 1.0f - ty, uz = 1.0f - tz;
    atomicAdd(grad + b, g * ux * uy * uz);
    atomicAdd(grad + b + 1, g * tx * uy * uz);
    atomicAdd(grad + b + sy, g * ux * ty * uz);
    atomicAdd(grad + b + sy + 1, g * tx * ty * uz);
    atomicAdd(grad + b + sz, g * ux * uy * tz);
    atomicAdd(grad + b + sz + 1, g * tx * uy * tz);
    atomicAdd(grad + b + sz + sy, g * ux * ty * tz);
    atomicAdd(grad + b + sz + sy + 1, g * tx * ty * tz);
}

constexpr int kAdjSlots = 8;  // loss, grad pose [6], contributing pixels

struct AdjArgs {
    VolParams p;
    PoseArg pa;
    Intr K;
    int W, H;
    const float* d_obs;  // observed depth, metres, W*H (<= 0 or non-finite: no observation)
    float step;          // march step in metres
    int backward;
    float* grad_tsdf;    // res^3, accumulated (+=); may be null
    float* depth_out;    // W*H ray depth (0 on a miss); may be null
    double* partials;    // [slot][block]
    // Empty-space skip bounds of this pose (brickSplat: per 2 x 2 pixel cell the nearest
    // and farthest distance of a marked brick the cell's rays can meet), or null; the
    // first pixel of each cell puts the cell back to its resting state after use.
    unsigned* dsafe;
    const int* skip_ok;
};

template <class S>
__global__ __launch_bounds__(256) void depthLossKernel(AdjVol<S> V, AdjArgs a) {
    __shared__ double s_w[4][kAdjSlots];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x = (int)(blockIdx.x % ((a.W + 15) / 16)) * 16 + (tid & 15);
    const int y = (int)(blockIdx.x / ((a.W + 15) / 16)) * 16 + (tid >> 4);
    double acc[kAdjSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (x < a.W && y < a.H) {
        const Rigid pose = rigidFrom(a.pa);
        const F3 cam = pose.t;
        const F3 ray_cam{((float)x - a.K.cx) / a.K.fx, ((float)y - a.K.cy) / a.K.fy, 1.0f};
        const float ray_norm = f3norm(ray_cam);
        const F3 rhat = f3div(ray_cam, ray_norm);
        const F3 dir = mulv(pose.R, rhat);
        float t_near = a.p.min_depth * ray_norm, t_far = a.p.max_depth * ray_norm;
        const float span = (float)(a.p.res - 1) * a.p.vs;
        const float cam_a[3] = {cam.x, cam.y, cam.z}, dir_a[3] = {dir.x, dir.y, dir.z};
        for (int k = 0; k < 3; ++k) {
            const float lo = a.p.origin[k], hi = a.p.origin[k] + span;
            const float inv = 1.0f / dir_a[k];
            float t0 = (lo - cam_a[k]) * inv, t1 = (hi - cam_a[k]) * inv;
            if (t0 > t1) {
                const float s = t0;
                t0 = t1;
                t1 = s;
            }
            if (__builtin_isnan(t0) || __builtin_isnan(t1)) {
                if (cam_a[k] < lo || cam_a[k] > hi) t_far = -1.0f;
                continue;
            }
            t_near = smax(t_near, t0);
            t_far = smin(t_far, t1);
        }
        // Lattice samples before t_safe lie in no marked brick, so all eight corners of
        // their cells are far (unobserved, or tsdf >= 0.999): such a sample is invalid or
        // positive, and the march passes it without stopping or hitting. Samples past
        // t_end are far too, so no hit can follow them.
        float t_safe = -__builtin_inff(), t_end = __builtin_inff();
        const int cell = (y >> 1) * ((a.W + 1) >> 1) + (x >> 1);
        if (a.dsafe) {
            const uint2 d = *reinterpret_cast<const uint2*>(a.dsafe + 2 * (size_t)cell);
            if (*a.skip_ok) {
                const float dn = f3norm(dir);
                t_safe = __uint_as_float(d.x) / dn;
                t_end = __uint_as_float(d.y) / dn;
            }
        }
        float depth = 0.0f;
        if (t_near < t_far) {
            // Samples sit on the lattice t_k = t0 + k h anchored at the pose-independent
            // t0 = min_depth |ray|, starting at the first lattice point inside the box, so
            // a small pose change moves the samples' values but not their parameters.
            const float t0 = a.p.min_depth * ray_norm;
            const int k0 = t_near > t0 ? (int)ceilf((t_near - t0) / a.step) : 0;
            const int k1 = (int)floorf((t_far - t0) / a.step);
            int k = k0;
            if (t0 + (float)k0 * a.step < t_safe && k0 < k1) {
                // start at the last lattice sample before t_safe (its value enters the next
                // crossing test); the ones before it are passed over unread
                int ks = (int)floorf((t_safe - t0) / a.step);
                ks = min(max(ks, k0), k1);
                while (ks > k0 && !(t0 + (float)ks * a.step < t_safe)) --ks;
                while (ks < k1 && t0 + (float)(ks + 1) * a.step < t_safe) ++ks;
                k = ks;
            }
            float t_prev = t0 + (float)k * a.step;
            // The march reads values only; the two samples of a bracket are read again
            // with their gradients once it is found.
            float fp = 0.0f;
            bool vp = sampleValue<S>(V, f3add(cam, f3mul(dir, t_prev)), fp);
            bool go = k <= k1 && !(vp && (!isfinite(fp) || fp < 0.0f));
            bool found = false;
            float t_b = 0.0f;
            while (go && k < k1 && t_prev < t_end) {
                ++k;
                const float t = t0 + (float)k * a.step;
                float fc = 0.0f;
                const bool vc = sampleValue<S>(V, f3add(cam, f3mul(dir, t)), fc);
                if (vc && !isfinite(fc)) break;
                if (vc && vp) {
                    if (fp > 0.0f && fc <= 0.0f) {
                        found = true;
                        t_b = t;
                        break;
                    }
                    if (fp < 0.0f && fc >= 0.0f) break;
                } else if (vc && fc < 0.0f) {
                    break;
                }
                t_prev = t;
                fp = fc;
                vp = vc;
            }
            TriSample sp, sb;
            if (found) {
                sp = sampleGrad<S>(V, f3add(cam, f3mul(dir, t_prev)));
                sb = sampleGrad<S>(V, f3add(cam, f3mul(dir, t_b)));
            }
            const float d_obs = a.d_obs ? a.d_obs[y * a.W + x] : 0.0f;
            if (found) {
                const float fa = sp.f, fb = sb.f, dt = t_b - t_prev;
                const float den = fa - fb;
                const float ts = t_prev + dt * (fa / den);
                depth = ts / ray_norm;
                if (d_obs > 0.0f && isfinite(d_obs)) {
                    const double r = (double)depth - (double)d_obs;
                    acc[0] = r * r;
                    acc[7] = 1.0;
                    if (a.backward) {
                        const float dLdt = (float)(2.0 * r) / ray_norm;
                        const float ga = dLdt * dt * (-fb) / (den * den);
                        const float gb = dLdt * dt * fa / (den * den);
                        const size_t sy = (size_t)V.res, sz = sy * sy;
                        if (a.grad_tsdf) {
                            scatterTrilinear(a.grad_tsdf, sp.base, sy, sz, sp.tx, sp.ty, sp.tz, ga);
                            scatterTrilinear(a.grad_tsdf, sb.base, sy, sz, sb.tx, sb.ty, sb.tz, gb);
                        }
                        // pose: g_s grad_f(p_s)^T R [I | -[q_s]x], q_s = t_s r_hat
                        const float gs[2] = {ga, gb};
                        const float tsv[2] = {t_prev, t_b};
                        const TriSample* ss[2] = {&sp, &sb};
                        for (int k = 0; k < 2; ++k) {
                            const F3 gw{ss[k]->g[0], ss[k]->g[1], ss[k]->g[2]};
                            const F3 gc = mulvT(pose.R, gw);  // R^T grad f: gradient in camera axes
                            const F3 q = f3mul(rhat, tsv[k]);
                            const F3 gr = f3cross(q, gc);     // -(gc^T [q]x) = (q x gc)^T
                            acc[1] += (double)gs[k] * gc.x;
                            acc[2] += (double)gs[k] * gc.y;
                            acc[3] += (double)gs[k] * gc.z;
                            acc[4] += (double)gs[k] * gr.x;
                            acc[5] += (double)gs[k] * gr.y;
                            acc[6] += (double)gs[k] * gr.z;
                        }
                    }
                }
            }
        }
        if (a.depth_out) a.depth_out[y * a.W + x] = depth;
        if (a.dsafe && ((x | y) & 1) == 0)
            *reinterpret_cast<uint2*>(a.dsafe + 2 * (size_t)cell) = make_uint2(0x7f800000u, 0u);
    }
    // Fixed-order fold: wave butterfly, then the four waves in order.
#pragma unroll
    for (int i = 0; i < kAdjSlots; ++i) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        acc[i] = v;
    }
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < kAdjSlots; ++i) s_w[wv][i] = acc[i];
    __syncthreads();
    if (tid < kAdjSlots)
        a.partials[(size_t)tid * gridDim.x + blockIdx.x] = ((s_w[0][tid] + s_w[1][tid]) + s_w[2][tid]) + s_w[3][tid];
}

// Sums the per-block partials in block order (one block, one slot per wave).
__global__ void depthLossFinish(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int s = wv; s < kAdjSlots; s += blockDim.x / 64) {
        double v = 0.0;
        for (int b = lane; b < nb; b += 64) v += partials[(size_t)s * nb + b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) out[s] = v;
    }
}

int depthLossBlocks(int W, int H) { return ((W + 15) / 16) * ((H + 15) / 16); }

hipError_t launchBrickSplat(const unsigned char* bricks, const VolParams& p, const PoseArg& pa, Intr K, int W, int H,
                            unsigned* dsafe, int* ok, hipStream_t st);

// bricks / dsafe / ok as for launchRaycast (null bricks: every lattice sample is read).
hipError_t launchDepthLoss(int format, const void* vox, const VolParams& p, const PoseArg& pa, Intr K, int W, int H,
                           const float* d_obs, float step, int backward, float* grad_tsdf, float* depth_out,
                           double* partials, double* out8, const unsigned char* bricks, unsigned* dsafe, int* ok,
                           hipStream_t st) {
    AdjArgs a;
    a.p = p;
    a.pa = pa;
    a.K = K;
    a.W = W;
    a.H = H;
    a.d_obs = d_obs;
    a.step = step;
    a.backward = backward;
    a.grad_tsdf = grad_tsdf;
    a.depth_out = depth_out;
    a.partials = partials;
    a.dsafe = nullptr;
    a.skip_ok = ok;
    if (bricks && dsafe && ok) {
        const hipError_t e = launchBrickSplat(bricks, p, pa, K, W, H, dsafe, ok, st);
        if (e != hipSuccess) return e;
        a.dsafe = dsafe;
    }
    const int nb = depthLossBlocks(W, H);
    if (format == kFmtF32) {
        AdjVol<StoreF32> V{static_cast<const float2*>(vox), p.res, p.vs, 1.0f / p.vs, p.origin[0], p.origin[1],
                           p.origin[2]};
        hipLaunchKernelGGL(depthLossKernel<StoreF32>, dim3(nb), dim3(256), 0, st, V, a);
    } else {
        AdjVol<StoreF16> V{static_cast<const uint32_t*>(vox), p.res, p.vs, 1.0f / p.vs, p.origin[0], p.origin[1],
                           p.origin[2]};
        hipLaunchKernelGGL(depthLossKernel<StoreF16>, dim3(nb), dim3(256), 0, st, V, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(depthLossFinish, dim3(1), dim3(256), 0, st, partials, nb, out8);
    return hipGetLastError();
}

}  // namespace azu
