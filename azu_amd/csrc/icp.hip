// Projective point-to-plane ICP on gfx950 (the reference's ICPTracker::track /
// trackLevel / buildLinearSystem, src/tracking/ICPTracker.cpp:40-354, with the
// constants and predicates of include/tracking/ICPShared.h:34-110).
//
// One kernel launch per ICP iteration (and a final solve-only launch), no host round
// trip. Launch k:
//   1. solves iteration k - 1 in every block: the FP64 partials that launch k - 1 wrote
//      are summed in a fixed order (deterministic, so every block gets the same sums),
//      and one lane runs the solve: adaptive damping from the extreme eigenvalues, FP64
//      Cholesky of A + lambda I, step caps, Rodrigues update, pose * delta,
//      re-projection onto SO(3), convergence test. Block 0 records the result. The
//      kernel boundary orders the partials, so no block waits for another;
//   2. then, unless the level has stopped, every thread walks its live pixels, associates
//      each with the model image, builds J = [R^T n_m ; v x R^T n_m], the Huber weight
//      and loss, and accumulates the 21 upper-triangle A terms, 6 b terms and the
//      residual as Kahan-compensated FP32 partial sums, plus 6 integer counters; each
//      wave sums its lanes' compensated sums by a fixed DPP tree, and the block adds its
//      wave sums in FP64 and writes one FP64 partial per slot.
// The tracker state (pose, per-level result, stop flag) lives in device memory, so all
// launches of a frame are enqueued back to back; once a level has stopped, its remaining
// launches return at once.
#include "common.hpp"

namespace azu {

constexpr int kIcpSlotsF = 28;  // 21 A + 6 b + residual
constexpr int kIcpSlotsI = 6;   // inliers, valid_live, valid_model, projected, dist_f, angle_f
constexpr int kIcpSlots = kIcpSlotsF + kIcpSlotsI;
constexpr int kIcpMaxBlocks = 256;
constexpr int kIcpThreads = 512;              // threads per block: two waves per SIMD
constexpr int kIcpWaves = kIcpThreads / 64;
constexpr int kSumGroups = 7;  // cross-block sum: 7 threads per slot (7 * 34 = 238 <= 256)

// Pose terms every thread of an iteration uses: the current pose (camera to world) and
// the transform from the live camera into the reference camera. Computed once per
// pose change (icpInit, the last block's update) into the state, which every block of
// the next launch reads at its start.
struct PoseDerived {
    M33 R_cw, R_rel;
    F3 t_cw, t_rel;
    int iterations;  // the state's packed step counts for the next launch
};

struct IcpState {
    float pose[16];
    float ref[16];
    float information[36];
    float error;
    float final_step;       // last accepted step of the current level (inf if none)
    float last_final_step;  // last accepted step over all levels (inf if none)
    int inliers;
    int converged;
    int counters[5];  // valid_live, valid_model, projected, dist_filtered, angle_filtered
    int iterations;   // accepted steps, one byte per level (level l at bit 8*l)
    // Phase timing of the last block, in 100 MHz clock ticks summed over the launches
    // that reached it: cross-block sum, solve; and the number of such launches. Work of
    // the same launches: interior live pixels visited, in-bounds model projections.
    unsigned long long ticks_sum, ticks_solve, ticks_n, work_px, work_proj;
    // Per launch parity p: launch k reads it[k & 1] and its block 0 writes it[1 - (k & 1)].
    struct Slot {
        PoseDerived pd;  // the pose terms the launch's pass uses (after its solve)
        int pending;     // the launch before ran a pass: its partials await a solve
    } it[2];
    int stop_level;  // level whose iteration loop has ended, -1 while none
    Rigid ref_inv;   // inverse of ref
};

struct IcpIterArgs {
    const float* lv;  // live level vertices / normals, packed xyz
    const float* ln;
    int W, H;
    const float* mv;  // model image vertices / normals (world), packed xyz
    const float* mn;
    int MW, MH;
    Intr Km;          // intrinsics scaled to the model image
    float angle_cos;  // cos(angle threshold), rounded to float like the reference
    float dist_sq;    // dist_threshold^2 in float
    int level, iter;
};

// DPP lane move: lanes whose source is out of range or in a masked row read 0, the
// identity of the sum below.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dppI(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, true);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dppF(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

// Wave64 float sum to lane 63, the same DPP sequence as waveSumInt below (a fixed tree
// of six adds per lane, so at most six roundings between a thread's sum and the wave's).
__device__ __forceinline__ float waveSumF(float v) {
    v += dppF<0xb1, 0xf>(v);
    v += dppF<0x4e, 0xf>(v);
    v += dppF<0x114, 0xf>(v);
    v += dppF<0x118, 0xf>(v);
    v += dppF<0x142, 0xa>(v);
    v += dppF<0x143, 0xc>(v);
    return v;
}

// Kahan-compensated accumulation of one term into the sum s with compensation c (the
// sum carried is s - c).
__device__ __forceinline__ void kahanAdd(float& s, float& c, float term) {
    const float y = term - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Wave64 integer sum to lane 63 with the gfx9 DPP sequence: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_shr:4, row_shr:8, row_bcast:15 (rows 1, 3), row_bcast:31
// (rows 2, 3).
__device__ __forceinline__ int waveSumInt(int v) {
    v += dppI<0xb1, 0xf>(v);
    v += dppI<0x4e, 0xf>(v);
    v += dppI<0x114, 0xf>(v);
    v += dppI<0x118, 0xf>(v);
    v += dppI<0x142, 0xa>(v);
    v += dppI<0x143, 0xc>(v);
    return v;
}

// ---------------------------------------------------------------------------
// FP64 small dense algebra for the single-lane solve.
// ---------------------------------------------------------------------------
// Cyclic Jacobi eigenvalues (Golub & Van Loan 8.5) of a symmetric 6x6; returns min, max.
__device__ void jacobiMinMax(const double* Ain, double& mn, double& mx) {
    double a[36];
    for (int i = 0; i < 36; ++i) a[i] = Ain[i];
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 6; ++i) {
            diag += a[i * 6 + i] * a[i * 6 + i];
            for (int j = i + 1; j < 6; ++j) off += a[i * 6 + j] * a[i * 6 + j];
        }
        if (off == 0.0 || off <= 1e-32 * diag) break;
        for (int p = 0; p < 6; ++p)
            for (int q = p + 1; q < 6; ++q) {
                const double apq = a[p * 6 + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * 6 + q] - a[p * 6 + p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; ++k) {
                    const double akp = a[k * 6 + p], akq = a[k * 6 + q];
                    a[k * 6 + p] = c * akp - s * akq;
                    a[k * 6 + q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; ++k) {
                    const double apk = a[p * 6 + k], aqk = a[q * 6 + k];
                    a[p * 6 + k] = c * apk - s * aqk;
                    a[q * 6 + k] = s * apk + c * aqk;
                }
            }
    }
    mn = a[0];
    mx = a[0];
    for (int i = 1; i < 6; ++i) {
        mn = fmin(mn, a[i * 6 + i]);
        mx = fmax(mx, a[i * 6 + i]);
    }
}

// dampingForHessian (ICPShared.h:86-100): lambda is the level's base (0.01 at level
// 0, 0.1 above) unless min_eig < 1e-4 or max_eig > 1e7 min_eig, then 1.0. A Cholesky
// certificate settles the common case without an eigen-solve: with g >= max_eig
// (Gershgorin) and s = max(1e-4, g / 1e7), a successful factorisation of A - s I
// proves min_eig > s, hence neither escalation condition.
//
// The solve runs three candidate factorisations at once, one per wave of the block
// (solveWave): wave 0 the certificate A - s I, wave 1 A + base I, wave 2 A + 1.0 I;
// waves 1 and 2 also solve for their step. One lane then picks the candidate the damping
// rule selects (icpSolve).
struct SolveLanes {
    int ok[3];
    int finite;
    double x[3][6];
};

__device__ void buildSystem(const double* sums, double* A, double* b) {
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            A[i * 6 + j] = sums[k];
            A[j * 6 + i] = sums[k];
            ++k;
        }
    for (int i = 0; i < 6; ++i) b[i] = sums[21 + i];
}

// Slot of A[i][j] among the 21 upper-triangle sums (row-major upper triangle).
__device__ __forceinline__ int symSlot(int i, int j) {
    const int r = i < j ? i : j, c = i < j ? j : i;
    return r * 6 - (r * (r - 1)) / 2 + (c - r);
}

__device__ __forceinline__ double readlaneD(double v, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Candidate `cand` (shift, Cholesky factor L L^T of the shifted A, and for cand > 0 the
// step from L y = b, L^T x = y) by one wave, lane i (< 6) holding row i of A and of L:
// column j's pivot s = A[j][j] - sum_k L[j][k]^2 is formed on lane j and the entries below
// it, A[i][j] - sum_k L[i][k] L[j][k] (k ascending), on lanes i > j with row j's entries
// broadcast from lane j; the pivot's 1 / sqrt(s) comes from the hardware estimate and one
// Newton step (relative error ~1e-16, far below the float step this solve produces), and
// sqrt(s) = s / sqrt(s). The forward substitution broadcasts y[j] as it is formed; the
// backward one (x[i] = (y[i] - sum_{j > i} L[j][i] x[j]) / L[i][i], j ascending) runs on
// broadcast values in every lane. Every element takes the same operations in the same
// order as a one-lane evaluation of those formulas (so the results are bit-identical to
// it), and a pivot s <= 0 leaves ok false.
__device__ void solveWave(const double* sums, int level, int cand, int lane, SolveLanes* out) {
    const double kMin = (double)1e-4f, kCond = (double)1e7f;
    const int i = lane < 6 ? lane : 5;  // lanes past 5 repeat row 5 (unused)
    double a[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) a[j] = sums[symSlot(i, j)];
    double shift;
    if (cand == 0) {
        // Gershgorin bound, rows in order: r_i = A[i][i] + sum_{j != i} |A[i][j]|.
        double diag = a[0];
#pragma unroll
        for (int j = 1; j < 6; ++j) diag = i == j ? a[j] : diag;
        double r = diag;
#pragma unroll
        for (int j = 0; j < 6; ++j) r = j != i ? r + fabs(a[j]) : r;
        double g = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) g = fmax(g, readlaneD(r, k));
        shift = -(fmax(kMin, g / kCond) * (1.0 + 1e-12));
        // Every A entry finite (the 21 sums).
        const unsigned long long fin = __ballot(lane < 21 && isfinite(sums[lane < 21 ? lane : 0]));
        if (lane == 0) out->finite = fin == 0x1FFFFFull ? 1 : 0;
    } else {
        shift = cand == 1 ? (level == 0 ? (double)0.01f : (double)0.1f) : 1.0;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) a[j] = i == j ? a[j] + shift : a[j];
    // Factor: column j's pivot s on lane j, the entries below it on lanes i > j.
    double L[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, rd[6];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double t = a[j];
#pragma unroll
        for (int k = 0; k < j; ++k) t -= L[k] * readlaneD(L[k], j);
        const double sj = readlaneD(t, j);
        ok = ok && sj > 0.0;
        const double r0 = __builtin_amdgcn_rsq(sj);
        const double inv = r0 * (1.5 - (0.5 * sj) * (r0 * r0));
        const double d = sj * inv;
        L[j] = i == j ? d : (i > j ? t * inv : 0.0);
        rd[j] = inv;
    }
    if (cand > 0 && ok) {
        double t = sums[21 + i], y[6], x[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            y[j] = readlaneD(t, j) * rd[j];
            t = i > j ? t - L[j] * y[j] : t;
        }
#pragma unroll
        for (int ii = 5; ii >= 0; --ii) {
            double u = y[ii];
#pragma unroll
            for (int j = ii + 1; j < 6; ++j) u -= readlaneD(L[ii], j) * x[j];
            x[ii] = u * rd[ii];
        }
        if (lane < 6) {
            double v = x[0];
#pragma unroll
            for (int k = 1; k < 6; ++k) v = lane == k ? x[k] : v;
            out->x[cand][lane] = v;
        }
    }
    if (lane == 0) out->ok[cand] = ok ? 1 : 0;
}

// Block 0's record of the solved system in the state (information = A in float, zero at
// a level's first iteration when it has fewer than 10 inliers; the gate counters), one
// entry per lane of one wave.
__device__ void recordSystem(IcpState* st, const double* sums, int iter, int lane) {
    const bool enough = (int)sums[kIcpSlotsF] >= 10;
    if (lane < 36 && (enough || iter == 0))
        st->information[lane] = enough ? (float)sums[symSlot(lane / 6, lane % 6)] : 0.0f;
    if (lane >= 36 && lane < 41) st->counters[lane - 36] = (int)sums[kIcpSlotsF + 1 + (lane - 36)];
}

// Nearest rotation by one-sided Jacobi SVD in FP64, U V^T with U's smallest-singular-
// value column negated if det < 0 (projectToSO3, ICPShared.h:102-110). General fallback.
__device__ void projectSO3Svd(const double* M, double* R) {
    double U[9], V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; ++i) U[i] = M[i];
    for (int sweep = 0; sweep < 60; ++sweep) {
        double offmax = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = 0; k < 3; ++k) {
                    al += U[k * 3 + p] * U[k * 3 + p];
                    be += U[k * 3 + q] * U[k * 3 + q];
                    ga += U[k * 3 + p] * U[k * 3 + q];
                }
                const double den = sqrt(al * be);
                if (den == 0.0) continue;
                offmax = fmax(offmax, fabs(ga) / den);
                if (fabs(ga) <= 1e-300) continue;
                const double ze = (be - al) / (2.0 * ga);
                const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
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
        sig[j] = sqrt(U[j] * U[j] + U[3 + j] * U[3 + j] + U[6 + j] * U[6 + j]);
        if (sig[j] > 0.0)
            for (int k = 0; k < 3; ++k) U[k * 3 + j] /= sig[j];
    }
    for (int j = 0; j < 3; ++j) {
        if (sig[j] > 0.0) continue;
        const int a = (j + 1) % 3, b = (j + 2) % 3;
        const double ca[3] = {U[a], U[3 + a], U[6 + a]}, cb[3] = {U[b], U[3 + b], U[6 + b]};
        const double c[3] = {ca[1] * cb[2] - ca[2] * cb[1], ca[2] * cb[0] - ca[0] * cb[2], ca[0] * cb[1] - ca[1] * cb[0]};
        const double l = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; ++k) U[k * 3 + j] = l > 0.0 ? c[k] / l : (k == j ? 1.0 : 0.0);
    }
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                R[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] + U[i * 3 + 2] * V[j * 3 + 2];
        const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                           R[2] * (R[3] * R[7] - R[4] * R[6]);
        if (pass == 1 || !(det < 0.0)) break;
        int js = 0;
        for (int j = 1; j < 3; ++j)
            if (sig[j] < sig[js]) js = j;
        for (int k = 0; k < 3; ++k) U[k * 3 + js] = -U[k * 3 + js];
    }
}

// The same projection for the matrices the update produces (a rotation times a
// rotation, orthogonal up to float rounding): the polar factor by Newton's iteration
// X <- (X + X^-T) / 2 (Higham, Functions of Matrices, 8.3); it equals U V^T when
// det > 0. Anything else takes the SVD route. With X = U (I + E), E symmetric, the step
// leaves U (I + E^2 / 2 + O(E^3)) and moves X by about |E|: once a step moves no entry
// by 1e-9, the iterate it produced is within FP64 rounding of U (E^2 / 2 < 1e-18).
// From float-rounded input (|E| ~ 1e-7) that is two steps.
__device__ void projectSO3(const double* M, double* R) {
    double X[9];
    for (int i = 0; i < 9; ++i) X[i] = M[i];
    for (int it = 0; it < 8; ++it) {
        const double c00 = X[4] * X[8] - X[5] * X[7], c01 = X[5] * X[6] - X[3] * X[8], c02 = X[3] * X[7] - X[4] * X[6];
        const double c10 = X[2] * X[7] - X[1] * X[8], c11 = X[0] * X[8] - X[2] * X[6], c12 = X[1] * X[6] - X[0] * X[7];
        const double c20 = X[1] * X[5] - X[2] * X[4], c21 = X[2] * X[3] - X[0] * X[5], c22 = X[0] * X[4] - X[1] * X[3];
        const double det = X[0] * c00 + X[1] * c01 + X[2] * c02;
        if (!(det > 0.5) || !isfinite(det)) {
            projectSO3Svd(M, R);
            return;
        }
        const double inv = 1.0 / det;  // X^-T = cofactor(X) / det
        const double Y[9] = {c00 * inv, c01 * inv, c02 * inv, c10 * inv, c11 * inv, c12 * inv,
                             c20 * inv, c21 * inv, c22 * inv};
        double dmax = 0.0;
        for (int i = 0; i < 9; ++i) {
            const double nx = 0.5 * (X[i] + Y[i]);
            dmax = fmax(dmax, fabs(nx - X[i]));
            X[i] = nx;
        }
        if (dmax < 1e-9) break;
    }
    for (int i = 0; i < 9; ++i) R[i] = X[i];
}


// Rodrigues' rotation of a rotation vector w with |w| <= 0.5 (the step cap), in FP64:
// R = cos t I + a [w]x + b w w^T with t = |w|, a = sin t / t and b = (1 - cos t) / t^2
// summed as their Taylor series in t^2 (seven terms: the first omitted one is below
// 0.25^7 / 15! < 5e-17), cos t = 1 - t^2 b. No square root, division or trigonometric
// call; the same matrix as the axis-angle form cI + s[k]x + (1 - c) k k^T to FP64 rounding.
__device__ __forceinline__ void rodrigues64(double wx, double wy, double wz, double* R) {
    const double t2 = wx * wx + wy * wy + wz * wz;
    const double a = 1.0 + t2 * (-1.0 / 6.0 + t2 * (1.0 / 120.0 + t2 * (-1.0 / 5040.0 + t2 * (1.0 / 362880.0 +
                     t2 * (-1.0 / 39916800.0 + t2 * (1.0 / 6227020800.0))))));
    const double b = 0.5 + t2 * (-1.0 / 24.0 + t2 * (1.0 / 720.0 + t2 * (-1.0 / 40320.0 + t2 * (1.0 / 3628800.0 +
                     t2 * (-1.0 / 479001600.0 + t2 * (1.0 / 87178291200.0))))));
    const double c = 1.0 - t2 * b;
    R[0] = c + b * wx * wx;
    R[4] = c + b * wy * wy;
    R[8] = c + b * wz * wz;
    R[1] = b * wx * wy - a * wz;
    R[3] = b * wx * wy + a * wz;
    R[2] = b * wx * wz + a * wy;
    R[6] = b * wx * wz - a * wy;
    R[5] = b * wy * wz - a * wx;
    R[7] = b * wy * wz + a * wx;
}

// The projection onto SO(3) of the products the update forms (a rotation times a
// rotation, orthogonal to float rounding: |M^T M - I| <= 1e-6): one Newton-Schulz step
// X = M (3I - M^T M) / 2 (Higham, Functions of Matrices, 8.3), which leaves an error of
// about 1.5 |M^T M - I|^2 < 2e-12, far below the float rounding of the result; the
// orthogonal polar factor it converges to is U V^T. Anything else (a reflection, a
// matrix further from orthogonal) takes projectSO3.
__device__ __forceinline__ void projectSO3Near(const double* M, double* R) {
    double G[9];  // M^T M
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) G[i * 3 + j] = M[i] * M[j] + M[3 + i] * M[3 + j] + M[6 + i] * M[6 + j];
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) e = fmax(e, fabs(G[i] - ((i % 4 == 0) ? 1.0 : 0.0)));
    const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                       M[2] * (M[3] * M[7] - M[4] * M[6]);
    if (!(e <= 1e-6) || !(det > 0.0)) {
        projectSO3(M, R);
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double s0 = ((j == 0) ? 3.0 : 0.0) - G[0 * 3 + j];
            const double s1 = ((j == 1) ? 3.0 : 0.0) - G[1 * 3 + j];
            const double s2 = ((j == 2) ? 3.0 : 0.0) - G[2 * 3 + j];
            R[i * 3 + j] = 0.5 * (M[i * 3 + 0] * s0 + M[i * 3 + 1] * s1 + M[i * 3 + 2] * s2);
        }
}

// The pose terms of a pose (column-major 4 x 4) and the inverse of the reference pose.
__device__ PoseDerived derivePose(const float* pose, const Rigid& rinv, int iterations) {
    const Rigid P = rigidFromPtr(pose);
    PoseDerived d;
    d.R_cw = P.R;
    d.t_cw = P.t;
    d.R_rel = mulm(rinv.R, P.R);
    d.t_rel = f3add(mulv(rinv.R, P.t), rinv.t);
    d.iterations = iterations;
    return d;
}

// One solve + pose update (trackLevel's loop body, ICPTracker.cpp:79-130), lane 0 of
// wave 0 once solveWave has run on waves 0-2; every block of a launch runs it on the same
// sums and gets the same result, and the writer (block 0) also records it in the state.
// It reads no global memory (ref_inv: the state's, read at the launch's start).
// sums: 28 FP64 slots (21 A upper, 6 b, residual) + 6 counters. out: the pose terms after
// the step (pd when there is none); returns true when the level's loop ends.
__device__ bool icpSolve(IcpState* st, bool writer, const double* sums, const SolveLanes* sl, const PoseDerived& pd,
                         const Rigid& ref_inv, int level, int iter, PoseDerived& out) {
    out = pd;
    // (The information matrix and the gate counters are recorded by recordSystem.)
    if (writer && iter == 0) {
        st->error = 0.0f;
        st->inliers = 0;
        st->converged = 0;
        st->final_step = __builtin_inff();
    }
    const int inliers = (int)sums[kIcpSlotsF + 0];
    if (inliers < 10) {  // covers buildLinearSystem's "no inlier" failure too
        if (writer) st->stop_level = level;
        return true;
    }
    const double residual = sums[27];
    int pick = 1;  // the level's base lambda
    if (!sl->finite) {
        pick = 2;
    } else if (!sl->ok[0]) {
        const double kMin = (double)1e-4f, kCond = (double)1e7f;
        double A[36], b[6], mn, mx;
        buildSystem(sums, A, b);
        jacobiMinMax(A, mn, mx);
        if (!isfinite(mn) || !isfinite(mx) || mn < kMin || mx > kCond * mn) pick = 2;
    }
    if (!sl->ok[pick]) {
        if (writer) st->stop_level = level;
        return true;
    }
    double xd[6];
    for (int i = 0; i < 6; ++i) xd[i] = sl->x[pick][i];
    float x[6];
    bool fin = true;
    for (int i = 0; i < 6; ++i) {
        x[i] = (float)xd[i];
        fin = fin && isfinite(x[i]);
    }
    if (!fin) {
        if (writer) st->stop_level = level;
        return true;
    }
    const float tn = sqrtf(sum3(x[0] * x[0], x[1] * x[1], x[2] * x[2]));
    const float rn = sqrtf(sum3(x[3] * x[3], x[4] * x[4], x[5] * x[5]));
    if (tn > 0.2f || rn > 0.5f) {
        if (writer) st->stop_level = level;
        return true;
    }
    float sq = 0.0f;
    for (int i = 0; i < 6; ++i) sq += x[i] * x[i];
    const float step_norm = sqrtf(sq);
    // pose <- pose * [R(w) | t] (ICPTracker.cpp:100-121): R(w) by Rodrigues' formula for
    // |w| > 1e-4 and the first-order I + [w]x below, as the reference's AngleAxis branch
    // does; the rotation product and its projection in FP64.
    const double rx = x[3], ry = x[4], rz = x[5];
    double Rd[9] = {1.0, -rz, ry, rz, 1.0, -rx, -ry, rx, 1.0};
    if (rx * rx + ry * ry + rz * rz > 1e-8) rodrigues64(rx, ry, rz, Rd);
    // The pose the solved iteration started from.
    const float t[3] = {pd.t_cw.x, pd.t_cw.y, pd.t_cw.z};
    double Md[9];
    float nt[3];
    for (int i = 0; i < 3; ++i) {
        const double p0 = pd.R_cw.m[i][0], p1 = pd.R_cw.m[i][1], p2 = pd.R_cw.m[i][2];
        for (int j = 0; j < 3; ++j) Md[i * 3 + j] = p0 * Rd[j] + p1 * Rd[3 + j] + p2 * Rd[6 + j];
        nt[i] = ((pd.R_cw.m[i][0] * x[0] + pd.R_cw.m[i][1] * x[1]) + pd.R_cw.m[i][2] * x[2]) + t[i];
    }
    double Rp[9];
    projectSO3Near(Md, Rp);
    float np[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) np[j * 4 + i] = (float)Rp[i * 3 + j];
        np[12 + i] = nt[i];
    }
    out = derivePose(np, ref_inv, pd.iterations + (1 << (8 * level)));
    const bool converged = step_norm < 5e-5f;
    if (writer) {
        // Rows 0-2; the bottom row stays as it was.
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 3; ++i) st->pose[j * 4 + i] = np[j * 4 + i];
        st->final_step = step_norm;
        st->last_final_step = step_norm;
        st->iterations = pd.iterations + (1 << (8 * level));
        st->error = (float)(residual / (double)(inliers > 1 ? inliers : 1));
        st->inliers = inliers;
        if (converged) {
            st->converged = 1;
            st->stop_level = level;
        }
    }
    return converged;
}

__global__ void icpInit(IcpState* st, PoseArg est, PoseArg ref) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    // Both poses are read before any write: either may point into the state itself.
    float e[16], r[16];
    for (int i = 0; i < 16; ++i) {
        e[i] = poseEl(est, i);
        r[i] = poseEl(ref, i);
    }
    for (int i = 0; i < 16; ++i) {
        st->pose[i] = e[i];
        st->ref[i] = r[i];
    }
    for (int i = 0; i < 36; ++i) st->information[i] = 0.0f;
    st->error = 0.0f;
    st->final_step = __builtin_inff();
    st->last_final_step = __builtin_inff();
    st->inliers = 0;
    st->converged = 0;
    for (int i = 0; i < 5; ++i) st->counters[i] = 0;
    st->stop_level = -1;
    st->iterations = 0;
    const Rigid rinv = rigidInverse(rigidFromPtr(r));
    st->ref_inv = rinv;
    for (int p = 0; p < 2; ++p) {
        st->it[p].pd = derivePose(e, rinv, 0);
        st->it[p].pending = 0;
    }
}

__global__ void icpClearTicks(IcpState* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        st->ticks_sum = st->ticks_solve = st->ticks_n = st->work_px = st->work_proj = 0;
}

// A * B for rigid 4x4 column-major poses, each entry summed left to right
// ((a0 b0 + a1 b1) + a2 b2) + a3 b3 with the implicit bottom row (0, 0, 0, 1).
__device__ void composeRigid(const float* A, const float* B, float* C) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) {
            const float s = ((A[0 * 4 + r] * B[c * 4 + 0] + A[1 * 4 + r] * B[c * 4 + 1]) + A[2 * 4 + r] * B[c * 4 + 2]) +
                            A[3 * 4 + r] * (c == 3 ? 1.0f : 0.0f);
            C[c * 4 + r] = s;
        }
    C[3] = C[7] = C[11] = 0.0f;
    C[15] = 1.0f;
}

// Constant-velocity motion model of the reference pipeline
// (src/app/PipelineController.cpp:1162-1187): predicted = cur * (last^-1 * cur), then
// last = cur. The inverse is the rigid one.
__global__ void predictPose(const float* __restrict__ cur, float* __restrict__ last, float* __restrict__ pred) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float c[16], l[16], li[16], rel[16], p[16];
    for (int i = 0; i < 16; ++i) {
        c[i] = cur[i];
        l[i] = last[i];
    }
    const Rigid L = rigidFromPtr(l);
    const Rigid Li = rigidInverse(L);
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) li[k * 4 + r] = Li.R.m[r][k];
        li[12 + r] = r == 0 ? Li.t.x : (r == 1 ? Li.t.y : Li.t.z);
    }
    li[3] = li[7] = li[11] = 0.0f;
    li[15] = 1.0f;
    composeRigid(li, c, rel);
    composeRigid(c, rel, p);
    for (int i = 0; i < 16; ++i) {
        pred[i] = p[i];
        last[i] = c[i];
    }
}


// The per-pixel part of buildLinearSystem (ICPTracker.cpp:236-327), in two stages so a
// thread can issue the memory reads of several pixels together: icpProject takes a live
// vertex to its model pixel (or -1), icpAccumulate runs the remaining gates on the
// gathered model vertex / normal and live normal and folds the pixel's terms into
// (hi, lo) and its gate outcome into cnt.
__device__ __forceinline__ int icpProject(const IcpIterArgs& a, const PoseDerived& pd, F3 lv, int* cnt) {
    if (lv.z <= 0.001f || !f3finite(lv)) return -1;
    const F3 vr = f3add(mulv(pd.R_rel, lv), pd.t_rel);
    if (vr.z <= 0.001f) return -1;
    cnt[1]++;
    const float inv_z = 1.0f / vr.z;
    const float mxf = a.Km.fx * vr.x * inv_z + a.Km.cx + 0.5f;
    const float myf = a.Km.fy * vr.y * inv_z + a.Km.cy + 0.5f;
    // floorToInt (CoordinateMath.h:27-49): finite and representable, then the bounds.
    if (!isfinite(mxf) || !isfinite(myf)) return -1;
    const float fmx = floorf(mxf), fmy = floorf(myf);
    if (!(fmx >= 0.0f && fmx < (float)a.MW && fmy >= 0.0f && fmy < (float)a.MH)) return -1;
    cnt[3]++;
    return (int)fmy * a.MW + (int)fmx;
}

__device__ __forceinline__ void icpAccumulate(const IcpIterArgs& a, const PoseDerived& pd, F3 lv, F3 mv, F3 mn,
                                              F3 ln, float* acc, float* comp, int* cnt) {
    if (!(f3finite(mv) && f3dot(mv, mv) > 1e-12f) || !(f3finite(mn) && f3dot(mn, mn) > 0.9f)) return;
    cnt[2]++;
    const F3 vlw = f3add(mulv(pd.R_cw, lv), pd.t_cw);
    const F3 dd = f3sub(vlw, mv);
    if (f3dot(dd, dd) > a.dist_sq) {
        cnt[4]++;
        return;
    }
    if (!(f3finite(ln) && f3dot(ln, ln) > 0.9f)) return;
    const F3 nlw = mulv(pd.R_cw, ln);
    if (fabsf(f3dot(nlw, mn)) < a.angle_cos) {
        cnt[5]++;
        return;
    }
    const float err = f3dot(mn, dd);
    if (!isfinite(err)) return;
    const F3 nlc = mulvT(pd.R_cw, mn);
    const F3 cr = f3cross(lv, nlc);
    const float J[6] = {nlc.x, nlc.y, nlc.z, cr.x, cr.y, cr.z};
    if (!(isfinite(J[0]) && isfinite(J[1]) && isfinite(J[2]) && isfinite(J[3]) && isfinite(J[4]) && isfinite(J[5])))
        return;
    const float ae = fabsf(err);
    const float w = (ae <= 0.02f) ? 1.0f : 0.02f / ae;
    const float we = err * w;
    const float loss = (ae <= 0.02f) ? ae * ae : 2.0f * 0.02f * ae - 0.02f * 0.02f;
    int k = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float wj = w * J[i];
#pragma unroll
        for (int j = i; j < 6; ++j) {
            kahanAdd(acc[k], comp[k], wj * J[j]);
            ++k;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) kahanAdd(acc[21 + i], comp[21 + i], -(J[i] * we));
    kahanAdd(acc[27], comp[27], loss);
    cnt[0]++;
}

__device__ __forceinline__ F3 load3(const float* __restrict__ p, int i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

constexpr int kPxBatch = 5;   // pixels whose reads one thread issues together

// Blocks an iteration over a W x H level is split into: one per kIcpThreads interior pixels,
// at most kIcpMaxBlocks.
__host__ __device__ int icpBlocksFor(int W, int H) {
    const long long npx = (long long)(W > 2 ? W - 2 : 0) * (H > 2 ? H - 2 : 0);
    long long b = (npx + kIcpThreads - 1) / kIcpThreads;
    if (b < 1) b = 1;
    if (b > kIcpMaxBlocks) b = kIcpMaxBlocks;
    return (int)b;
}

// One batch of a thread's live pixels: level indices (-1 past the level's end) and the
// live vertices and normals there.
struct PassBatch {
    int idx[kPxBatch];
    F3 lv[kPxBatch], ln[kPxBatch];
};

// Pixels p0, p0 + stride, ... of the level's interior; every read is issued before any
// use.
__device__ __forceinline__ void loadBatch(const IcpIterArgs& a, int p0, int stride, PassBatch& pb) {
    const int iw = a.W - 2, ih = a.H - 2;
    const int npx = (iw > 0 && ih > 0) ? iw * ih : 0;
#pragma unroll
    for (int b = 0; b < kPxBatch; ++b) {
        const int p = p0 + b * stride;
        pb.idx[b] = p < npx ? (1 + p / iw) * a.W + 1 + p % iw : -1;
    }
#pragma unroll
    for (int b = 0; b < kPxBatch; ++b) pb.lv[b] = load3(a.lv, pb.idx[b] < 0 ? 0 : pb.idx[b]);
#pragma unroll
    for (int b = 0; b < kPxBatch; ++b) pb.ln[b] = load3(a.ln, pb.idx[b] < 0 ? 0 : pb.idx[b]);
}

struct IcpShared {
    float wsum[kIcpWaves][kIcpSlotsF];  // per-wave sums
    int cnt[kIcpWaves][kIcpSlotsI];
    double grp[kSumGroups][kIcpSlots];
    double sum[kIcpSlots];
    SolveLanes lanes;
    PoseDerived pd;
    int stop;
};


// One block's share of an iteration: the pixel pass over its pixels, the block fold and
// its FP64 partials (written by wave 0). The level is split over the launch's blocks;
// first: the thread's first batch (loadBatch at p0 = blockIdx.x * kIcpThreads + tid, stride =
// gridDim.x * kIcpThreads), already requested.
__device__ void icpBlockPass(const IcpIterArgs& a, const PoseDerived& pd, double* __restrict__ partials,
                             IcpShared& sh, const PassBatch& first) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    float acc[kIcpSlotsF], comp[kIcpSlotsF];
#pragma unroll
    for (int i = 0; i < kIcpSlotsF; ++i) {
        acc[i] = 0.0f;
        comp[i] = 0.0f;
    }
    int cnt[kIcpSlotsI] = {0, 0, 0, 0, 0, 0};
    const int iw = a.W - 2, ih = a.H - 2;
    const int npx = (iw > 0 && ih > 0) ? iw * ih : 0;
    const int stride = (int)gridDim.x * kIcpThreads;
    // Each thread visits pixels p0, p0 + stride, ... in that order, kPxBatch at a time:
    // the batch's live reads, then all of its model gathers, are issued before their use.
    PassBatch pb = first;
    const int p_first = blockIdx.x * kIcpThreads + tid;
    for (int p0 = p_first; p0 < npx; p0 += kPxBatch * stride) {
        if (p0 != p_first) loadBatch(a, p0, stride, pb);
        int midx[kPxBatch];
#pragma unroll
        for (int b = 0; b < kPxBatch; ++b) midx[b] = pb.idx[b] < 0 ? -1 : icpProject(a, pd, pb.lv[b], cnt);
        F3 mv[kPxBatch], mn[kPxBatch];
#pragma unroll
        for (int b = 0; b < kPxBatch; ++b) {
            const int m = midx[b] < 0 ? 0 : midx[b];
            mv[b] = load3(a.mv, m);
            mn[b] = load3(a.mn, m);
        }
#pragma unroll
        for (int b = 0; b < kPxBatch; ++b)
            if (midx[b] >= 0) icpAccumulate(a, pd, pb.lv[b], mv[b], mn[b], pb.ln[b], acc, comp, cnt);
    }

    // Block fold: each thread's Kahan sum (s - c), a fixed DPP tree per wave, then the
    // wave sums in FP64 in wave order. Counters are exact integers.
    float wsum[kIcpSlotsF];
#pragma unroll
    for (int i = 0; i < kIcpSlotsF; ++i) wsum[i] = waveSumF(acc[i] - comp[i]);
#pragma unroll
    for (int i = 0; i < kIcpSlotsI; ++i) cnt[i] = waveSumInt(cnt[i]);
    if (lane == 63) {
#pragma unroll
        for (int i = 0; i < kIcpSlotsF; ++i) sh.wsum[wv][i] = wsum[i];
#pragma unroll
        for (int i = 0; i < kIcpSlotsI; ++i) sh.cnt[wv][i] = cnt[i];
    }
    __syncthreads();
    if (tid < kIcpSlots) {
        double v;
        if (tid < kIcpSlotsF) {
            v = (double)sh.wsum[0][tid];
#pragma unroll
            for (int w = 1; w < kIcpWaves; ++w) v += (double)sh.wsum[w][tid];
        } else {
            const int i = tid - kIcpSlotsF;
            int n = 0;
#pragma unroll
            for (int w = 0; w < kIcpWaves; ++w) n += sh.cnt[w][i];
            v = (double)n;
        }
        // Block-major (coalesced both ways); the next launch reads them.
        partials[(size_t)blockIdx.x * kIcpSlots + tid] = v;
    }
}

// Cross-block sum, first step: thread (g, s) reads slot s of blocks g, g + 7, g + 14, ...
// of the previous launch's partials (threads past 7 * 34: zeros). Reads past the
// previous launch's nblk blocks go to the zero row kIcpMaxBlocks, which no launch
// writes, so every read is unconditional and all are in flight at once.
constexpr int kMaxPer = (kIcpMaxBlocks + kSumGroups - 1) / kSumGroups;
constexpr int kPartialRows = kIcpMaxBlocks + 1;
__device__ __forceinline__ void loadPartials(const double* __restrict__ partials, int nblk, double* x) {
    const int tid = threadIdx.x;
    const int s = tid % kIcpSlots, g = tid / kIcpSlots;
    const unsigned nb = tid < kSumGroups * kIcpSlots ? (unsigned)nblk : 0u;
    const double* src = partials + s;
#pragma unroll
    for (int i = 0; i < kMaxPer; ++i) {
        const unsigned bi = (unsigned)(g + kSumGroups * i);
        x[i] = src[(size_t)(bi < nb ? bi : (unsigned)kIcpMaxBlocks) * kIcpSlots];
    }
}

// The previous iteration's solve, in every block: FP64 cross-block sum of its partials
// (gsum: the thread's loadPartials values added in block order; the seven group sums
// are then added in group order), then the solve. The order is fixed, so every block
// gets the same sums, and the same solve. Returns true when the level's loop ends; the
// new pose terms are left in sh.pd.
__device__ bool blockSolve(IcpState* __restrict__ st, double gsum, IcpShared& sh, const PoseDerived& pd,
                           const Rigid& ref_inv, int level, int iter, bool writer, unsigned long long* ticks) {
    const int tid = threadIdx.x;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    if (tid < kSumGroups * kIcpSlots) sh.grp[tid / kIcpSlots][tid % kIcpSlots] = gsum;
    __syncthreads();
    if (tid < kIcpSlots) {
        double acc = 0.0;
        for (int g = 0; g < kSumGroups; ++g) acc += sh.grp[g][tid];
        sh.sum[tid] = acc;
    }
    __syncthreads();
    // The three candidate factorisations on waves 0-2 at once, block 0's record of the
    // system on wave 3, then the rest of the solve on one lane.
    const unsigned long long t_sum = __builtin_amdgcn_s_memrealtime();
    const int wv = tid >> 6, lane = tid & 63;
    if (wv < 3 && (int)sh.sum[kIcpSlotsF] >= 10) solveWave(sh.sum, level, wv, lane, &sh.lanes);
    if (wv == 3 && writer) recordSystem(st, sh.sum, iter, lane);
    __syncthreads();
    if (tid == 0) {
        PoseDerived out;
        sh.stop = icpSolve(st, writer, sh.sum, &sh.lanes, pd, ref_inv, level, iter, out) ? 1 : 0;
        sh.pd = out;
        ticks[0] = t_sum - t0;
        ticks[1] = __builtin_amdgcn_s_memrealtime() - t_sum;
    }
    __syncthreads();
    return sh.stop != 0;
}

// One launch per ICP iteration, plus a final solve-only launch: first the previous
// iteration's solve (every block, from the partials the previous launch wrote; the kernel
// boundary orders them, so no block waits on another), then this iteration's pixel pass
// into fresh partials. prev.level < 0: nothing to solve; a.level < 0: no pass.
struct IcpPrev {
    int level, iter, npx, nblk;
};

__global__ __launch_bounds__(kIcpThreads) void icpIter(IcpIterArgs a, IcpPrev prev, IcpState* __restrict__ st,
                                               const double* __restrict__ part_prev, double* __restrict__ part_cur,
                                               int par) {
    __shared__ IcpShared sh;
    const int tid = threadIdx.x;
    const bool writer = blockIdx.x == 0;
    // Every kernel argument is in registers before the first state read: one scalar
    // wait covers all scalar loads in flight, so an argument first loaded after the
    // state read was issued would wait out the state's round trip too. The empty asm
    // statements pass the level through as if they depended on every argument, and the
    // level decides the first branch.
    int level = a.level;
    asm("" : "+s"(level) : "s"(a.lv), "s"(a.ln), "s"(a.mv), "s"(a.mn), "s"(a.W), "s"(a.H), "s"(a.MW), "s"(a.MH));
    asm("" : "+s"(level) : "s"(a.Km.fx), "s"(a.Km.fy), "s"(a.Km.cx), "s"(a.Km.cy), "s"(a.angle_cos), "s"(a.dist_sq),
        "s"(a.iter), "s"((int)gridDim.x));
    asm("" : "+s"(level) : "s"(prev.level), "s"(prev.iter), "s"(prev.npx), "s"(prev.nblk), "s"(st), "s"(part_prev),
        "s"(part_cur), "s"(par));
    // Read at the start: the stop level (block 0 of this launch may set it, to a value
    // every block also reaches by its own solve) and this launch's slot (written by the
    // launch before; block 0 writes the other one).
    // The first launch of a level always runs its pass (a level stops only in its own
    // launches); its first batch of live pixels is requested before anything else.
    // Later launches of a level request it once the state shows the level still running
    // (otherwise the launch has nothing to do), in time to arrive under the solve.
    const bool level_start = level >= 0 && level != prev.level;
    const int p_first = blockIdx.x * kIcpThreads + tid, stride = (int)gridDim.x * kIcpThreads;
    PassBatch first;
    if (level_start) loadBatch(a, p_first, stride, first);
    const int stop = st->stop_level;
    const int pending = st->it[par].pending;
    PoseDerived pd = st->it[par].pd;
    const Rigid ref_inv = st->ref_inv;
    // The previous launch's partials (used only if it ran a pass) are requested before
    // the state arrives.
    // (Only the threads of the sum's groups read them; the waves past them skip the reads.)
    double x[kMaxPer];
#pragma unroll
    for (int i = 0; i < kMaxPer; ++i) x[i] = 0.0;
    if (prev.level >= 0 && tid < (kSumGroups * kIcpSlots + 63) / 64 * 64) loadPartials(part_prev, prev.nblk, x);
    const bool late = !level_start && a.level >= 0 && stop != a.level;
    bool stopped = false;
    const bool solve = pending && prev.level >= 0;
    unsigned long long ticks[2] = {0, 0};
    if (solve) {
        // The thread's share of the sum is taken before the late request of the live
        // pixels, so that waiting for the partials does not wait for those too (the
        // empty asm keeps the sum here).
        double gsum = x[0];
#pragma unroll
        for (int i = 1; i < kMaxPer; ++i) gsum += x[i];
        asm volatile("" ::"v"(gsum));
        if (late) loadBatch(a, p_first, stride, first);
        stopped = blockSolve(st, gsum, sh, pd, ref_inv, prev.level, prev.iter, writer, ticks);
        pd = sh.pd;
    } else if (late) {
        loadBatch(a, p_first, stride, first);
    }
    const bool pass = a.level >= 0 && stop != a.level && !(stopped && prev.level == a.level);
    if (writer && tid == 0) {
        st->it[1 - par].pd = pd;
        st->it[1 - par].pending = pass ? 1 : 0;
        if (solve) {
            // Diagnostics, after the block's last barrier of the solve: atomics whose
            // result is unused, so nothing waits on them before the pass.
            __hip_atomic_fetch_add(&st->ticks_sum, ticks[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&st->ticks_solve, ticks[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&st->ticks_n, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&st->work_px, (unsigned long long)prev.npx, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&st->work_proj, (unsigned long long)sh.sum[kIcpSlotsF + 3], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!pass) return;
    icpBlockPass(a, pd, part_cur, sh, first);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
size_t icpStateBytes() { return sizeof(IcpState); }

// Two halves (the launch's own and the previous launch's partials), each with a zero row
// after its kIcpMaxBlocks block rows; the buffer is zeroed once at allocation.
size_t icpPartialsBytes(int) { return 2 * sizeof(double) * kIcpSlots * (size_t)kPartialRows; }

hipError_t launchIcpInit(void* st, const PoseArg& est, const PoseArg& ref, hipStream_t s) {
    hipLaunchKernelGGL(icpInit, dim3(1), dim3(64), 0, s, (IcpState*)st, est, ref);
    return hipGetLastError();
}

// Overwrites the tracked pose (the state's first 16 floats) with a host pose passed by
// value, leaving the motion model's previous pose and the last result's other fields.
__global__ void writeTrackedPose(float* __restrict__ dst, PoseArg p) {
    if (blockIdx.x == 0 && threadIdx.x < 16) dst[threadIdx.x] = p.m[threadIdx.x];
}

hipError_t launchWritePose(void* st, const PoseArg& p, hipStream_t s) {
    hipLaunchKernelGGL(writeTrackedPose, dim3(1), dim3(64), 0, s, static_cast<float*>(st), p);
    return hipGetLastError();
}

hipError_t launchPredict(const float* cur, float* last, float* pred, hipStream_t s) {
    hipLaunchKernelGGL(predictPose, dim3(1), dim3(64), 0, s, cur, last, pred);
    return hipGetLastError();
}

// Launch k of a track (k = 0, 1, ...): level < 0 is the final, solve-only launch (one
// block); prev_level < 0: no previous iteration. partials: icpPartialsBytes.
hipError_t launchIcpIter(const float* lv, const float* ln, int W, int H, const float* mv, const float* mn, int MW,
                         int MH, Intr Km, float angle_cos, float dist_sq, int level, int iter, int prev_level,
                         int prev_iter, int prev_W, int prev_H, int k, void* st, double* partials, hipStream_t s) {
    IcpIterArgs a;
    a.lv = lv;
    a.ln = ln;
    a.W = W;
    a.H = H;
    a.mv = mv;
    a.mn = mn;
    a.MW = MW;
    a.MH = MH;
    a.Km = Km;
    a.angle_cos = angle_cos;
    a.dist_sq = dist_sq;
    a.level = level;
    a.iter = iter;
    IcpPrev prev{prev_level, prev_iter, 0, 1};
    if (prev_level >= 0) {
        prev.npx = (prev_W > 2 && prev_H > 2) ? (prev_W - 2) * (prev_H - 2) : 0;
        prev.nblk = icpBlocksFor(prev_W, prev_H);
    }
    const size_t half = (size_t)kIcpSlots * kPartialRows;
    double* cur = partials + (size_t)(k & 1) * half;
    const double* old = partials + (size_t)((k + 1) & 1) * half;
    const unsigned blocks = level >= 0 ? (unsigned)icpBlocksFor(W, H) : 1u;
    hipLaunchKernelGGL(icpIter, dim3(blocks), dim3(kIcpThreads), 0, s, a, prev, (IcpState*)st, old, cur, k & 1);
    return hipGetLastError();
}

hipError_t launchIcpClearTicks(void* st, hipStream_t s) {
    hipLaunchKernelGGL(icpClearTicks, dim3(1), dim3(64), 0, s, (IcpState*)st);
    return hipGetLastError();
}

void icpTicksRead(const void* hostCopy, unsigned long long* ticks) {
    const IcpState* s = static_cast<const IcpState*>(hostCopy);
    ticks[0] = s->ticks_sum;
    ticks[1] = s->ticks_solve;
    ticks[2] = s->ticks_n;
    ticks[3] = s->work_px;
    ticks[4] = s->work_proj;
}

// Host view of the state (copied back once per frame).
void icpStateRead(const void* hostCopy, float* pose, float* information, float* error, float* final_step,
                  int* inliers, int* converged, int* counters, int* iterations) {
    const IcpState* s = static_cast<const IcpState*>(hostCopy);
    for (int i = 0; i < 16; ++i) pose[i] = s->pose[i];
    for (int i = 0; i < 36; ++i) information[i] = s->information[i];
    *error = s->error;
    *final_step = s->last_final_step;
    *inliers = s->inliers;
    *converged = s->converged;
    for (int i = 0; i < 5; ++i) counters[i] = s->counters[i];
    *iterations = s->iterations;
}

}  // namespace azu
