// Marching cubes on gfx950, bit-identical to the reference's CPU mesher
// (MarchingCubes::extract, src/meshing/MarchingCubes.cpp:115-382): same corner support
// rule, lower-endpoint crossing, blended unit normals, reversed (2,1,0) winding, triangle
// cap, and the same welded vertex order (slices ascending z, cubes y then x, triangles in
// table order, first insertion of an edge key wins).
//
// Passes (cubes and slots in emission order: cubes in linear z-y-x order, then
// triangles in table order, then slots):
//   1. crossMask, over all (R-1)^3 cubes: one bit per cube, set when its corner code has a
//      sign change (neither 0 nor 255); fewer than 1 % of the cubes of a fused scene are
//      set, so every later pass runs over full waves of crossing cubes or slots;
//   2-3. exclusive scan of the per-word bit counts and expansion into the ordered list
//      of crossing cubes;
//   4. classify, per crossing cube: corner code, the 12-bit mask of usable crossing edges
//      (both corners supported, finite crossing, finite endpoint normals, non-degenerate
//      blend) and the number of triangles whose three edges are usable;
//   5. exclusive scan of the triangle counts -> each cube's first triangle index;
//   6. per emitted slot below the triangle cap: its edge key (axis, lower voxel);
//   7. the owner of each edge key: its smallest slot (atomic minimum), which is the first
//      insertion of that key in emission order;
//   8-9. first-occurrence flags (slot == owner of its key), scanned into welded vertex
//      ids;
//   10. write, per slot: the index is its owner's vertex id; an owning slot also writes
//      the vertex.
// Arithmetic is IEEE single precision without contraction, in the CPU mesher's order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "mc_tables.h"

namespace azu {
namespace mc {

// Device copies of the shared tables: corner offsets, edge corners, and the triangle
// rows as edge ids terminated by -1 (uploaded once from azu_mc::kTriRows).
__constant__ int8_t dCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int8_t dEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int8_t dTri[256][16];
// dEdge for indices known at compile time (unrolled loops over the 12 edges).
constexpr int kEdgeCorner[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

__device__ __forceinline__ int edgeMaskDev(int cube) {
    int m = 0;
    for (int e = 0; e < 12; ++e)
        if (((cube >> dEdge[e][0]) & 1) != ((cube >> dEdge[e][1]) & 1)) m |= 1 << e;
    return m;
}

struct Corner {
    float v;
    bool sup;
};

template <class S> struct McVol {
    const typename S::Word* vox;
    int R;
    float vs;
    float ox, oy, oz;
    __device__ size_t at(int x, int y, int z) const { return ((size_t)z * R + y) * R + x; }
};

// sampleCorner (MarchingCubes.cpp:81-85)
template <class S>
__device__ __forceinline__ Corner corner(const McVol<S>& V, int x, int y, int z) {
    const typename S::Word w = V.vox[V.at(x, y, z)];
    const float t = S::tsdf(w), wt = S::weight(w);
    if (!isfinite(t) || wt <= 0.001f) return {1.0f, false};
    return {t, true};
}

__device__ __forceinline__ float n3(F3 v) { return sqrtf(v.x * v.x + (v.y * v.y + v.z * v.z)); }

// voxelNormal (MarchingCubes.cpp:136-172); NaN vector when unusable.
template <class S>
__device__ F3 voxelNormal(const McVol<S>& V, int x, int y, int z) {
    const float nan = __builtin_nanf("");
    auto usable = [&](int xi, int yi, int zi, float& out) {
        if (xi < 0 || xi >= V.R || yi < 0 || yi >= V.R || zi < 0 || zi >= V.R) return false;
        const Corner c = corner(V, xi, yi, zi);
        if (!c.sup) return false;
        out = c.v;
        return true;
    };
    float self = 0.0f;
    if (!usable(x, y, z, self)) return F3{nan, nan, nan};
    float g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ox = a == 0, oy = a == 1, oz = a == 2;
        float hi = 0.0f, lo = 0.0f;
        const bool hh = usable(x + ox, y + oy, z + oz, hi);
        const bool hl = usable(x - ox, y - oy, z - oz, lo);
        if (hh && hl) g[a] = 0.5f * (hi - lo);
        else if (hh) g[a] = hi - self;
        else if (hl) g[a] = self - lo;
        else g[a] = 0.0f;
    }
    const F3 gv{g[0], g[1], g[2]};
    if (!f3finite(gv)) return F3{nan, nan, nan};
    const float len = n3(gv);
    if (!isfinite(len) || len <= 1e-6f) return F3{nan, nan, nan};
    return F3{gv.x / len, gv.y / len, gv.z / len};
}

struct EdgeOut {
    bool ok;
    F3 pos, nrm;
};

// One crossing edge of cube (x, y, z) (MarchingCubes.cpp:254-313): crossingFromLower on
// the canonical lower endpoint, blended unit normal.
// k0 and k1 are the corners dEdge[e][0] and dEdge[e][1].
template <class S>
__device__ EdgeOut edgeVertex(const McVol<S>& V, int x, int y, int z, int e, Corner k0, Corner k1) {
    EdgeOut o;
    o.ok = false;
    const int c0 = dEdge[e][0], c1 = dEdge[e][1];
    if (!k0.sup || !k1.sup) return o;
    int lo = c0, up = c1;
    bool c0low = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (dCorner[c0][a] != dCorner[c1][a]) {
            c0low = dCorner[c0][a] <= dCorner[c1][a];
            lo = c0low ? c0 : c1;
            up = c0low ? c1 : c0;
        }
    const int lx = x + dCorner[lo][0], ly = y + dCorner[lo][1],
              lz = z + dCorner[lo][2];
    const int ux = x + dCorner[up][0], uy = y + dCorner[up][1],
              uz = z + dCorner[up][2];
    const F3 pl{V.ox + (float)lx * V.vs, V.oy + (float)ly * V.vs, V.oz + (float)lz * V.vs};
    const F3 pu{V.ox + (float)ux * V.vs, V.oy + (float)uy * V.vs, V.oz + (float)uz * V.vs};
    const float vl = c0low ? k0.v : k1.v, vu = c0low ? k1.v : k0.v;
    if (!isfinite(vl) || !isfinite(vu) || !f3finite(pl) || !f3finite(pu)) return o;
    F3 p;
    float t;
    if (fabsf(vl) < 1e-6f) {
        p = pl;
        t = 0.0f;
    } else if (fabsf(vu) < 1e-6f) {
        p = pu;
        t = 1.0f;
    } else {
        const float diff = vl - vu;
        if (!isfinite(diff) || fabsf(diff) < 1e-6f) {
            p = pl;
            t = 0.0f;
        } else {
            const float q = vl / diff;
            t = smax(0.0f, smin(1.0f, q));
            p = F3{pl.x + t * (pu.x - pl.x), pl.y + t * (pu.y - pl.y), pl.z + t * (pu.z - pl.z)};
        }
    }
    const F3 nl = voxelNormal(V, lx, ly, lz), nu = voxelNormal(V, ux, uy, uz);
    if (!f3finite(nl) || !f3finite(nu)) return o;
    const F3 bl{nl.x + t * (nu.x - nl.x), nl.y + t * (nu.y - nl.y), nl.z + t * (nu.z - nl.z)};
    const float bl_len = n3(bl);
    if (!isfinite(bl_len) || bl_len <= 1e-6f) return o;
    o.ok = true;
    o.pos = p;
    o.nrm = F3{bl.x / bl_len, bl.y / bl_len, bl.z / bl_len};
    return o;
}

__device__ __forceinline__ int triEdge(const int8_t* row, int i) { return row[i]; }

__device__ __forceinline__ void edgeKey(int x, int y, int z, int e, int& lx, int& ly, int& lz, int& axis) {
    const int c0 = dEdge[e][0], c1 = dEdge[e][1];
    axis = 0;
    for (int k = 0; k < 3; ++k)
        if (dCorner[c0][k] != dCorner[c1][k]) axis = k;
    const int lo = dCorner[c0][axis] <= dCorner[c1][axis] ? c0 : c1;
    lx = x + dCorner[lo][0];
    ly = y + dCorner[lo][1];
    lz = z + dCorner[lo][2];
}

// Pass 1 geometry: a block covers kMaskX cubes of one row (y, z) for kMaskZ layers of z;
// wave w of the block owns mask word w of the row segment.
constexpr int kMaskX = 512;
constexpr int kMaskZ = 8;

// Pass 1: a bit per cube, set when its eight corner signs (value below zero; an
// unsupported voxel reads +1) are not all equal, i.e. the corner code is neither 0 nor
// 255 — the only cubes with a crossing edge, and every cube sharing an emitted edge with
// a set cube is set too, since it holds the same two corners. Lane x holds the signs of
// voxel columns x and x + 1 of rows y and y + 1 for the block's layers. Blocks are
// dispatched round-robin over the 8 XCDs; each XCD takes a contiguous run of rows so the
// voxel rows shared by neighbouring rows are read from its own L2.
template <class S>
__global__ __launch_bounds__(512) void crossMask(McVol<S> V, int nb, unsigned long long* __restrict__ mask,
                                                 uint32_t* __restrict__ wcount) {
    const int C = V.R - 1;
    const unsigned nblk = gridDim.x, q = nblk / 8, r = nblk % 8, xcd = blockIdx.x % 8;
    const unsigned lb = xcd * q + min(xcd, r) + blockIdx.x / 8;
    const int bx = (int)(lb % (unsigned)nb);
    const unsigned t = lb / (unsigned)nb;
    const int y = (int)(t % (unsigned)C), zc = (int)(t / (unsigned)C);
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int x = bx * kMaskX + w * 64 + lane;
    const int z0 = zc * kMaskZ, z1 = min(z0 + kMaskZ, C);
    // lanes past the row read its last voxel (clamped) and set no bit
    const int xa = min(x, C), xb = min(x + 1, C);
    auto neg = [&](int xi, int yi, int zi) -> uint32_t { return corner(V, xi, yi, zi).v < 0.0f ? 1u : 0u; };
    // bits 0-3: voxels (x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1) of layer zi
    auto layer = [&](int zi) -> uint32_t {
        return neg(xa, y, zi) | neg(xb, y, zi) << 1 | neg(xa, y + 1, zi) << 2 | neg(xb, y + 1, zi) << 3;
    };
    // all layers of the block are loaded before any is used, so the loads overlap
    uint32_t lay[kMaskZ + 1];
#pragma unroll
    for (int k = 0; k <= kMaskZ; ++k) lay[k] = layer(min(z0 + k, z1));
#pragma unroll
    for (int k = 0; k < kMaskZ; ++k) {
        const int z = z0 + k;
        if (z >= z1) break;
        const uint32_t all = lay[k] | lay[k + 1] << 4;
        const unsigned long long m = __ballot(x < C && all != 0u && all != 255u);
        if (lane == 0) {
            const size_t wi = (((size_t)z * C + y) * nb + bx) * 8 + w;
            mask[wi] = m;
            wcount[wi] = (uint32_t)__popcll(m);
        }
    }
}

// Pass 3: the set cubes' linear indices in ascending order, one thread per mask word.
__global__ __launch_bounds__(256) void expandMask(const unsigned long long* __restrict__ mask,
                                                  const uint32_t* __restrict__ woff, size_t nwords, int C, int nb,
                                                  uint32_t* __restrict__ cand) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nwords) return;
    unsigned long long m = mask[i];
    if (!m) return;
    const size_t seg = i / 8, row = seg / (size_t)nb;
    const uint32_t x0 = (uint32_t)((seg % (size_t)nb) * kMaskX + (i % 8) * 64);
    const uint32_t c0 = (uint32_t)(row * (size_t)C) + x0;
    uint32_t o = woff[i];
    while (m) {
        cand[o++] = c0 + (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
    }
}

__device__ __forceinline__ void cubeXYZ(uint32_t c, uint32_t C, int& x, int& y, int& z) {
    const uint32_t row = c / C;
    x = (int)(c - row * C);
    z = (int)(row / C);
    y = (int)(row - (uint32_t)z * C);
}

// Pass 4, per selected cube: corner code, the mask of usable crossing edges, and the
// number of triangles whose three edges are usable. Thread nc writes the zero that
// closes the count array for the exclusive scan.
template <class S>
__global__ __launch_bounds__(256) void classify(McVol<S> V, const uint32_t* __restrict__ cand, uint32_t nc,
                                                uint32_t* __restrict__ info, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) {
        if (i == nc) count[nc] = 0;
        return;
    }
    int x, y, z;
    cubeXYZ(cand[i], (uint32_t)(V.R - 1), x, y, z);
    Corner cs[8];
    int cube = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        cs[k] = corner(V, x + dCorner[k][0], y + dCorner[k][1], z + dCorner[k][2]);
        if (cs[k].v < 0.0f) cube |= 1 << k;
    }
    const int emask = edgeMaskDev(cube);
    int ok = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if ((emask >> e) & 1)
            if (edgeVertex(V, x, y, z, e, cs[kEdgeCorner[e][0]], cs[kEdgeCorner[e][1]]).ok) ok |= 1 << e;
    int tris = 0;
    if (ok) {
        const int8_t* row = dTri[cube];
        for (int t = 0; row[t] >= 0; t += 3)
            if (((ok >> triEdge(row, t)) & 1) && ((ok >> triEdge(row, t + 1)) & 1) && ((ok >> triEdge(row, t + 2)) & 1))
                ++tris;
    }
    info[i] = (uint32_t)cube | ((uint32_t)ok << 8);
    count[i] = (uint32_t)tris;
}

// Pass 6, per selected cube: for every emitted slot s = 3 (first triangle + j) + k below
// the cap (slot k of triangle j holds table edge 2 - k), the key of its edge (axis, lower
// voxel) and a reference {cube, local edge}; the edge's owner entry is reset for pass 7.
__global__ __launch_bounds__(256) void slotKeys(const uint32_t* __restrict__ cand, uint32_t nc,
                                                const uint32_t* __restrict__ info, const uint32_t* __restrict__ count,
                                                const uint32_t* __restrict__ tri_off, int R, unsigned long long cap,
                                                unsigned long long* __restrict__ key, uint2* __restrict__ ref,
                                                uint32_t* __restrict__ owner) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc || count[i] == 0) return;
    int x, y, z;
    cubeXYZ(cand[i], (uint32_t)(R - 1), x, y, z);
    const uint32_t inf = info[i];
    const int cube = inf & 0xff, ok = (int)(inf >> 8) & 0xfff;
    const int8_t* row = dTri[cube];
    unsigned long long g = tri_off[i];
    for (int t = 0; row[t] >= 0 && g < cap; t += 3) {
        const int te[3] = {triEdge(row, t), triEdge(row, t + 1), triEdge(row, t + 2)};
        if (!((ok >> te[0]) & 1) || !((ok >> te[1]) & 1) || !((ok >> te[2]) & 1)) continue;
        for (int k = 0; k < 3; ++k) {
            int lx, ly, lz, axis;
            edgeKey(x, y, z, te[2 - k], lx, ly, lz, axis);
            const unsigned long long kk = (((unsigned long long)axis * R + lz) * R + ly) * R + lx;
            key[3 * g + k] = kk;
            ref[3 * g + k] = make_uint2(i, (uint32_t)te[2 - k]);
            owner[kk] = 0xffffffffu;
        }
        ++g;
    }
}

// Pass 7: the owner of an edge is its smallest slot. Slot order is emission order (cubes
// ascending, then triangles, then slots), so this is the first insertion of the edge key.
__global__ __launch_bounds__(256) void ownerMin(const unsigned long long* __restrict__ key, uint32_t nslots,
                                                uint32_t* __restrict__ owner) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) atomicMin(&owner[key[s]], s);
}

// Pass 8: first-occurrence flag of every slot; thread nslots closes the array with 0.
__global__ __launch_bounds__(256) void ownerFlags(const unsigned long long* __restrict__ key, uint32_t nslots,
                                                  const uint32_t* __restrict__ owner, uint32_t* __restrict__ flags) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) flags[s] = owner[key[s]] == s ? 1u : 0u;
    else if (s == nslots) flags[s] = 0u;
}

// Pass 10, per slot: the index of its welded vertex; an owning slot also writes the
// vertex (position and blended normal of its crossing).
template <class S>
__global__ __launch_bounds__(256) void writeMesh(McVol<S> V, const uint32_t* __restrict__ cand,
                                                 const unsigned long long* __restrict__ key,
                                                 const uint2* __restrict__ ref, uint32_t nslots,
                                                 const uint32_t* __restrict__ owner, const uint32_t* __restrict__ vid,
                                                 float* __restrict__ pos, float* __restrict__ nrm,
                                                 uint32_t* __restrict__ idx) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t own = owner[key[s]];
    const uint32_t v = vid[own];
    idx[s] = v;
    if (own != s) return;
    const uint2 rf = ref[s];
    int x, y, z;
    cubeXYZ(cand[rf.x], (uint32_t)(V.R - 1), x, y, z);
    const int e = (int)rf.y, c0 = dEdge[e][0], c1 = dEdge[e][1];
    const Corner k0 = corner(V, x + dCorner[c0][0], y + dCorner[c0][1], z + dCorner[c0][2]);
    const Corner k1 = corner(V, x + dCorner[c1][0], y + dCorner[c1][1], z + dCorner[c1][2]);
    const EdgeOut eo = edgeVertex(V, x, y, z, e, k0, k1);
    pos[3 * (size_t)v + 0] = eo.pos.x;
    pos[3 * (size_t)v + 1] = eo.pos.y;
    pos[3 * (size_t)v + 2] = eo.pos.z;
    nrm[3 * (size_t)v + 0] = eo.nrm.x;
    nrm[3 * (size_t)v + 1] = eo.nrm.y;
    nrm[3 * (size_t)v + 2] = eo.nrm.z;
}

// ---------------------------------------------------------------------------
// Global point cloud (TSDFVolume::extractGlobalPointCloud, TSDFVolume.cpp:571-632, and
// its GPU form compactPointsKernel, TSDFVolume_cuda.cu:409-447): every voxel with
// weight > 1 and |tsdf| < 0.2, as origin + (x, y, z) * voxel_size, in linear voxel
// order. One wave per x-row of the volume: the row count is a ballot popcount per 64
// voxels; the write pass places each point at its row's offset plus the lane's rank.

template <class S>
__device__ __forceinline__ bool pcKeep(typename S::Word w) {
    return S::weight(w) > 1.0f && fabsf(S::tsdf(w)) < 0.2f;
}

template <class S>
__global__ __launch_bounds__(256) void pointCount(McVol<S> V, uint32_t* __restrict__ row_count) {
    const size_t row = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= (size_t)V.R * V.R) return;
    const typename S::Word* src = V.vox + row * V.R;
    uint32_t n = 0;
    // four 64-voxel chunks per step, loads issued before the ballots
    for (int x0 = 0; x0 < V.R; x0 += 256) {
        typename S::Word w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = x0 + 64 * k + lane;
            w[k] = x < V.R ? src[x] : S::empty();
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) n += (uint32_t)__popcll(__ballot(pcKeep<S>(w[k])));
    }
    if (lane == 0) row_count[row] = n;
}

template <class S>
__global__ __launch_bounds__(256) void pointWrite(McVol<S> V, const uint32_t* __restrict__ row_off,
                                                  unsigned long long capacity, float* __restrict__ out) {
    const size_t row = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= (size_t)V.R * V.R) return;
    const int z = (int)(row / (size_t)V.R), y = (int)(row % (size_t)V.R);
    const float py = V.oy + (float)y * V.vs, pz = V.oz + (float)z * V.vs;
    const typename S::Word* src = V.vox + row * V.R;
    unsigned long long base = row_off[row];
    for (int x0 = 0; x0 < V.R; x0 += 256) {
        typename S::Word w[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int x = x0 + 64 * c + lane;
            w[c] = x < V.R ? src[x] : S::empty();
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int x = x0 + 64 * c + lane;
            const bool keep = pcKeep<S>(w[c]);
            const unsigned long long mask = __ballot(keep);
            const unsigned rank =
                __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            const unsigned long long k = base + rank;
            if (keep && k < capacity) {
                out[3 * k + 0] = V.ox + (float)x * V.vs;
                out[3 * k + 1] = py;
                out[3 * k + 2] = pz;
            }
            base += (unsigned long long)__popcll(mask);
        }
    }
}

}  // namespace mc

struct McResult {
    unsigned long long ntri, nvert;
    bool truncated;
    float* pos;  // device
    float* nrm;
    uint32_t* idx;
};

static hipError_t uploadTriTable() {
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [] {
        int8_t host[256][16];
        for (int c = 0; c < 256; ++c) {
            const char* r = azu_mc::kTriRows[c];
            int n = 0;
            for (; r[n] != '\0' && n < 15; ++n) host[c][n] = (int8_t)(r[n] <= '9' ? r[n] - '0' : r[n] - 'a' + 10);
            for (; n < 16; ++n) host[c][n] = -1;
        }
        status = hipMemcpyToSymbol(HIP_SYMBOL(mc::dTri), host, sizeof(host));
    });
    return status;
}

// Runs the whole extraction; on success the caller owns res->pos / nrm / idx (hipFree).
template <class S>
static hipError_t mcRun(const mc::McVol<S>& V, unsigned long long max_tris, McResult* res, hipStream_t st) {
    res->ntri = res->nvert = 0;
    res->truncated = false;
    res->pos = res->nrm = nullptr;
    res->idx = nullptr;
    const int C = V.R - 1;
    if (C < 1) return hipSuccess;
    const size_t n = (size_t)C * C * C;
    if (n >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    {
        const hipError_t te = uploadTriTable();
        if (te != hipSuccess) return te;
    }
    const int nb = (C + mc::kMaskX - 1) / mc::kMaskX, zchunks = (C + mc::kMaskZ - 1) / mc::kMaskZ;
    const size_t nwords = (size_t)C * C * nb * 8;
    if (nwords + 1 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    unsigned long long *mask = nullptr, *key = nullptr;
    uint32_t *wcount = nullptr, *woff = nullptr, *cand = nullptr, *info = nullptr, *count = nullptr, *off = nullptr,
             *owner = nullptr, *flags = nullptr, *vid = nullptr;
    uint2* ref = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    auto release = [](auto*& q) {
        if (q) (void)hipFree(q);
        q = nullptr;
    };
    auto cleanup = [&]() {
        release(mask);
        release(key);
        release(wcount);
        release(woff);
        release(cand);
        release(info);
        release(count);
        release(off);
        release(owner);
        release(flags);
        release(vid);
        release(ref);
        release(tmp);
    };
    // exclusive scan of m + 1 counts (the last one zero): out[m] is the total
    auto scan = [&](const uint32_t* in, uint32_t* out, size_t m, uint32_t* total) -> hipError_t {
        size_t need = 0;
        hipError_t r = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)(m + 1), st);
        if (r == hipSuccess && (need > tmp_bytes || !tmp)) {
            release(tmp);
            tmp_bytes = std::max<size_t>(need, 1);
            r = hipMalloc(&tmp, tmp_bytes);
        }
        if (r == hipSuccess) r = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)(m + 1), st);
        if (r == hipSuccess) r = hipMemcpyAsync(total, out + m, 4, hipMemcpyDeviceToHost, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        return r;
    };
    auto blocks = [](size_t m) { return dim3((unsigned)((m + 255) / 256)); };
    // passes 1-3: crossing-cube bits, their scan, the ordered list of crossing cubes
    hipError_t e = hipMalloc(&mask, nwords * 8);
    if (e == hipSuccess) e = hipMalloc(&wcount, (nwords + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&woff, (nwords + 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::crossMask<S>, dim3((unsigned)((size_t)nb * C * zchunks)), dim3(512), 0, st, V, nb,
                           mask, wcount);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(wcount + nwords, 0, 4, st);
    uint32_t nc = 0;
    if (e == hipSuccess) e = scan(wcount, woff, nwords, &nc);
    if (e == hipSuccess && nc > 0) e = hipMalloc(&cand, (size_t)nc * 4);
    if (e == hipSuccess && nc > 0) {
        hipLaunchKernelGGL(mc::expandMask, blocks(nwords), dim3(256), 0, st, mask, woff, nwords, C, nb, cand);
        e = hipGetLastError();
    }
    release(wcount);
    if (e != hipSuccess || nc == 0) {
        cleanup();
        return e;
    }
    // passes 4-5: classify the crossing cubes, scan their triangle counts
    e = hipMalloc(&info, (size_t)nc * 4);
    if (e == hipSuccess) e = hipMalloc(&count, ((size_t)nc + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&off, ((size_t)nc + 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::classify<S>, blocks((size_t)nc + 1), dim3(256), 0, st, V, cand, nc, info, count);
        e = hipGetLastError();
    }
    uint32_t total = 0;
    if (e == hipSuccess) e = scan(count, off, nc, &total);
    if (e != hipSuccess) {
        cleanup();
        return e;
    }
    const unsigned long long ntri = std::min<unsigned long long>(total, max_tris);
    res->truncated = total > max_tris;
    res->ntri = ntri;
    if (ntri == 0) {
        cleanup();
        return hipSuccess;
    }
    const size_t nslots = 3 * (size_t)ntri;
    if (nslots + 1 >= ((size_t)1 << 31)) {
        cleanup();
        return hipErrorInvalidValue;
    }
    // passes 6-9: slot edge keys, edge owners, first-occurrence flags, welded vertex ids
    const size_t nedges = 3 * (size_t)V.R * V.R * V.R;
    e = hipMalloc(&key, nslots * 8);
    if (e == hipSuccess) e = hipMalloc(&ref, nslots * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc(&owner, nedges * 4);
    if (e == hipSuccess) e = hipMalloc(&flags, (nslots + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&vid, (nslots + 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::slotKeys, blocks(nc), dim3(256), 0, st, cand, nc, info, count, off, V.R, ntri, key, ref,
                           owner);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::ownerMin, blocks(nslots), dim3(256), 0, st, key, (uint32_t)nslots, owner);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::ownerFlags, blocks(nslots + 1), dim3(256), 0, st, key, (uint32_t)nslots, owner,
                           flags);
        e = hipGetLastError();
    }
    uint32_t nvert = 0;
    if (e == hipSuccess) e = scan(flags, vid, nslots, &nvert);
    // pass 10: vertices and indices
    if (e == hipSuccess) e = hipMalloc(&res->pos, (size_t)nvert * 12 + 12);
    if (e == hipSuccess) e = hipMalloc(&res->nrm, (size_t)nvert * 12 + 12);
    if (e == hipSuccess) e = hipMalloc(&res->idx, nslots * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::writeMesh<S>, blocks(nslots), dim3(256), 0, st, V, cand, key, ref, (uint32_t)nslots,
                           owner, vid, res->pos, res->nrm, res->idx);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    res->nvert = nvert;
    cleanup();
    if (e != hipSuccess) {
        if (res->pos) (void)hipFree(res->pos);
        if (res->nrm) (void)hipFree(res->nrm);
        if (res->idx) (void)hipFree(res->idx);
        res->pos = res->nrm = nullptr;
        res->idx = nullptr;
    }
    return e;
}

hipError_t launchMarchingCubes(int format, const void* vox, const VolParams& p, unsigned long long max_tris,
                               McResult* res, hipStream_t st) {
    if (format == kFmtF32) {
        mc::McVol<StoreF32> V{static_cast<const float2*>(vox), p.res, p.vs, p.origin[0], p.origin[1], p.origin[2]};
        return mcRun(V, max_tris, res, st);
    }
    mc::McVol<StoreF16> V{static_cast<const uint32_t*>(vox), p.res, p.vs, p.origin[0], p.origin[1], p.origin[2]};
    return mcRun(V, max_tris, res, st);
}

template <class S>
static hipError_t pcRun(const mc::McVol<S>& V, float* d_out, unsigned long long capacity, unsigned long long* total,
                        hipStream_t st) {
    *total = 0;
    const size_t rows = (size_t)V.R * V.R;
    if (V.R < 1) return hipSuccess;
    if (rows + 1 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    uint32_t *cnt = nullptr, *off = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipMalloc(&cnt, (rows + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&off, (rows + 1) * 4);
    const unsigned blocks = (unsigned)((rows + 3) / 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::pointCount<S>, dim3(blocks), dim3(256), 0, st, V, cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(cnt + rows, 0, 4, st);
    // exclusive scan over rows + 1 entries: off[rows] is the total (below 2^32: at most
    // 2^31 voxels)
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, off, (int)(rows + 1), st);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, (int)(rows + 1), st);
    if (e == hipSuccess && d_out && capacity > 0) {
        hipLaunchKernelGGL(mc::pointWrite<S>, dim3(blocks), dim3(256), 0, st, V, off, capacity, d_out);
        e = hipGetLastError();
    }
    uint32_t n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, off + rows, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) *total = n;
    if (cnt) (void)hipFree(cnt);
    if (off) (void)hipFree(off);
    if (tmp) (void)hipFree(tmp);
    return e;
}

hipError_t launchPointCloud(int format, const void* vox, const VolParams& p, float* d_out,
                            unsigned long long capacity, unsigned long long* total, hipStream_t st) {
    if ((size_t)p.res * p.res * p.res >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    if (format == kFmtF32) {
        mc::McVol<StoreF32> V{static_cast<const float2*>(vox), p.res, p.vs, p.origin[0], p.origin[1], p.origin[2]};
        return pcRun(V, d_out, capacity, total, st);
    }
    mc::McVol<StoreF16> V{static_cast<const uint32_t*>(vox), p.res, p.vs, p.origin[0], p.origin[1], p.origin[2]};
    return pcRun(V, d_out, capacity, total, st);
}

}  // namespace azu
