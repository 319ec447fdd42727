// Marching cubes on gfx950, bit-identical to the reference's CPU mesher
// (MarchingCubes::extract, src/meshing/MarchingCubes.cpp:115-382): same corner support
// rule, lower-endpoint crossing, blended unit normals, reversed (2,1,0) winding, triangle
// cap, and the same welded vertex order (slices ascending z, cubes y then x, triangles in
// table order, first insertion of an edge key wins).
//
// Passes (all over the (R-1)^3 cubes in linear z-y-x order):
//   1. classify: corner code, the 12-bit mask of usable crossing edges (both corners
//      supported, finite crossing, finite endpoint normals, non-degenerate blend) and the
//      number of triangles whose three edges are usable;
//   2. exclusive scan of the triangle counts -> each cube's first triangle index;
//   3. first-occurrence flags of every emitted vertex slot: a slot owns its edge key when
//      no earlier slot of its cube and no emitted triangle of an earlier cube sharing the
//      edge (at most three: the edge's lower voxel minus the other two axes) uses it;
//   4. exclusive scan of the flags -> welded vertex ids;
//   5. write: an owning slot writes its vertex; every slot writes its index, an owned
//      slot by looking up its owner (the earliest cube and slot using the edge).
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
template <class S>
__device__ EdgeOut edgeVertex(const McVol<S>& V, int x, int y, int z, int e, const Corner* cs) {
    EdgeOut o;
    o.ok = false;
    const int c0 = dEdge[e][0], c1 = dEdge[e][1];
    if (!cs[c0].sup || !cs[c1].sup) return o;
    int lo = c0, up = c1;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (dCorner[c0][a] != dCorner[c1][a]) {
            const bool c0low = dCorner[c0][a] <= dCorner[c1][a];
            lo = c0low ? c0 : c1;
            up = c0low ? c1 : c0;
        }
    const int lx = x + dCorner[lo][0], ly = y + dCorner[lo][1],
              lz = z + dCorner[lo][2];
    const int ux = x + dCorner[up][0], uy = y + dCorner[up][1],
              uz = z + dCorner[up][2];
    const F3 pl{V.ox + (float)lx * V.vs, V.oy + (float)ly * V.vs, V.oz + (float)lz * V.vs};
    const F3 pu{V.ox + (float)ux * V.vs, V.oy + (float)uy * V.vs, V.oz + (float)uz * V.vs};
    const float vl = cs[lo].v, vu = cs[up].v;
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

__device__ __forceinline__ int rowLen(const int8_t* row) {
    int n = 0;
    while (n < 16 && row[n] >= 0) ++n;
    return n;
}

// Local edge of a cube whose lower corner offset is (dx, dy, dz) along `axis`, or -1.
__device__ __forceinline__ int localEdge(int axis, int dx, int dy, int dz) {
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const int c0 = dEdge[e][0], c1 = dEdge[e][1];
        int a = 0;
        for (int k = 0; k < 3; ++k)
            if (dCorner[c0][k] != dCorner[c1][k]) a = k;
        if (a != axis) continue;
        const int lo = dCorner[c0][a] <= dCorner[c1][a] ? c0 : c1;
        if (dCorner[lo][0] == dx && dCorner[lo][1] == dy &&
            dCorner[lo][2] == dz)
            return e;
    }
    return -1;
}

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

// First emitted triangle j and slot k (slot k holds table edge 2 - k) of a cube that use
// local edge `e`, among triangles with global index < cap. Returns false if none.
__device__ bool firstUse(uint32_t info, uint32_t tri_off, unsigned long long cap, int e, int& jo, int& ko) {
    const int cube = info & 0xff;
    const int ok = (int)(info >> 8) & 0xfff;
    const int8_t* row = dTri[cube];
    const int n = rowLen(row);
    int j = 0;
    for (int t = 0; t < n; t += 3) {
        const int te[3] = {triEdge(row, t), triEdge(row, t + 1), triEdge(row, t + 2)};
        if (!((ok >> te[0]) & 1) || !((ok >> te[1]) & 1) || !((ok >> te[2]) & 1)) continue;
        if ((unsigned long long)tri_off + j >= cap) return false;
        for (int k = 0; k < 3; ++k)
            if (te[2 - k] == e) {
                jo = j;
                ko = k;
                return true;
            }
        ++j;
    }
    return false;
}

template <class S>
__global__ __launch_bounds__(256) void classify(McVol<S> V, uint32_t* __restrict__ info, uint32_t* __restrict__ count) {
    const int C = V.R - 1;
    const size_t n = (size_t)C * C * C;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(c % C), y = (int)((c / C) % C), z = (int)(c / ((size_t)C * C));
        Corner cs[8];
        int cube = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cs[k] = corner(V, x + dCorner[k][0], y + dCorner[k][1],
                           z + dCorner[k][2]);
            if (cs[k].v < 0.0f) cube |= 1 << k;
        }
        const int emask = edgeMaskDev(cube);
        int ok = 0;
        for (int e = 0; e < 12; ++e)
            if ((emask >> e) & 1)
                if (edgeVertex(V, x, y, z, e, cs).ok) ok |= 1 << e;
        int tris = 0;
        if (ok) {
            const int8_t* row = dTri[cube];
            for (int t = 0; row[t] >= 0; t += 3)
                if (((ok >> triEdge(row, t)) & 1) && ((ok >> triEdge(row, t + 1)) & 1) && ((ok >> triEdge(row, t + 2)) & 1))
                    ++tris;
        }
        info[c] = (uint32_t)cube | ((uint32_t)ok << 8);
        count[c] = (uint32_t)tris;
    }
}

// The owner (cube, triangle, slot) of the edge used by slot (j, k) of cube (x, y, z).
// Returns the owner's global slot index.
__device__ unsigned long long ownerSlot(const uint32_t* info, const uint32_t* tri_off, int C, unsigned long long cap,
                                        int x, int y, int z, int e, int j_self, int k_self) {
    int lx, ly, lz, axis;
    edgeKey(x, y, z, e, lx, ly, lz, axis);
    // cubes sharing the edge: origin = lower voxel - d, d over the two other axes
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    int best_cube = -1;
    long long best_lin = -1;
    int bj = 0, bk = 0;
    // scan in ascending linear order: d = (1,1), (0,1), (1,0), (0,0) along (a1, a2) sorted by linear index
    for (int m = 0; m < 4; ++m) {
        int d[3] = {0, 0, 0};
        d[a1] = m & 1;
        d[a2] = (m >> 1) & 1;
        const int ox = lx - d[0], oy = ly - d[1], oz = lz - d[2];
        if (ox < 0 || oy < 0 || oz < 0 || ox >= C || oy >= C || oz >= C) continue;
        const long long lin = ((long long)oz * C + oy) * C + ox;
        const int le = localEdge(axis, d[0], d[1], d[2]);
        int jo, ko;
        if (!firstUse(info[lin], tri_off[lin], cap, le, jo, ko)) continue;
        if (best_cube < 0 || lin < best_lin) {
            best_cube = 1;
            best_lin = lin;
            bj = jo;
            bk = ko;
        }
    }
    if (best_cube < 0) {  // cannot happen: the calling slot itself uses the edge
        return 3ull * ((unsigned long long)tri_off[((long long)z * C + y) * C + x] + j_self) + k_self;
    }
    return 3ull * ((unsigned long long)tri_off[best_lin] + bj) + bk;
}

__global__ __launch_bounds__(256) void firstFlags(const uint32_t* __restrict__ info, const uint32_t* __restrict__ count,
                                                  const uint32_t* __restrict__ tri_off, int C, unsigned long long cap,
                                                  uint32_t* __restrict__ flags) {
    const size_t n = (size_t)C * C * C;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        if (count[c] == 0) continue;
        const int x = (int)(c % C), y = (int)((c / C) % C), z = (int)(c / ((size_t)C * C));
        const uint32_t inf = info[c];
        const int cube = inf & 0xff, ok = (int)(inf >> 8) & 0xfff;
        const int8_t* row = dTri[cube];
        int j = 0;
        for (int t = 0; row[t] >= 0; t += 3) {
            const int te[3] = {triEdge(row, t), triEdge(row, t + 1), triEdge(row, t + 2)};
            if (!((ok >> te[0]) & 1) || !((ok >> te[1]) & 1) || !((ok >> te[2]) & 1)) continue;
            const unsigned long long g = (unsigned long long)tri_off[c] + j;
            if (g >= cap) break;
            for (int k = 0; k < 3; ++k) {
                const unsigned long long s = 3ull * g + k;
                const unsigned long long own = ownerSlot(info, tri_off, C, cap, x, y, z, te[2 - k], j, k);
                flags[s] = own == s ? 1u : 0u;
            }
            ++j;
        }
    }
}

template <class S>
__global__ __launch_bounds__(256) void writeMesh(McVol<S> V, const uint32_t* __restrict__ info,
                                                 const uint32_t* __restrict__ count,
                                                 const uint32_t* __restrict__ tri_off,
                                                 const uint32_t* __restrict__ vid, unsigned long long cap,
                                                 float* __restrict__ pos, float* __restrict__ nrm,
                                                 uint32_t* __restrict__ idx) {
    const int C = V.R - 1;
    const size_t n = (size_t)C * C * C;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        if (count[c] == 0) continue;
        const int x = (int)(c % C), y = (int)((c / C) % C), z = (int)(c / ((size_t)C * C));
        const uint32_t inf = info[c];
        const int cube = inf & 0xff, ok = (int)(inf >> 8) & 0xfff;
        Corner cs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            cs[k] = corner(V, x + dCorner[k][0], y + dCorner[k][1],
                           z + dCorner[k][2]);
        const int8_t* row = dTri[cube];
        int j = 0;
        for (int t = 0; row[t] >= 0; t += 3) {
            const int te[3] = {triEdge(row, t), triEdge(row, t + 1), triEdge(row, t + 2)};
            if (!((ok >> te[0]) & 1) || !((ok >> te[1]) & 1) || !((ok >> te[2]) & 1)) continue;
            const unsigned long long g = (unsigned long long)tri_off[c] + j;
            if (g >= cap) break;
            for (int k = 0; k < 3; ++k) {
                const unsigned long long s = 3ull * g + k;
                const unsigned long long own = ownerSlot(info, tri_off, C, cap, x, y, z, te[2 - k], j, k);
                const uint32_t v = vid[own];
                idx[s] = v;
                if (own == s) {
                    const EdgeOut eo = edgeVertex(V, x, y, z, te[2 - k], cs);
                    pos[3 * (size_t)v + 0] = eo.pos.x;
                    pos[3 * (size_t)v + 1] = eo.pos.y;
                    pos[3 * (size_t)v + 2] = eo.pos.z;
                    nrm[3 * (size_t)v + 0] = eo.nrm.x;
                    nrm[3 * (size_t)v + 1] = eo.nrm.y;
                    nrm[3 * (size_t)v + 2] = eo.nrm.z;
                }
            }
            ++j;
        }
    }
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

static unsigned gridFor(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 131072); }

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
    uint32_t *info = nullptr, *count = nullptr, *off = nullptr, *flags = nullptr, *vid = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipMalloc(&info, n * 4);
    if (e == hipSuccess) e = hipMalloc(&count, (n + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&off, (n + 1) * 4);
    auto cleanup = [&]() {
        if (info) (void)hipFree(info);
        if (count) (void)hipFree(count);
        if (off) (void)hipFree(off);
        if (flags) (void)hipFree(flags);
        if (vid) (void)hipFree(vid);
        if (tmp) (void)hipFree(tmp);
    };
    if (e != hipSuccess) {
        cleanup();
        return e;
    }
    hipLaunchKernelGGL(mc::classify<S>, dim3(gridFor(n)), dim3(256), 0, st, V, info, count);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemsetAsync(count + n, 0, 4, st);
    // exclusive scan over n + 1 entries: off[n] is the total
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, count, off, (int)(n + 1), st);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, count, off, (int)(n + 1), st);
    uint32_t total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, off + n, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
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
    e = hipMalloc(&flags, (nslots + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&vid, (nslots + 1) * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::firstFlags, dim3(gridFor(n)), dim3(256), 0, st, info, count, off, C, ntri, flags);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemsetAsync(flags + nslots, 0, 4, st);
    size_t tmp2 = 0;
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, flags, vid, (int)(nslots + 1), st);
    if (e == hipSuccess && tmp2 > tmp_bytes) {
        (void)hipFree(tmp);
        tmp = nullptr;
        tmp_bytes = tmp2;
        e = hipMalloc(&tmp, tmp_bytes);
    }
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flags, vid, (int)(nslots + 1), st);
    uint32_t nvert = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nvert, vid + nslots, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMalloc(&res->pos, (size_t)nvert * 12 + 12);
    if (e == hipSuccess) e = hipMalloc(&res->nrm, (size_t)nvert * 12 + 12);
    if (e == hipSuccess) e = hipMalloc(&res->idx, nslots * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mc::writeMesh<S>, dim3(gridFor(n)), dim3(256), 0, st, V, info, count, off, vid, ntri,
                           res->pos, res->nrm, res->idx);
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
