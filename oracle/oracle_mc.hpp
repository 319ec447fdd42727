// TEST INFRASTRUCTURE ONLY: CPU oracle of MarchingCubes::extract
// (src/meshing/MarchingCubes.cpp:115-382) over a float tsdf / weight volume. Output is
// welded positions, normals and a triangle index list in the reference's order:
// slices ascending z, cubes y then x, triangles in table order, each emitted with the
// reversed (2,1,0) winding, vertices welded by edge identity with the first insertion
// winning. Vertex colours are not produced: the volume carries no colour (every
// vertex would get the empty colour's byte 128).
#pragma once

#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../azu_amd/csrc/mc_tables.h"

namespace orc {

struct MeshOut {
    std::vector<float> pos, nrm;  // xyz triples
    std::vector<uint32_t> idx;
    bool truncated = false;
};

namespace mcd {
struct Corner {
    float value;
    bool supported;
};
// sampleCorner (MarchingCubes.cpp:81-85): weight <= 0.001 or non-finite -> (+1, unsupported)
inline Corner sampleCorner(float tsdf, float weight) {
    if (!std::isfinite(tsdf) || weight <= 0.001f) return {1.0f, false};
    return {tsdf, true};
}
struct N3 {
    float x, y, z;
};
inline bool fin(N3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
inline float nrm3(N3 v) { return std::sqrt(v.x * v.x + (v.y * v.y + v.z * v.z)); }
}  // namespace mcd

// extract (MarchingCubes.cpp:174-382)
inline MeshOut mcExtract(const float* tsdf, const float* weight, int R, float vs, const float* origin,
                         size_t max_tris) {
    using namespace mcd;
    MeshOut out;
    if (R < 2) return out;
    const size_t sy = static_cast<size_t>(R), sz = static_cast<size_t>(R) * R;
    auto at = [&](int x, int y, int z) { return static_cast<size_t>(z) * sz + static_cast<size_t>(y) * sy + x; };
    auto usable = [&](int x, int y, int z, float& v) {
        if (x < 0 || x >= R || y < 0 || y >= R || z < 0 || z >= R) return false;
        const Corner c = sampleCorner(tsdf[at(x, y, z)], weight[at(x, y, z)]);
        if (!c.supported) return false;
        v = c.value;
        return true;
    };
    const float nanf_ = std::numeric_limits<float>::quiet_NaN();
    // voxelNormal (MarchingCubes.cpp:136-172)
    auto voxelNormal = [&](int x, int y, int z) -> N3 {
        float self = 0;
        if (!usable(x, y, z, self)) return {nanf_, nanf_, nanf_};
        float g[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            const int ox = a == 0, oy = a == 1, oz = a == 2;
            float hi = 0, lo = 0;
            const bool hh = usable(x + ox, y + oy, z + oz, hi);
            const bool hl = usable(x - ox, y - oy, z - oz, lo);
            if (hh && hl) g[a] = 0.5f * (hi - lo);
            else if (hh) g[a] = hi - self;
            else if (hl) g[a] = self - lo;
            else g[a] = 0.0f;
        }
        const N3 gv{g[0], g[1], g[2]};
        if (!fin(gv)) return {nanf_, nanf_, nanf_};
        const float len = nrm3(gv);
        if (!std::isfinite(len) || len <= 1e-6f) return {nanf_, nanf_, nanf_};
        return {gv.x / len, gv.y / len, gv.z / len};
    };

    struct Key {
        int x, y, z, a;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z && a == o.a; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            return (static_cast<size_t>(k.z) * 1315423911u) ^ (static_cast<size_t>(k.y) * 2654435761u) ^
                   (static_cast<size_t>(k.x) * 97531u) ^ static_cast<size_t>(k.a);
        }
    };
    struct SliceOut {
        std::vector<N3> pos, nrm;
        std::vector<Key> key;
    };
    std::vector<SliceOut> slices(R - 1);

#pragma omp parallel for schedule(dynamic, 4)
    for (int z = 0; z < R - 1; ++z) {
        SliceOut& sl = slices[z];
        for (int y = 0; y < R - 1; ++y)
            for (int x = 0; x < R - 1; ++x) {
                Corner corner[8];
                for (int c = 0; c < 8; ++c) {
                    const size_t i = at(x + azu_mc::kCornerOffset[c][0], y + azu_mc::kCornerOffset[c][1],
                                        z + azu_mc::kCornerOffset[c][2]);
                    corner[c] = sampleCorner(tsdf[i], weight[i]);
                }
                int cube = 0;
                for (int c = 0; c < 8; ++c)
                    if (corner[c].value < 0.0f) cube |= 1 << c;
                const int emask = azu_mc::edgeMask(cube);
                if (emask == 0) continue;
                N3 cpos[8];
                for (int c = 0; c < 8; ++c)
                    cpos[c] = {origin[0] + static_cast<float>(x + azu_mc::kCornerOffset[c][0]) * vs,
                               origin[1] + static_cast<float>(y + azu_mc::kCornerOffset[c][1]) * vs,
                               origin[2] + static_cast<float>(z + azu_mc::kCornerOffset[c][2]) * vs};
                N3 cn[8];
                bool cn_done[8] = {};
                auto cornerNormal = [&](int c) {
                    if (!cn_done[c]) {
                        cn[c] = voxelNormal(x + azu_mc::kCornerOffset[c][0], y + azu_mc::kCornerOffset[c][1],
                                            z + azu_mc::kCornerOffset[c][2]);
                        cn_done[c] = true;
                    }
                    return cn[c];
                };
                N3 epos[12], enrm[12];
                Key ekey[12];
                bool eok[12] = {};
                for (int e = 0; e < 12; ++e) {
                    if (!(emask & (1 << e))) continue;
                    const int c0 = azu_mc::kEdgeCorners[e][0], c1 = azu_mc::kEdgeCorners[e][1];
                    if (!corner[c0].supported || !corner[c1].supported) continue;
                    int axis = 0, lo = c0, up = c1;
                    for (int a = 0; a < 3; ++a)
                        if (azu_mc::kCornerOffset[c0][a] != azu_mc::kCornerOffset[c1][a]) {
                            axis = a;
                            const bool c0low = azu_mc::kCornerOffset[c0][a] <= azu_mc::kCornerOffset[c1][a];
                            lo = c0low ? c0 : c1;
                            up = c0low ? c1 : c0;
                            break;
                        }
                    ekey[e] = Key{x + azu_mc::kCornerOffset[lo][0], y + azu_mc::kCornerOffset[lo][1],
                                  z + azu_mc::kCornerOffset[lo][2], axis};
                    // crossingFromLower (MarchingCubes.cpp:115-134)
                    const float vl = corner[lo].value, vu = corner[up].value;
                    const N3 pl = cpos[lo], pu = cpos[up];
                    if (!std::isfinite(vl) || !std::isfinite(vu) || !fin(pl) || !fin(pu)) continue;
                    N3 p;
                    float t;
                    if (std::abs(vl) < 1e-6f) { p = pl; t = 0.0f; }
                    else if (std::abs(vu) < 1e-6f) { p = pu; t = 1.0f; }
                    else {
                        const float diff = vl - vu;
                        if (!std::isfinite(diff) || std::abs(diff) < 1e-6f) { p = pl; t = 0.0f; }
                        else {
                            const float q = vl / diff;
                            t = std::max(0.0f, std::min(1.0f, q));
                            p = {pl.x + t * (pu.x - pl.x), pl.y + t * (pu.y - pl.y), pl.z + t * (pu.z - pl.z)};
                        }
                    }
                    const N3 nl = cornerNormal(lo), nu = cornerNormal(up);
                    if (!fin(nl) || !fin(nu)) continue;
                    const N3 bl{nl.x + t * (nu.x - nl.x), nl.y + t * (nu.y - nl.y), nl.z + t * (nu.z - nl.z)};
                    const float bl_len = nrm3(bl);
                    if (!std::isfinite(bl_len) || bl_len <= 1e-6f) continue;
                    epos[e] = p;
                    enrm[e] = {bl.x / bl_len, bl.y / bl_len, bl.z / bl_len};
                    eok[e] = true;
                }
                const char* row = azu_mc::kTriRows[cube];
                for (int t = 0; row[t] != '\0'; t += 3) {
                    int te[3];
                    for (int k = 0; k < 3; ++k) {
                        const char ch = row[t + k];
                        te[k] = (ch <= '9') ? ch - '0' : ch - 'a' + 10;
                    }
                    if (!eok[te[0]] || !eok[te[1]] || !eok[te[2]]) continue;
                    for (int i = 2; i >= 0; --i) {
                        sl.pos.push_back(epos[te[i]]);
                        sl.nrm.push_back(enrm[te[i]]);
                        sl.key.push_back(ekey[te[i]]);
                    }
                }
            }
    }
    std::unordered_map<Key, uint32_t, KeyHash> weld;
    size_t tri = 0;
    for (int z = 0; z < R - 1 && !out.truncated; ++z) {
        const SliceOut& sl = slices[z];
        for (size_t i = 0; i + 2 < sl.key.size(); i += 3) {
            if (tri >= max_tris) { out.truncated = true; break; }
            for (int k = 0; k < 3; ++k) {
                const Key& key = sl.key[i + k];
                auto it = weld.find(key);
                uint32_t gi;
                if (it != weld.end()) gi = it->second;
                else {
                    gi = static_cast<uint32_t>(out.pos.size() / 3);
                    weld.emplace(key, gi);
                    out.pos.insert(out.pos.end(), {sl.pos[i + k].x, sl.pos[i + k].y, sl.pos[i + k].z});
                    out.nrm.insert(out.nrm.end(), {sl.nrm[i + k].x, sl.nrm[i + k].y, sl.nrm[i + k].z});
                }
                out.idx.push_back(gi);
            }
            ++tri;
        }
    }
    return out;
}

}  // namespace orc
