// kfusion::meshing::MarchingCubes for the MI355X backend: a header-only C++ host class
// over the C ABI of include/azu_hip.h, shaped like the reference's mesher
// (include/meshing/MarchingCubes.h:32-60, MeshData.h:13-40 of iphands/azu). Extraction
// runs on the GPU and produces the CPU extractor's mesh: the same vertices, normals and
// indices in the same order, and the same truncation rule under the triangle cap.
// Colour is not stored by the volume, so MeshData::colors stays empty.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include <cmath>
#include <string>

#include "../azu_hip.h"
#include "../azu_types.h"
#include "../tsdf/TSDFVolume.h"

namespace kfusion {
namespace meshing {

// MeshData (MeshData.h:13-40): positions / normals are Eigen::Vector3f (kfusion::Vec3f),
// packed xyz floats.
struct MeshData {
    std::vector<Vec3f> positions;
    std::vector<Vec3f> normals;
    std::vector<uint8_t> colors;
    std::vector<uint32_t> indices;
    bool truncated = false;

    void clear() {
        positions.clear();
        normals.clear();
        colors.clear();
        indices.clear();
        truncated = false;
    }
    bool empty() const { return positions.empty(); }
    size_t triangleCount() const { return indices.size() / 3; }
    bool hasColors() const { return !colors.empty(); }

    // MeshData::validate (reference include/meshing/MeshData.h:49-97): the mesh can be
    // indexed and serialised without out-of-bounds reads or non-finite coordinates.
    // Checks in this order and reports the first failure through `reason` (may be
    // null): positions empty, indices empty, index count not a multiple of 3, an index
    // out of range, a non-finite position, a normal count neither 0 nor the vertex
    // count, a non-finite normal, a colour count neither 0 nor 3 per vertex.
    bool validate(std::string* reason) const {
        auto reject = [&](const std::string& why) {
            if (reason) *reason = why;
            return false;
        };
        auto finite3 = [](const Vec3f& v) {
            return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
        };
        const size_t nv = positions.size(), ni = indices.size();
        if (nv == 0) return reject("positions is empty");
        if (ni == 0) return reject("indices is empty while positions.size()=" + std::to_string(nv));
        if (ni % 3 != 0) return reject("indices.size()=" + std::to_string(ni) + " is not a multiple of 3");
        for (size_t i = 0; i < ni; ++i)
            if (static_cast<size_t>(indices[i]) >= nv)
                return reject("indices[" + std::to_string(i) + "]=" + std::to_string(indices[i]) +
                              " is out of range for positions.size()=" + std::to_string(nv));
        for (size_t i = 0; i < nv; ++i)
            if (!finite3(positions[i])) return reject("positions[" + std::to_string(i) + "] has a non-finite component");
        if (!normals.empty() && normals.size() != nv)
            return reject("normals.size()=" + std::to_string(normals.size()) + " is neither 0 nor positions.size()=" +
                          std::to_string(nv));
        for (size_t i = 0; i < normals.size(); ++i)
            if (!finite3(normals[i])) return reject("normals[" + std::to_string(i) + "] has a non-finite component");
        const size_t nc = colors.size();
        if (nc != 0 && (nc % 3 != 0 || nc / 3 != nv))
            return reject("colors.size()=" + std::to_string(nc) + " is neither 0 nor 3 * positions.size() (" +
                          std::to_string(nv) + " vertices)");
        return true;
    }
};

using ProgressCallback = std::function<void(float)>;

class MarchingCubes {
public:
    // The reference's CPU and GPU triangle budget (MarchingCubes.h:48-56).
    void setMaxTriangles(size_t n) { max_triangles_ = n; }
    size_t maxTriangles() const { return max_triangles_; }

    // extractGPU: the mesh on the device, downloaded into MeshData. Runs on the volume's
    // stream (TSDFVolume::setStream), after the work queued there.
    std::shared_ptr<MeshData> extractGPU(const tsdf::TSDFVolume& volume) {
        azu_mesh* m = nullptr;
        detail::check(azu_mc_extract(volume.handle(), max_triangles_, &m, volume.stream()),
                      "MarchingCubes::extractGPU");
        auto out = std::make_shared<MeshData>();
        unsigned long long nv = 0, nt = 0;
        int tr = 0;
        int rc = azu_mesh_info(m, &nv, &nt, &tr);
        if (rc == AZU_OK) {
            out->positions.resize(nv);
            out->normals.resize(nv);
            out->indices.resize(3 * nt);
            out->truncated = tr != 0;
            rc = azu_mesh_download(m, out->positions.empty() ? nullptr : out->positions[0].data(),
                                   out->normals.empty() ? nullptr : out->normals[0].data(),
                                   out->indices.empty() ? nullptr : out->indices.data());
        }
        azu_mesh_destroy(m);
        detail::check(rc, "MarchingCubes::extractGPU download");
        return out;
    }

    // extract (MarchingCubes.h:37-38): the same GPU extraction; the callback, if any, is
    // called once with 1.0 from the calling thread when the mesh is complete.
    std::shared_ptr<MeshData> extract(const tsdf::TSDFVolume& volume, ProgressCallback cb = nullptr) {
        auto m = extractGPU(volume);
        if (cb) cb(1.0f);
        return m;
    }

private:
    size_t max_triangles_ = 2000000;
};

}  // namespace meshing
}  // namespace kfusion
