// kfusion::export_io::PLYExporter for the MI355X backend: header-only host C++ writing a
// meshing::MeshData (the GPU mesher's output, include/meshing/MarchingCubes.h) as PLY,
// with the reference's schema and refusal rules (include/export/PLYExporter.h,
// src/export/PLYExporter.cpp of iphands/azu):
//   header  "ply", "format <binary_little_endian|ascii> 1.0", "comment KinectFusionQt
//           export", vertex element with float x y z [nx ny nz] [uchar red green blue],
//           face element "property list uchar int vertex_indices", "end_header";
//   binary  explicit little-endian bytes (a big-endian host is refused);
//   ascii   one record per line, floats as "%.9g" of the widened double;
//   refusal a mesh that fails MeshData::validate is reported on stderr and no file is
//           created; a write or close failure removes the partial file.
// The whole file is built in memory and written with one call.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>

#include "../meshing/MarchingCubes.h"

namespace kfusion {
namespace export_io {

class PLYExporter {
public:
    static bool writeBinary(const meshing::MeshData& mesh, const std::string& filepath) {
        const uint32_t one = 1;
        unsigned char first;
        std::memcpy(&first, &one, 1);
        if (first != 1) {
            std::cerr << "[PLY] Big-endian host: refusing to write a little-endian payload to " << filepath << "\n";
            return false;
        }
        Layout l;
        if (!plan(mesh, filepath, l)) return false;
        std::string out = header(l, "binary_little_endian");
        out.reserve(out.size() + l.nvert * l.stride() + l.nface * 13);
        auto u32 = [&](uint32_t v) {
            for (int k = 0; k < 4; ++k) out.push_back(static_cast<char>((v >> (8 * k)) & 0xFFu));
        };
        auto f32 = [&](float f) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u32(u);
        };
        for (size_t i = 0; i < l.nvert; ++i) {
            for (int c = 0; c < 3; ++c) f32(mesh.positions[i][c]);
            if (l.normals)
                for (int c = 0; c < 3; ++c) f32(mesh.normals[i][c]);
            if (l.colors)
                for (int c = 0; c < 3; ++c) out.push_back(static_cast<char>(mesh.colors[3 * i + c]));
        }
        for (size_t t = 0; t < l.nface; ++t) {
            out.push_back(static_cast<char>(3));
            for (int c = 0; c < 3; ++c) u32(mesh.indices[3 * t + c]);
        }
        if (!writeAll(out, filepath, true)) return false;
        std::cout << "[PLY] Exported " << l.nvert << " vertices, " << l.nface << " triangles to: " << filepath
                  << "\n";
        return true;
    }

    static bool writeASCII(const meshing::MeshData& mesh, const std::string& filepath) {
        Layout l;
        if (!plan(mesh, filepath, l)) return false;
        std::string out = header(l, "ascii");
        char buf[40];
        auto num = [&](float v, bool first) {
            if (!first) out.push_back(' ');
            std::snprintf(buf, sizeof(buf), "%.9g", static_cast<double>(v));
            out += buf;
        };
        for (size_t i = 0; i < l.nvert; ++i) {
            for (int c = 0; c < 3; ++c) num(mesh.positions[i][c], c == 0);
            if (l.normals)
                for (int c = 0; c < 3; ++c) num(mesh.normals[i][c], false);
            if (l.colors)
                for (int c = 0; c < 3; ++c) out += " " + std::to_string(static_cast<int>(mesh.colors[3 * i + c]));
            out.push_back('\n');
        }
        for (size_t t = 0; t < l.nface; ++t) {
            out += "3";
            for (int c = 0; c < 3; ++c) out += " " + std::to_string(mesh.indices[3 * t + c]);
            out.push_back('\n');
        }
        if (!writeAll(out, filepath, false)) return false;
        std::cout << "[PLY] ASCII exported " << l.nvert << " vertices, " << l.nface << " triangles to: " << filepath
                  << "\n";
        return true;
    }

private:
    struct Layout {
        size_t nvert = 0, nface = 0;
        bool normals = false, colors = false;
        size_t stride() const { return 12 + (normals ? 12 : 0) + (colors ? 3 : 0); }
    };

    static bool plan(const meshing::MeshData& mesh, const std::string& filepath, Layout& l) {
        std::string why;
        if (!mesh.validate(&why)) {
            std::cerr << "[PLY] Rejected " << filepath << ": " << why << "\n";
            return false;
        }
        l.nvert = mesh.positions.size();
        l.nface = mesh.indices.size() / 3;
        l.normals = mesh.normals.size() == l.nvert;
        l.colors = mesh.colors.size() == 3 * l.nvert;
        return true;
    }

    static std::string header(const Layout& l, const char* format) {
        std::string h = "ply\nformat ";
        h += format;
        h += " 1.0\ncomment KinectFusionQt export\n";
        h += "element vertex " + std::to_string(l.nvert) + "\n";
        h += "property float x\nproperty float y\nproperty float z\n";
        if (l.normals) h += "property float nx\nproperty float ny\nproperty float nz\n";
        if (l.colors) h += "property uchar red\nproperty uchar green\nproperty uchar blue\n";
        h += "element face " + std::to_string(l.nface) + "\n";
        h += "property list uchar int vertex_indices\nend_header\n";
        return h;
    }

    static bool writeAll(const std::string& out, const std::string& filepath, bool binary) {
        std::ofstream f(filepath, binary ? (std::ios::binary | std::ios::trunc) : std::ios::trunc);
        if (!f.is_open()) {
            std::cerr << "[PLY] Cannot open: " << filepath << "\n";
            return false;
        }
        f.write(out.data(), static_cast<std::streamsize>(out.size()));
        f.flush();
        const bool wrote = f.good();
        f.close();
        if (!wrote || f.fail()) {
            std::cerr << "[PLY] Write error on: " << filepath << "\n";
            std::error_code ec;
            std::filesystem::remove(filepath, ec);
            return false;
        }
        return true;
    }
};

}  // namespace export_io
}  // namespace kfusion
