// PLY writer cases (include/export/PLYExporter.h), after the reference's
// tests/ply_writer_contract.cpp: a valid mesh is written in both formats (the files are
// then parsed by tests/test_ply_cpu.py), repeated writes are byte-identical, and each
// invalid mesh class is refused with its reason and creates no file. usage: <out dir>.
#include "export/PLYExporter.h"

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <limits>
#include <string>

using kfusion::export_io::PLYExporter;
using kfusion::meshing::MeshData;

namespace {
int g_fail = 0, g_checks = 0;
void check(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++g_fail;
    }
}
std::string slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}
MeshData quad(bool normals, bool colors) {
    MeshData m;
    m.positions = {{1.0f, 0.0f, 0.5f}, {0.0f, 1.0f, 0.5f}, {-0.25f, 0.1f, 0.75f}, {1e-7f, -3.5f, 123.456f}};
    if (normals) m.normals = {{0, 0, 1}, {0, 0, -1}, {0.6f, 0.8f, 0}, {1, 0, 0}};
    if (colors) m.colors = {255, 0, 10, 1, 2, 3, 128, 64, 32, 7, 8, 9};
    m.indices = {0, 1, 2, 2, 3, 0};
    return m;
}
void refused(const MeshData& m, const std::string& dir, const std::string& tag, const std::string& reason) {
    std::string why;
    check(!m.validate(&why), tag + ": invalid");
    check(why.find(reason) != std::string::npos, tag + ": reason '" + why + "' names " + reason);
    const std::string b = dir + "/" + tag + ".bin.ply", a = dir + "/" + tag + ".ascii.ply";
    check(!PLYExporter::writeBinary(m, b) && !std::filesystem::exists(b), tag + ": binary refused, no file");
    check(!PLYExporter::writeASCII(m, a) && !std::filesystem::exists(a), tag + ": ascii refused, no file");
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const MeshData full = quad(true, true), bare = quad(false, false);
    check(PLYExporter::writeBinary(full, dir + "/full.bin.ply"), "full binary written");
    check(PLYExporter::writeASCII(full, dir + "/full.ascii.ply"), "full ascii written");
    check(PLYExporter::writeBinary(bare, dir + "/bare.bin.ply"), "bare binary written");
    check(PLYExporter::writeASCII(bare, dir + "/bare.ascii.ply"), "bare ascii written");
    check(PLYExporter::writeBinary(full, dir + "/again.bin.ply") &&
              slurp(dir + "/again.bin.ply") == slurp(dir + "/full.bin.ply"),
          "binary byte-identical on repeat");
    check(PLYExporter::writeASCII(full, dir + "/again.ascii.ply") &&
              slurp(dir + "/again.ascii.ply") == slurp(dir + "/full.ascii.ply"),
          "ascii byte-identical on repeat");

    refused(MeshData{}, dir, "empty", "positions is empty");
    MeshData m = bare;
    m.indices.clear();
    refused(m, dir, "no_indices", "indices is empty");
    m = bare;
    m.indices.push_back(1);
    refused(m, dir, "trailing_index", "not a multiple of 3");
    m = bare;
    m.indices[4] = 4;
    refused(m, dir, "out_of_range", "out of range");
    m = bare;
    m.positions[2][1] = std::numeric_limits<float>::quiet_NaN();
    refused(m, dir, "non_finite", "non-finite");
    m = full;
    m.normals.pop_back();
    refused(m, dir, "normal_count", "normals.size()");
    m = full;
    m.normals[1][0] = std::numeric_limits<float>::infinity();
    refused(m, dir, "normal_non_finite", "non-finite");
    m = full;
    m.colors.pop_back();
    refused(m, dir, "color_count", "colors.size()");

    std::printf("ply_writer_contract: %s (%d checks)\n", g_fail ? "FAIL" : "PASS", g_checks);
    return g_fail ? 1 : 0;
}
