"""CPU checks of the C ABI: the library loads and exports every entry point that
include/azu_hip.h declares (no compute calls: this container has no GPU), and the C++
host headers compile against it."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "azu_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(azu_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "azu_integrate" in syms and "azu_raycast" in syms and "azu_tracker_track" in syms
    assert len(syms) >= 40


def test_library_exports_every_declared_symbol():
    from azu_amd import native
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (azu_[a-z0-9_]+)\b", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_every_declared_symbol():
    from azu_amd import native
    bound = set(native._SIGS) | set(native._EXTRA_SIGS)
    missing = [s for s in declared_symbols() if s not in bound]
    assert not missing, missing


def test_abi_version_and_error_string():
    from azu_amd import native
    assert native.lib.azu_abi_version() == 1
    assert isinstance(native.lib.azu_last_error(), bytes)


def test_argument_errors_without_device():
    """Argument validation happens before any device work, so it is testable here."""
    import ctypes
    from azu_amd import native as N
    h = ctypes.c_void_p()
    assert N.lib.azu_vol_create(None, 0, ctypes.byref(h)) == N.AZU_ERR_ARG
    p = N.TsdfParams()
    p.resolution, p.voxel_size, p.truncation, p.max_weight = 64, 0.01, 0.03, 0.5
    assert N.lib.azu_vol_create(ctypes.byref(p), 0, ctypes.byref(h)) == N.AZU_ERR_ARG
    assert b"max_weight" in N.lib.azu_last_error()
    p.max_weight = 128.0
    assert N.lib.azu_vol_create(ctypes.byref(p), 7, ctypes.byref(h)) == N.AZU_ERR_ARG
    K = (ctypes.c_float * 4)(525, 525, 319.5, 239.5)
    ip = N.IcpParams()
    assert N.lib.azu_tracker_create(2, 2, K, ctypes.byref(ip), ctypes.byref(h)) == N.AZU_ERR_ARG


# A per-frame loop written against the host classes the way the reference's pipeline
# uses TSDFVolume / ICPTracker / MarchingCubes; compiled and linked, not run.
CPP_LOOP = r"""
#include "tsdf/TSDFVolume.h"
#include "tracking/ICPTracker.h"
#include "tracking/TrackingPolicy.h"
#include "meshing/MarchingCubes.h"
#include "export/PLYExporter.h"
#include <array>
#include <cstdio>
struct Mat4 { float m[16]; const float* data() const { return m; } };
int run(const float* d_depth, int W, int H, int frames) {
    using namespace kfusion;
    tsdf::TSDFParams p;
    p.resolution = 512;
    p.voxel_size = 0.005f;
    tsdf::TSDFVolume vol(p, tsdf::Storage::F16_U16);
    tracking::ICPTracker icp;
    tracking::CameraIntrinsics k;
    icp.setIntrinsics(k);
    tracking::ModelFrame model(W, H);
    Mat4 pose{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    for (int f = 0; f < frames; ++f) {
        bool integrate = true;
        if (f > 0) {
            model.render(vol, pose, k);
            tracking::ICPResult r = icp.trackGPU(d_depth, nullptr, W, H, model, pose, model.pose);
            const tracking::TrackQuality q = tracking::classifyTracking(r, pose);
            if (q != tracking::TrackQuality::Failed) {
                const tracking::ObservableMotion m = tracking::keepObservableMotion(pose.m, r.pose.data(),
                                                                                    r.information.data());
                for (int i = 0; i < 16; ++i) pose.m[i] = m.pose.data()[i];
            }
            integrate = q == tracking::TrackQuality::Good;
        }
        if (integrate) vol.integrateGPU(d_depth, nullptr, pose, k.fx, k.fy, k.cx, k.cy, W, H);
    }
    meshing::MarchingCubes mc;
    auto mesh = mc.extract(vol, [](float x) { std::printf("%f\n", x); });
    std::vector<Vec3f> cloud;
    std::vector<uint8_t> colors;
    vol.extractGlobalPointCloud(cloud, colors);
    export_io::PLYExporter::writeBinary(*mesh, "mesh.ply");
    return static_cast<int>(mesh->triangleCount());
}
int main(int argc, char**) { return argc > 99 ? run(nullptr, 640, 480, 2) : 0; }
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "include", "tsdf", "TSDFVolume.h")),
                    reason="host headers absent")
def test_cpp_host_headers_compile(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(CPP_LOOP)
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", os.path.join(ROOT, "azu_amd"), "-lazu_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "azu_amd")], check=True)
