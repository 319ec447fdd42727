"""Builds the native libraries in-tree.

* ``azu_amd/libazu_hip.so``: the gfx950 kernels and the C ABI of include/azu_hip.h
  (hipcc, one object per source, linked as a shared library).
* ``azu_amd/libazu_scene.so``: the synthetic depth generator (host-only, g++).
* ``oracle/liboracle.so``: the CPU oracle used by the tests (make in oracle/).
* ``tests/cpp/bin/host_classes``: the C++ host classes driven through a frame loop and the
  reference's volume / tracker contracts (g++, linked against libazu_hip.so).

Every HIP translation unit is compiled with -ffp-contract=off and IEEE-exact fp32
division / square root (the FP32 parity kernels rely on both) and without packed-FP32
instructions.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")

HIP_SOURCES = ["capi.hip", "integrate.hip", "raycast.hip", "frame.hip", "icp.hip", "adjoint.hip", "mc.hip", "condition.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AZU_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), p.stdout))
    return p.stdout


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(srcs, objs)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(HERE, "libazu_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_scene():
    lib = os.path.join(HERE, "libazu_scene.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", lib + ".tmp", os.path.join(CSRC, "scene.cpp")])
    os.replace(lib + ".tmp", lib)
    return lib


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_test_programs():
    """tests/cpp/bin/host_classes: a C++ program driving the host classes of include/
    (run by tests/test_gpu_host_classes.py)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "bin")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "host_classes")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "host_classes.cpp"), "-o", exe + ".tmp", "-L", HERE, "-lazu_hip",
          "-Wl,-rpath,$ORIGIN/../../../azu_amd"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_all(verbose=False):
    return [build_hip(verbose), build_scene(), build_oracle(), build_test_programs()]


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print("built", p)
