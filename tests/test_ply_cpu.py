"""PLY export of meshes (include/export/PLYExporter.h; reference
src/export/PLYExporter.cpp): the C++ cases of tests/ply_writer_contract.cpp, then the
written files parsed here against the schema the reference writes (exact header text,
little-endian records, "%.9g" ASCII floats) and the mesh they came from."""
import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

POS = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.5], [-0.25, 0.1, 0.75], [1e-7, -3.5, 123.456]], np.float32)
NRM = np.array([[0, 0, 1], [0, 0, -1], [0.6, 0.8, 0], [1, 0, 0]], np.float32)
COL = np.array([[255, 0, 10], [1, 2, 3], [128, 64, 32], [7, 8, 9]], np.uint8)
IDX = np.array([[0, 1, 2], [2, 3, 0]], np.uint32)


def header(fmt, normals, colors):
    h = "ply\nformat %s 1.0\ncomment KinectFusionQt export\nelement vertex 4\n" % fmt
    h += "property float x\nproperty float y\nproperty float z\n"
    if normals:
        h += "property float nx\nproperty float ny\nproperty float nz\n"
    if colors:
        h += "property uchar red\nproperty uchar green\nproperty uchar blue\n"
    return h + "element face 2\nproperty list uchar int vertex_indices\nend_header\n"


def run_cases(tmp_path):
    exe = tmp_path / "ply"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "ply_writer_contract.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASS" in out.stdout


def test_ply_files_match_the_schema_and_the_mesh(tmp_path):
    run_cases(tmp_path)
    for normals, colors, name in [(True, True, "full"), (False, False, "bare")]:
        b = (tmp_path / (name + ".bin.ply")).read_bytes()
        h = header("binary_little_endian", normals, colors).encode()
        assert b[:len(h)] == h
        body = bytearray()
        for i in range(4):
            body += struct.pack("<3f", *POS[i])
            if normals:
                body += struct.pack("<3f", *NRM[i])
            if colors:
                body += bytes(COL[i])
        for t in IDX:
            body += struct.pack("<BIII", 3, *t)
        assert b[len(h):] == bytes(body)
        a = (tmp_path / (name + ".ascii.ply")).read_text()
        h = header("ascii", normals, colors)
        assert a.startswith(h)
        lines = a[len(h):].splitlines()
        assert len(lines) == 6
        for i in range(4):
            f = lines[i].split()
            vals = ["%.9g" % float(v) for v in POS[i]]
            if normals:
                vals += ["%.9g" % float(v) for v in NRM[i]]
            if colors:
                vals += [str(int(v)) for v in COL[i]]
            assert f == vals
            assert np.array_equal(np.float32([float(x) for x in f[:3]]), POS[i])  # round trip
        assert lines[4:] == ["3 0 1 2", "3 2 3 0"]
