"""CPU builds of the host-side parsers (csrc/mesh_io.cpp: OBJ, PLY and the
ground-truth JSON, which read untrusted files) and of the oracle with
AddressSanitizer and UndefinedBehaviorSanitizer (g++/gcc, host code only),
run on malformed files and random scenes.  Every malformed file must give
STQ_ERR_INVALID_ARGUMENT (never STQ_ERR_INTERNAL) with no sanitizer report;
the well-formed fixtures read as before."""
import os
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def build(tmp_path, out, cmd):
    exe = str(tmp_path / out)
    subprocess.run(cmd + ["-o", exe], check=True, capture_output=True)
    return exe


def no_report(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text and "LeakSanitizer" not in text, text


MALFORMED = {
    "quad.obj": b"v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n",
    "zero_index.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
    "huge_index.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 99999999999\n",
    "huge_negative.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -99999999999\n",
    "overflow_index.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 999999999999999999999999\n",
    "past_end.obj": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
    "bad_vertex.obj": b"v 0 zero 0\n",
    "not_ply.ply": b"plx\nformat ascii 1.0\nend_header\n",
    "big_endian.ply": b"ply\nformat binary_big_endian 1.0\nend_header\n",
    "no_header_end.ply": b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n",
    "negative_count.ply": b"ply\nformat ascii 1.0\nelement vertex -3\nproperty float x\nend_header\n",
    "no_xyz.ply": b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n1\n",
    "unknown_type.ply": b"ply\nformat ascii 1.0\nelement vertex 1\nproperty quux x\nend_header\n1\n",
    "truncated_ascii.ply": (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                            b"property float z\nend_header\n0 0 0\n1 0\n"),
    "huge_list.ply": (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                      b"property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
                      b"0 0 0\n1 0 0\n0 1 0\n9999999 0 1 2\n"),
    "face_index_double.ply": (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                              b"property float z\nelement face 1\nproperty list uchar double vertex_indices\n"
                              b"end_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 1e300\n"),
    "face_index_fraction.ply": (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                                b"property float z\nelement face 1\nproperty list uchar float vertex_indices\n"
                                b"end_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 1.5\n"),
    "face_index_negative.ply": (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                                b"property float z\nelement face 1\nproperty list uchar int vertex_indices\n"
                                b"end_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 -2\n"),
    "truncated_binary.ply": (b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty double x\n"
                             b"property double y\nproperty double z\nend_header\n" + struct.pack("<5d", 0, 0, 0, 1, 0)),
    "nan_index_binary.ply": (b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\n"
                             b"property float y\nproperty float z\nelement face 1\n"
                             b"property list uchar float vertex_indices\nend_header\n"
                             + struct.pack("<9f", 0, 0, 0, 1, 0, 0, 0, 1, 0) + struct.pack("<B3f", 3, 0, 1, float("nan"))),
    "overflow.json": b"[[1, 99999999999999999999999]]",
    "sign_only.json": b"[[1, -]]",
    "truncated.json": b"[[1, 2], [3",
    "not_pairs.json": b"[[1, 2, 3]]",
    "garbage.json": b"\x00\xff\x10 not json",
    "empty.json": b"",
}


def test_mesh_io_under_sanitizers(tmp_path):
    exe = build(tmp_path, "mesh_io_driver",
                ["g++", "-std=c++17"] + SAN + ["-I", os.path.join(ROOT, "include"),
                                               os.path.join(ROOT, "tests", "cpp", "mesh_io_driver.cpp"),
                                               os.path.join(ROOT, "broadphase-gpu_amd", "csrc", "mesh_io.cpp")])
    paths = []
    for name, data in MALFORMED.items():
        p = tmp_path / name
        p.write_bytes(data)
        paths.append(str(p))
    good = [os.path.join(GOLDEN, f) for f in ("tetra.obj", "tetra_ascii.ply", "tetra_binary.ply",
                                               "truth_tiny.json")]
    (tmp_path / "empty_mesh.obj").write_bytes(b"")
    good_empty = [str(tmp_path / "empty_mesh.obj")]
    r = subprocess.run([exe] + paths + good + good_empty, capture_output=True, text=True, timeout=120, env=ENV)
    no_report(r)
    assert r.returncode == 0, r.stderr
    status = {os.path.basename(ln.split()[0]): int(ln.split()[1]) for ln in r.stdout.splitlines()}
    print(status)
    for name in MALFORMED:
        assert status[name] == 1, (name, status[name])  # STQ_ERR_INVALID_ARGUMENT
    for g in good:
        assert status[os.path.basename(g)] == 0, g
    assert status["empty_mesh.obj"] == 0  # an empty file is an empty mesh


def test_oracle_under_sanitizers(tmp_path):
    exe = build(tmp_path, "oracle_driver",
                ["gcc", "-std=c11", "-fopenmp"] + SAN + [os.path.join(ROOT, "tests", "cpp", "oracle_driver.c"),
                                                         os.path.join(ROOT, "oracle", "stq_oracle.c"), "-lm"])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    no_report(r)
    assert r.returncode == 0 and "oracle driver: ok" in r.stdout, r.stdout + r.stderr
