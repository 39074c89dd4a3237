"""The C++ drop-in (include/stq/gpu/stq_dropin.hpp) compiles with plain g++ (no
HIP headers), links against libstq_amd.so and, on the GPU, returns the
reference pair set when called like the reference's app/main.cu:141-144."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from stq_amd import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_demo(tmp_path):
    exe = str(tmp_path / "dropin_demo")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_demo.cpp"), "-L", libdir, "-lstq_amd",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    return exe


def test_dropin_compiles_without_hip(tmp_path):
    exe = build_demo(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_matches_oracle(tmp_path):
    exe = build_demo(tmp_path)
    b = scenes.cloth_scene(g=80, seed=3, sigma=1e-2)
    bin_in, bin_out = tmp_path / "boxes.bin", tmp_path / "pairs.bin"
    b.tofile(bin_in)
    r = subprocess.run([exe, str(bin_in), str(bin_out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(bin_out, dtype=np.int32).reshape(-1, 2)
    ref = oracle.reference_set(b)
    assert f"pairs {len(ref)}" in r.stdout
    assert np.array_equal(oracle.canonical(got), ref)
