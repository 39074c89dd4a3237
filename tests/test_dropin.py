"""The C++ drop-ins compile with plain g++ (no HIP headers) and link against
libstq_amd.so.  On the GPU:
  * include/stq/gpu/stq_dropin.hpp, called like the reference's
    app/main.cu:141-144 in a loop over tidstart, returns the reference pair
    set: whole scene in one call, a preset MAX_OVERLAP_CUTOFF (several row
    ranges), and a memlimit that makes the sizing halve the cutoff
    (src/memory.cuh:171-188);
  * include/stq/cpu/stq_cpu_dropin.hpp, called exactly as app/main.cpp:147-151
    calls sweep_cpu_single_batch, returns (a.id, b.id) pairs with a sorted
    before b, whose set is the reference set."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from stq_amd import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp_path, name):
    exe = str(tmp_path / name)
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-L", libdir, "-lstq_amd",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    return exe


def run_demo(tmp_path, name, boxes, *extra):
    exe = build(tmp_path, name)
    bin_in, bin_out = tmp_path / "boxes.bin", tmp_path / "pairs.bin"
    boxes.tofile(bin_in)
    r = subprocess.run([exe, str(bin_in), str(bin_out)] + [str(x) for x in extra], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return np.fromfile(bin_out, dtype=np.int32).reshape(-1, 2), r.stdout


def field(stdout, key):
    toks = stdout.split()
    return int(toks[toks.index(key) + 1])


def test_dropins_compile_without_hip(tmp_path):
    for name in ("dropin_demo", "dropin_cpu_demo"):
        assert os.path.exists(build(tmp_path, name))


@pytest.mark.gpu
def test_dropin_matches_oracle(tmp_path):
    b = scenes.cloth_scene(g=80, seed=3, sigma=1e-2)
    got, out = run_demo(tmp_path, "dropin_demo", b)
    ref = oracle.reference_set(b)
    assert f"pairs {len(ref)}" in out and field(out, "calls") == 1
    assert np.array_equal(oracle.canonical(got), ref)


@pytest.mark.gpu
def test_dropin_preset_cutoff_row_ranges(tmp_path):
    b = scenes.cloth_scene(g=120, seed=4, sigma=1e-2)
    n = len(b)
    got, out = run_demo(tmp_path, "dropin_demo", b, 0, n // 7)
    ref = oracle.reference_set(b)
    assert field(out, "calls") == 8  # ceil(n / (n // 7)) ranges
    assert len(got) == len(ref), "a pair reported by two ranges, or missed"
    assert np.array_equal(oracle.canonical(got), ref)


@pytest.mark.gpu
def test_dropin_memlimit_halves_the_cutoff(tmp_path):
    import stq_amd
    with stq_amd.Context(0) as ctx:
        import ctypes
        free_b, total_b = ctypes.c_size_t(0), ctypes.c_size_t(0)
        assert ctx._L.stq_device_memory(ctx.ptr, ctypes.byref(free_b), ctypes.byref(total_b)) == 0
        used_gib = (total_b.value - free_b.value) / 2 ** 30
    # a limit one or two GiB above the memory in use: room for at most ~7.7M
    # pairs with their narrow-phase records, fewer than the scene has
    memlimit = int(np.ceil(used_gib)) + 1
    b = scenes.cloth_scene(g=560, seed=5)
    ref = oracle.reference_set(b)
    assert len(ref) > 8_000_000
    got, out = run_demo(tmp_path, "dropin_demo", b, memlimit)
    assert field(out, "calls") > 1 and field(out, "cutoff") < len(b)
    assert len(got) == len(ref)
    assert np.array_equal(oracle.canonical(got), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [0, 4])
def test_cpu_dropin_single_batch_loop(tmp_path, batch):
    b = scenes.cloth_scene(g=90, seed=6, sigma=1e-2)
    n = len(b)
    extra = [n // batch] if batch else []
    got, out = run_demo(tmp_path, "dropin_cpu_demo", b, *extra)
    ref = oracle.reference_set(b)
    assert field(out, "count") == len(ref) and field(out, "left") == 0
    assert field(out, "batches") >= (batch + 1 if batch else 2)
    assert np.array_equal(oracle.canonical(got), ref)
    # the reference orientation: a is sorted before b (min x, then position)
    ka = b["min"][got[:, 0], 0]
    kb = b["min"][got[:, 1], 0]
    assert np.all((ka < kb) | ((ka == kb) & (got[:, 0] < got[:, 1])))
