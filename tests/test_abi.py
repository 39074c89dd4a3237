"""CPU-side checks of the C-ABI library: it loads, exports every entry point
that include/stq.h declares, and its structs match the header.  No compute
call is made (there is no GPU here)."""
import ctypes
import os
import re

import numpy as np

from stq_amd import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "stq.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(stq_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 16
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_struct_sizes_match_header():
    L = _lib.lib()
    assert L.stq_sizeof_aabb() == 72 == scenes.AABB_DTYPE.itemsize
    assert L.stq_sizeof_options() == ctypes.sizeof(_lib.Options) == 64
    assert L.stq_sizeof_stats() == ctypes.sizeof(_lib.Stats)


def test_aabb_field_offsets():
    d = scenes.AABB_DTYPE
    assert d.fields["id"][1] == 0
    assert d.fields["min"][1] == 8
    assert d.fields["max"][1] == 32
    assert d.fields["vids"][1] == 56
    assert d.fields["ref_id"][1] == 68


def test_defaults_and_status_strings():
    L = _lib.lib()
    o = _lib.Options()
    assert L.stq_options_default(ctypes.byref(o)) == 0
    assert (o.precision, o.axis_policy, o.sizing) == (0, 0, 0)
    for s in range(8):
        assert L.stq_status_string(s)
    assert L.stq_options_default(None) == _lib.STQ_ERR_INVALID_ARGUMENT


def test_null_arguments_rejected_without_gpu():
    L = _lib.lib()
    cnt = ctypes.c_int64()
    assert L.stq_find_pairs_device(None, None, 0, None, 0, ctypes.byref(cnt)) == _lib.STQ_ERR_INVALID_ARGUMENT
    assert L.stq_destroy(None) == _lib.STQ_ERR_INVALID_ARGUMENT
    assert L.stq_get_stats(None, None) == _lib.STQ_ERR_INVALID_ARGUMENT
    L.stq_free_host(None)


def test_product_library_does_not_link_the_oracle():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"oracle_" not in data
    assert b"liboracle" not in data


def test_scene_generators_shapes():
    c = scenes.cloth_scene(g=10, seed=0)
    V, E, F = scenes.cloth_mesh(10)
    assert len(c) == len(V) + len(E) + len(F)
    assert len(E) == 2 * 9 * 10 + 9 * 9 and len(F) == 2 * 9 * 9
    # the 708 x 708 cloth has exactly 3,001,923 boxes
    g = 708
    assert g * g + (2 * (g - 1) * g + (g - 1) ** 2) + 2 * (g - 1) ** 2 == 3_001_923
    u = scenes.uniform_scene(100, seed=0)
    assert np.all(u["min"] <= u["max"])
