"""Mesh input and the ground-truth recall count (host code of the library;
no GPU needed), against the hand-derived fixtures in tests/golden/.

The edge order follows igl::edges as restated in the library (by larger
index, then smaller); libigl is not in this image, so that order is checked
against the fixture and the rule, not against libigl itself (parity
unpinned for the order; the edge set is pinned).
"""
import json
import os

import numpy as np
import pytest

import stq_amd

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def expected():
    with open(os.path.join(GOLD, "mesh_expected.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["tetra.obj", "tetra_ascii.ply", "tetra_binary.ply"])
def test_read_mesh(expected, name):
    V, F = stq_amd.read_mesh(os.path.join(GOLD, name))
    assert np.array_equal(V, np.array(expected["V"]))
    assert np.array_equal(F, np.array(expected["F"]))
    assert np.array_equal(stq_amd.mesh_edges(F), np.array(expected["E"]))


def test_obj_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    V = rng.normal(size=(500, 3))
    F = rng.integers(0, 500, (900, 3))
    p = tmp_path / "m.obj"
    with open(p, "w") as fh:
        for v in V:
            fh.write("v %r %r %r\n" % tuple(float(x) for x in v))
        for f in F:
            fh.write("f %d %d %d\n" % tuple(f + 1))
    V2, F2 = stq_amd.read_mesh(p)
    assert np.array_equal(V2, V) and np.array_equal(F2, F)


def test_mesh_edges_rule():
    """Every face edge once, as (smaller, larger), ordered by (larger, smaller)."""
    rng = np.random.default_rng(4)
    F = rng.integers(0, 200, (1000, 3))
    E = stq_amd.mesh_edges(F)
    want = set()
    for f in F:
        for k in range(3):
            a, b = int(f[k]), int(f[(k + 1) % 3])
            if a != b:
                want.add((min(a, b), max(a, b)))
    assert len(E) == len(want) and set(map(tuple, E.tolist())) == want
    key = E[:, 1].astype(np.int64) * 1000 + E[:, 0]
    assert np.all(np.diff(key) > 0) and np.all(E[:, 0] < E[:, 1])


@pytest.mark.parametrize("text", ["f 1 2\n", "f 1 2 3 4\n", "f 1 x 3\n", "f 1 2 9\n"])
def test_bad_obj_rejected(tmp_path, text):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\n" + text)
    with pytest.raises(stq_amd.StqError):
        stq_amd.read_mesh(p)


def test_missing_and_unknown_files(tmp_path):
    with pytest.raises(stq_amd.StqError):
        stq_amd.read_mesh(tmp_path / "none.obj")
    (tmp_path / "m.stl").write_text("solid\n")
    with pytest.raises(stq_amd.StqError):
        stq_amd.read_mesh(tmp_path / "m.stl")


def test_recall():
    with open(os.path.join(GOLD, "recall_expected.json")) as fh:
        r = json.load(fh)
    truth = os.path.join(GOLD, "truth_tiny.json")
    assert list(stq_amd.recall(np.array(r["pairs"]), truth)) == r["as_given"]
    assert list(stq_amd.recall(np.array(r["pairs"]), truth, canonical=True)) == r["canonical"]
    assert stq_amd.recall(np.zeros((0, 2), np.int32), truth) == (0, 3)


def test_recall_rejects_bad_json(tmp_path):
    for text in ["[[1, 2], [3]]", "{\"a\": 1}", "[[1, 2]", ""]:
        p = tmp_path / "t.json"
        p.write_text(text)
        with pytest.raises(stq_amd.StqError):
            stq_amd.recall(np.zeros((1, 2), np.int32), p)
