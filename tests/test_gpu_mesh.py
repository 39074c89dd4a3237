"""Swept boxes of a moving mesh built on the GPU, against the oracle's
restatement of the reference construction (bit for bit, both builds), and
the fused mesh-to-pairs call against the pair-set oracle."""
import json

import numpy as np
import pytest

import stq_amd
from stq_amd import _lib, scenes
from oracle import oracle

pytestmark = pytest.mark.gpu

SPECIAL = np.array([0.0, -0.0, 1e-45, -1e-45, 1e-40, 0.1, -0.1, 1.0, 3.0e38, -3.0e38, 3.5e38, -3.5e38,
                    1e300, -1e300, 16777217.0, 0.333333333333], dtype=np.float64)


@pytest.fixture(scope="module")
def ctx():
    c = stq_amd.Context(0)
    yield c
    c.close()


def _mesh(seed, nv=3000):
    rng = np.random.default_rng(seed)
    V0 = rng.uniform(-2, 2, (nv, 3))
    V0.ravel()[:len(SPECIAL)] = SPECIAL
    V1 = V0 + rng.normal(0, 1e-2, V0.shape)
    V1.ravel()[len(SPECIAL):2 * len(SPECIAL)] = SPECIAL[::-1]
    E = rng.integers(0, nv, (5000, 2))
    F = rng.integers(0, nv, (4000, 3))
    return V0, V1, E, F


def _same_boxes(a, b):
    for k in ("id", "vids", "ref_id"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("min", "max"):  # bit patterns (NaN-free here; -0.0 vs +0.0 distinguished)
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


@pytest.mark.parametrize("float_build", [False, True])
@pytest.mark.parametrize("radius", [0.0, 1e-3, 1e-45])
def test_build_matches_oracle(ctx, float_build, radius):
    V0, V1, E, F = _mesh(4)
    got = ctx.build_swept_boxes(V0, V1, E, F, radius=radius, float_build=float_build)
    st = ctx.stats()  # the call's record count and its kernel time on the stream
    assert st["n_boxes"] == len(V0) + len(E) + len(F) and st["ms_total"] > 0
    ref = oracle.build_swept_boxes(V0, V1, E, F, radius=radius, float_build=float_build)
    _same_boxes(got, ref)


def test_build_rejects_bad_indices(ctx):
    V0, V1, E, F = _mesh(5, nv=100)
    E, F = E % 100, F % 100
    E[7, 0] = 100
    with pytest.raises(stq_amd.StqError) as e:
        ctx.build_swept_boxes(V0, V1, E, F)
    assert e.value.status == _lib.STQ_ERR_INVALID_ARGUMENT
    E[7, 0] = 0
    F[9, 2] = -3
    with pytest.raises(stq_amd.StqError):
        ctx.find_pairs_mesh(V0, V1, E, F)
    # the context stays usable
    F[9, 2] = 0
    assert len(ctx.find_pairs_mesh(V0, V1, E, F)) >= 0


@pytest.mark.parametrize("prec", [stq_amd.PRECISION_F64, stq_amd.PRECISION_F32_CONFIRM])
def test_mesh_pairs(ctx, prec):
    """Cloth mesh 150 x 150 with an inflation radius: the fused call equals the
    pair-set oracle on the oracle-built boxes."""
    V0, E, F = scenes.cloth_mesh(150)
    V1 = V0 + np.random.default_rng(8).normal(0, 4e-3, V0.shape)
    ctx.set_options(precision=prec)
    got = ctx.find_pairs_mesh(V0, V1, E, F, radius=1e-3)
    ref = oracle.reference_set(oracle.build_swept_boxes(V0, V1, E, F, radius=1e-3))
    assert len(ref) > 10000
    c = oracle.canonical(got)
    assert len(c) == len(got) and np.array_equal(c, ref)


def test_mesh_pairs_full_cloth(ctx):
    """The 708 x 708 cloth from its mesh equals the box path on the same boxes,
    pair for pair and in order."""
    V0, E, F = scenes.cloth_mesh(708)
    V1 = V0 + np.random.default_rng(1).normal(0.0, 2e-3, size=V0.shape)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    got = ctx.find_pairs_mesh(V0, V1, E, F)
    boxes = scenes.cloth_scene()
    assert np.array_equal(got, ctx.find_pairs(boxes))


def _mesh_pairs(ctx, g=80, seed=11, sigma=5e-3):
    V0, E, F = scenes.cloth_mesh(g)
    V1 = V0 + np.random.default_rng(seed).normal(0, sigma, V0.shape)
    boxes = oracle.build_swept_boxes(V0, V1, E, F)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    pairs = ctx.find_pairs_mesh(V0, V1, E, F)
    return V0, V1, boxes, pairs


def test_orient_pairs_matches_oracle(ctx):
    V0, V1, boxes, pairs = _mesh_pairs(ctx)
    assert len(pairs) > 1000
    got = ctx.orient_pairs(boxes, pairs)
    assert np.array_equal(got, oracle.orient_pairs(boxes, pairs))
    # orienting twice changes nothing; a vertex-vertex pair is rejected
    assert np.array_equal(ctx.orient_pairs(boxes, got), got)
    with pytest.raises(stq_amd.StqError):
        ctx.orient_pairs(boxes, np.array([[0, 1]], np.int32))


def test_ccd_queries(ctx):
    """Every record holds the pair's four vertices (vertex-face: the vertex,
    then the face's; edge-edge: first edge, then second) at t0 and t1."""
    V0, V1, boxes, pairs = _mesh_pairs(ctx, seed=12)
    q = ctx.ccd_queries(boxes, V0, V1, pairs, min_separation=1e-4)
    st = ctx.stats()
    assert st["n_boxes"] == len(pairs) and st["ms_total"] > 0
    o = oracle.orient_pairs(boxes, pairs)
    va, vb = boxes["vids"][o[:, 0]], boxes["vids"][o[:, 1]]
    vf = vb[:, 2] >= 0  # second box is a face: vertex-face
    verts = np.where(vf[:, None], np.stack([va[:, 0], vb[:, 0], vb[:, 1], vb[:, 2]], 1),
                     np.stack([va[:, 0], va[:, 1], vb[:, 0], vb[:, 1]], 1))
    assert np.array_equal(q["vs"], V0[verts]) and np.array_equal(q["ve"], V1[verts])
    assert np.all(q["ms"] == 1e-4)
    assert not q["err"].any() and not q["tol"].any() and not q["nbr_checks"].any()
    assert vf.any() and (~vf).any()
    # a vertex id outside the vertex arrays is reported
    with pytest.raises(stq_amd.StqError):
        ctx.ccd_queries(boxes, V0[:10], V1[:10], pairs)


def test_mesh_files_end_to_end(ctx, tmp_path):
    """t0/t1 meshes written as OBJ and PLY, read back, edges extracted, pairs
    found on the GPU, recall against the oracle's pair list as JSON: all found."""
    V0, _, F = scenes.cloth_mesh(60)
    V1 = V0 + np.random.default_rng(13).normal(0, 4e-3, V0.shape)
    with open(tmp_path / "t0.obj", "w") as fh:
        for v in V0:
            fh.write("v %r %r %r\n" % tuple(float(x) for x in v))
        for f in F:
            fh.write("f %d %d %d\n" % tuple(f + 1))
    with open(tmp_path / "t1.ply", "wb") as fh:
        fh.write(b"ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % len(V1) +
                 b"property double x\nproperty double y\nproperty double z\n" +
                 b"element face %d\nproperty list uchar int vertex_indices\nend_header\n" % len(F))
        fh.write(np.ascontiguousarray(V1, dtype="<f8").tobytes())
        rec = np.zeros(len(F), dtype=[("n", "u1"), ("i", "<i4", (3,))])
        rec["n"], rec["i"] = 3, F
        fh.write(rec.tobytes())
    A0, G0 = stq_amd.read_mesh(tmp_path / "t0.obj")
    A1, G1 = stq_amd.read_mesh(tmp_path / "t1.ply")
    assert np.array_equal(A0, V0) and np.array_equal(A1, V1) and np.array_equal(G0, G1)
    E = stq_amd.mesh_edges(G0)
    ctx.set_options(precision=stq_amd.PRECISION_F32_CONFIRM)
    pairs = ctx.find_pairs_mesh(A0, A1, E, G0)
    ref = oracle.reference_set(oracle.build_swept_boxes(A0, A1, E, G0))
    (tmp_path / "truth.json").write_text(json.dumps(ref.tolist()))
    assert stq_amd.recall(pairs, tmp_path / "truth.json") == (len(ref), len(ref))
    assert len(pairs) == len(ref)
