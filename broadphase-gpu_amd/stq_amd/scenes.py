"""Synthetic box scenes for the STQ broad phase.

Every scene is a numpy structured array with the reference FP64 box layout
(``stq::gpu::Aabb`` compiled with CCD_USE_DOUBLE, src/aabb.cuh:15-43): 72 bytes
per box, ``id`` = input position, vertex ids in the reference padding
convention (src/aabb.cu:80,113; src/cpu/aabb.cpp:34,63-64):

* vertex ``{v, -v-1, -v-1}``
* edge   ``{a, b, -a-1}``
* face   ``{a, b, c}``

Generators (seeded numpy PCG64, so a seed fixes the scene on any machine with
this numpy):

* ``uniform_scene``  -- 100,000-box uniform scene: centres U[0,1)^3, per-axis
  half-extent 10^U(-4,-2), 1-3 vertex ids from a pool of 60,000.
* ``cloth_scene``    -- swept boxes of a g x g cloth over [0,1]^2 with
  z = 0.05 sin(6x) sin(6y), one diagonal per quad, every vertex displaced by
  N(0, sigma^2) per axis over the step; one box per vertex, edge and face
  (hull of the t0/t1 positions), ordered vertices, edges, faces like
  ``constructBoxes`` (src/io.cu:25-27).  g = 708 gives N = 3,001,923.
* ``clustered_scene`` -- 8M boxes around 64 Gaussian clusters (sigma 0.02),
  half-extents 5e-5 (1 + Pareto(1.5)) capped at 0.05 per axis.
* ``batched_scenes`` -- 512 scenes, sizes log-uniform in [1e3, 5e5], each a
  uniform-like or cloth-like scene with a density factor rho ~ U[0.5, 4].
* ``learned_stream`` -- scenes of 5e4..4e6 boxes with rho log-uniform in
  [0.25, 8] (the learned-sizing stream).

Density factor (the batched and learned streams): a uniform-like scene of n
boxes uses half-extents scaled by (rho * 1e5 / n)^(1/3), which keeps the mean
number of neighbours per box at rho times that of the 100,000-box uniform
scene; a cloth-like scene of about n boxes uses a grid of
g = round(sqrt(n / 6)) and a displacement sigma of 2e-3 * (707 / (g - 1))
* rho^(1/3), which keeps the motion relative to the grid spacing.
"""
from __future__ import annotations

import numpy as np

AABB_DTYPE = np.dtype(
    [
        ("id", "<i4"),
        ("pad0", "<i4"),
        ("min", "<f8", (3,)),
        ("max", "<f8", (3,)),
        ("vids", "<i4", (3,)),
        ("ref_id", "<i4"),
    ],
    align=False,
)
assert AABB_DTYPE.itemsize == 72

UNIFORM_N = 100_000
VID_POOL = 60_000


def empty_boxes(n: int) -> np.ndarray:
    b = np.zeros(n, dtype=AABB_DTYPE)
    b["id"] = np.arange(n, dtype=np.int32)
    b["ref_id"] = b["id"]
    return b


def _vids_random(rng: np.random.Generator, n: int, pool: int) -> np.ndarray:
    """1-3 ids per box from U{0..pool-1}, padded by the reference convention."""
    k = rng.integers(1, 4, size=n)
    ids = rng.integers(0, pool, size=(n, 3), dtype=np.int64)
    v = np.empty((n, 3), dtype=np.int64)
    v[:, 0] = ids[:, 0]
    # vertex {v,-v-1,-v-1}; edge {a,b,-a-1}; face {a,b,c}
    v[:, 1] = np.where(k >= 2, ids[:, 1], -ids[:, 0] - 1)
    v[:, 2] = np.where(k == 3, ids[:, 2], -ids[:, 0] - 1)
    return v.astype(np.int32)


def uniform_scene(n: int = UNIFORM_N, seed: int = 0, extent_scale: float = 1.0,
                  vid_pool: int = VID_POOL) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3))
    h = 10.0 ** rng.uniform(-4.0, -2.0, size=(n, 3)) * extent_scale
    b = empty_boxes(n)
    b["min"] = c - h
    b["max"] = c + h
    b["vids"] = _vids_random(rng, n, vid_pool)
    return b


def cloth_mesh(g: int):
    """Vertex grid, edges and faces of a g x g cloth (one diagonal per quad)."""
    lin = np.linspace(0.0, 1.0, g)
    X, Y = np.meshgrid(lin, lin, indexing="ij")
    V = np.stack([X.ravel(), Y.ravel(),
                  0.05 * np.sin(6.0 * X.ravel()) * np.sin(6.0 * Y.ravel())], axis=1)
    idx = np.arange(g * g, dtype=np.int64).reshape(g, g)
    eh = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    ev = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    ed = np.stack([idx[:-1, :-1].ravel(), idx[1:, 1:].ravel()], axis=1)
    E = np.concatenate([eh, ev, ed], axis=0)
    f1 = np.stack([idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()], axis=1)
    f2 = np.stack([idx[:-1, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()], axis=1)
    F = np.concatenate([f1, f2], axis=0)
    return V, E, F


def swept_boxes(V0: np.ndarray, V1: np.ndarray, E: np.ndarray, F: np.ndarray) -> np.ndarray:
    """FP64 swept boxes: hull of the t0/t1 positions of each simplex's vertices
    (the CCD_USE_DOUBLE branch of src/aabb.cu:63-174, without inflation).
    ref_id is the index within the box's kind, as the reference records it
    (src/aabb.cu:94,128,170)."""
    nv, ne, nf = len(V0), len(E), len(F)
    b = empty_boxes(nv + ne + nf)
    b["ref_id"] = np.concatenate([np.arange(nv), np.arange(ne), np.arange(nf)]).astype(np.int32)
    vmin = np.minimum(V0, V1)
    vmax = np.maximum(V0, V1)
    b["min"][:nv] = vmin
    b["max"][:nv] = vmax
    vi = np.arange(nv, dtype=np.int64)
    b["vids"][:nv] = np.stack([vi, -vi - 1, -vi - 1], axis=1)
    s = nv
    b["min"][s:s + ne] = np.minimum(vmin[E[:, 0]], vmin[E[:, 1]])
    b["max"][s:s + ne] = np.maximum(vmax[E[:, 0]], vmax[E[:, 1]])
    b["vids"][s:s + ne] = np.stack([E[:, 0], E[:, 1], -E[:, 0] - 1], axis=1)
    s += ne
    b["min"][s:] = np.minimum(np.minimum(vmin[F[:, 0]], vmin[F[:, 1]]), vmin[F[:, 2]])
    b["max"][s:] = np.maximum(np.maximum(vmax[F[:, 0]], vmax[F[:, 1]]), vmax[F[:, 2]])
    b["vids"][s:] = F
    return b


def cloth_scene(g: int = 708, seed: int = 1, sigma: float = 2e-3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    V0, E, F = cloth_mesh(g)
    V1 = V0 + rng.normal(0.0, sigma, size=V0.shape)
    return swept_boxes(V0, V1, E, F)


def clustered_scene(n: int = 8_000_000, seed: int = 2, clusters: int = 64,
                    sigma: float = 0.02, hmin: float = 5e-5, hcap: float = 0.05,
                    alpha: float = 1.5, vid_pool: int = VID_POOL) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.random((clusters, 3))
    which = rng.integers(0, clusters, size=n)
    c = centres[which] + rng.normal(0.0, sigma, size=(n, 3))
    h = np.minimum(hmin * (1.0 + rng.pareto(alpha, size=(n, 3))), hcap)
    b = empty_boxes(n)
    b["min"] = c - h
    b["max"] = c + h
    b["vids"] = _vids_random(rng, n, vid_pool)
    return b


def density_scene(kind: int, n: int, rho: float, seed: int) -> np.ndarray:
    """One scene of a batched/learned stream: kind 0 = uniform-like, 1 = cloth-like."""
    if kind == 0:
        scale = (rho * UNIFORM_N / max(n, 1)) ** (1.0 / 3.0)
        return uniform_scene(n, seed=seed, extent_scale=scale)
    g = max(2, int(round(np.sqrt(n / 6.0))))
    sigma = 2e-3 * (707.0 / (g - 1)) * rho ** (1.0 / 3.0)
    return cloth_scene(g, seed=seed, sigma=sigma)


def batched_scenes(count: int = 512, seed: int = 3, nmin: float = 1e3,
                   nmax: float = 5e5, rho_lo: float = 0.5, rho_hi: float = 4.0):
    """List of (kind, rho, boxes) for the batched stream."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(10.0 ** rng.uniform(np.log10(nmin), np.log10(nmax), size=count)).astype(np.int64)
    kinds = rng.integers(0, 2, size=count)
    rhos = rng.uniform(rho_lo, rho_hi, size=count)
    out = []
    for i in range(count):
        out.append((int(kinds[i]), float(rhos[i]),
                    density_scene(int(kinds[i]), int(sizes[i]), float(rhos[i]), seed * 100003 + i)))
    return out


def learned_stream_params(count: int = 2000, seed: int = 4, nmin: float = 5e4,
                          nmax: float = 4e6, rho_lo: float = 0.25, rho_hi: float = 8.0):
    """(kind, n, rho, seed) tuples of the learned-sizing stream."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(10.0 ** rng.uniform(np.log10(nmin), np.log10(nmax), size=count)).astype(np.int64)
    kinds = rng.integers(0, 2, size=count)
    rhos = 10.0 ** rng.uniform(np.log10(rho_lo), np.log10(rho_hi), size=count)
    return [(int(kinds[i]), int(sizes[i]), float(rhos[i]), seed * 100003 + i) for i in range(count)]


def concat_scenes(scenes):
    """Concatenate scenes into one box array plus int64 offsets (nscenes+1).
    Ids inside each scene stay scene-local positions."""
    sizes = [len(s) for s in scenes]
    offs = np.zeros(len(scenes) + 1, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)
    allb = np.concatenate(scenes) if scenes else empty_boxes(0)
    return allb, offs
