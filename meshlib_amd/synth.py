"""Seeded synthetic inputs for the distance-field path.

* ``bumpy_icosphere(level, bumps, seed)`` -- closed, outward-oriented, star-shaped
  triangle mesh: an icosahedron split ``level`` times (each triangle into four,
  midpoints pushed back to the unit sphere), then every vertex scaled by
  ``r = 1 + sum_j a_j * exp(-theta_j^2 / w_j^2)`` with ``a_j ~ U(0.05, 0.2)``,
  ``w_j ~ U(0.1, 0.4)`` and bump centres ``c_j`` uniform on the sphere.
  ``F = 20 * 4**level`` triangles, ``V = 10 * 4**level + 2`` vertices.
* ``grid_for_mesh(verts, dims, expand)`` -- the voxel-centre grid used by
  ``meshToDistanceVolume`` (reference ``source/MRVoxels/MRMeshToDistanceVolume.cpp:55``):
  ``p(x, y, z) = A * (x, y, z) + b`` with ``A = diag(voxelSize)`` and
  ``b = origin + 0.5 * voxelSize``.
* ``projection_queries(verts, tris, n, seed)`` -- half uniform in the bbox grown
  by 20 %, half a surface point plus N(0, (0.02 * bbox diagonal)^2) noise.

All randomness comes from ``numpy.random.Generator(PCG64(seed))``; geometry is
computed in float64 and rounded once to float32.
"""
from __future__ import annotations

import numpy as np

_PHI = (1.0 + 5.0 ** 0.5) / 2.0


def _icosahedron():
    v = np.array([
        [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
        [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
        [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
    ], dtype=np.float64)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # make every face counter-clockwise seen from outside (outward normal)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, a + b + c) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def icosphere(level: int):
    """Unit icosphere split ``level`` times; returns float64 verts, int64 faces."""
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * (nv + 1) + hi
        uniq, inv = np.unique(key, return_inverse=True)
        ulo = uniq // (nv + 1)
        uhi = uniq % (nv + 1)
        mid = v[ulo] + v[uhi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], axis=0)
        m = (inv + nv).reshape(3, -1)  # m[0]=mid(a,b), m[1]=mid(b,c), m[2]=mid(c,a)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([
            np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], axis=0)
    return v, f


def bumpy_icosphere(level: int, bumps: int, seed: int):
    """Star-shaped closed mesh; returns (verts float32 [V,3], tris int32 [F,3])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    amp = rng.uniform(0.05, 0.2, size=bumps)
    wid = rng.uniform(0.1, 0.4, size=bumps)
    cen = rng.standard_normal(size=(bumps, 3))
    cen /= np.linalg.norm(cen, axis=1, keepdims=True)
    v, f = icosphere(level)
    r = np.ones(v.shape[0], dtype=np.float64)
    for j in range(bumps):
        th = np.arccos(np.clip(v @ cen[j], -1.0, 1.0))
        r += amp[j] * np.exp(-(th * th) / (wid[j] * wid[j]))
    verts = (v * r[:, None]).astype(np.float32)
    return np.ascontiguousarray(verts), np.ascontiguousarray(f.astype(np.int32))


def mesh_bbox(verts: np.ndarray):
    lo = verts.min(axis=0).astype(np.float64)
    hi = verts.max(axis=0).astype(np.float64)
    return lo, hi


def grid_for_mesh(verts: np.ndarray, dims, expand: float = 1.1):
    """Voxel-centre grid over the mesh bbox with half-extent scaled by ``expand``.

    Returns ``(dims int32[3], xf float32[12])`` where ``xf`` is the row-major
    3x3 matrix ``A`` followed by the translation ``b`` (AffineXf3f layout).
    """
    dims = np.asarray(dims, dtype=np.int32)
    lo, hi = mesh_bbox(verts)
    c = 0.5 * (lo + hi)
    half = 0.5 * (hi - lo) * expand
    origin = (c - half).astype(np.float32)
    vox = ((2.0 * half) / dims).astype(np.float32)
    b = origin + np.float32(0.5) * vox  # float32 arithmetic, as the reference basis
    xf = np.zeros(12, dtype=np.float32)
    xf[0], xf[4], xf[8] = vox[0], vox[1], vox[2]
    xf[9:12] = b
    return dims, xf, vox


def projection_queries(verts: np.ndarray, tris: np.ndarray, n: int, seed: int,
                       sigma_frac: float = 0.02, grow: float = 1.2):
    """``n`` float32 query points: half uniform in the grown bbox, half near the surface."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = mesh_bbox(verts)
    c = 0.5 * (lo + hi)
    half = 0.5 * (hi - lo) * grow
    n_u = n // 2
    n_s = n - n_u
    pu = rng.uniform(c - half, c + half, size=(n_u, 3))
    v = verts.astype(np.float64)
    a, b, cc = v[tris[:, 0]], v[tris[:, 1]], v[tris[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, cc - a), axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    fi = np.searchsorted(cdf, rng.random(n_s), side="right")
    fi = np.minimum(fi, tris.shape[0] - 1)
    u = rng.random(n_s)
    w = rng.random(n_s)
    su = np.sqrt(u)
    ps = ((1.0 - su)[:, None] * a[fi] + (su * (1.0 - w))[:, None] * b[fi]
          + (su * w)[:, None] * cc[fi])
    sigma = sigma_frac * float(np.linalg.norm(hi - lo))
    ps += rng.normal(0.0, sigma, size=(n_s, 3))
    pts = np.concatenate([pu, ps], axis=0).astype(np.float32)
    return np.ascontiguousarray(pts)


def batched_workload_specs(items: int = 64, seed: int = 4):
    """The uneven batched workload: per item (split level, bumps, mesh seed, grid side, query count,
    query seed) -- levels uniform in 4..8, grid sides from {32, 64, 96, 128, 192}, 2^18..2^21 queries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    levels = rng.integers(4, 9, size=items)
    sides = rng.choice([32, 64, 96, 128, 192], size=items)
    nqs = rng.integers(1 << 18, (1 << 21) + 1, size=items)
    return [(int(levels[i]), 32, 1000 + i, int(sides[i]), int(nqs[i]), 2000 + i) for i in range(items)]
