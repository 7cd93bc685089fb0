"""The walks drop a queued leaf test when a lower bound of the triangle's distance (its plane slab
and edge half-spaces, with rounding margins) already exceeds the lane's best.  The bound must never
exceed the leaf test's own result: these cases put queries on and right next to the surface
(vertices, edge midpoints, face points pushed off by 1e-7..1e-2 along the normal), on meshes far
from the origin and with degenerate triangles (zero area, collinear, needles, slivers), and compare
projection records and distance magnitudes bit for bit with the oracle."""
import sys

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(offset, scale, degenerate, seed):
    v, f = synth.bumpy_icosphere(3, 8, seed)
    v = (v.astype(np.float64) * scale + offset).astype(np.float32)
    if degenerate:
        rng = np.random.default_rng(seed)
        nv = v.shape[0]
        extra_v, extra_f = [], []
        for i in range(40):
            a = v[rng.integers(nv)].astype(np.float64)
            b = v[rng.integers(nv)].astype(np.float64)
            kind = i % 4
            if kind == 0:
                c = a                                          # zero area (repeated vertex)
            elif kind == 1:
                c = a + 0.37 * (b - a)                         # collinear
            elif kind == 2:
                b = a + scale * 1e-6 * rng.normal(size=3)      # needle
                c = a + scale * 0.05 * rng.normal(size=3)
            else:
                c = 0.5 * (a + b) + scale * 1e-6 * rng.normal(size=3)   # sliver
            base = nv + len(extra_v)
            extra_v += [a, b, c]
            extra_f.append([base, base + 1, base + 2])
        v = np.concatenate([v, np.asarray(extra_v, np.float64).astype(np.float32)])
        f = np.concatenate([f, np.asarray(extra_f, np.int32)])
    return v, f


def _near_surface(v, f, n, seed):
    rng = np.random.default_rng(seed)
    vv = v.astype(np.float64)
    a, b, c = vv[f[:, 0]], vv[f[:, 1]], vv[f[:, 2]]
    nrm = np.cross(b - a, c - a)
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.where(ln > 0, ln, 1), 0)
    ext = np.linalg.norm(vv.max(0) - vv.min(0))
    fi = rng.integers(0, f.shape[0], n)
    u, w = rng.random(n), rng.random(n)
    flip = u + w > 1
    u[flip], w[flip] = 1 - u[flip], 1 - w[flip]
    on = a[fi] + u[:, None] * (b[fi] - a[fi]) + w[:, None] * (c[fi] - a[fi])
    h = ext * 10.0 ** rng.uniform(-7, -2, n) * rng.choice([-1, 1], n)
    parts = [on + h[:, None] * nrm[fi], on, vv[f[:2000, 0]], 0.5 * (vv[f[:2000, 0]] + vv[f[:2000, 1]])]
    return np.concatenate(parts).astype(np.float32)


@pytest.mark.parametrize("offset,scale,degenerate", [(0.0, 1.0, False), (0.0, 1.0, True),
                                                     (1000.0, 1.0, True), (-3.0e4, 25.0, False),
                                                     (0.5, 1e-3, True)])
def test_filter_keeps_every_winner(offset, scale, degenerate):
    v, f = _mesh(offset, scale, degenerate, 7)
    say(f"leaf filter: offset {offset} scale {scale} degenerate {degenerate}: {f.shape[0]} faces")
    gm = M.Mesh(v, f)
    om = O.Mesh(v, f)
    pts = _near_surface(v, f, 30000, 3)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    g = pr.findProjections(pts)
    o = om.find_projections(pts)
    bad = np.nonzero(np.any(g.view(np.uint8).reshape(-1, 32) != o.view(np.uint8).reshape(-1, 32), axis=1))[0]
    assert bad.size == 0, (bad[:5], g[bad[:3]], o[bad[:3]])
    dims, xf, vox = synth.grid_for_mesh(v, (40, 44, 36), expand=1.02)
    fwn = M.FastWindingNumber(gm)
    os_ = om.sdf_grid(dims, xf)
    try:
        for mode in ("on", "off", "auto"):      # auto: by the grid's size against the tree's depth
            M.set_leaf_filter(mode)
            s = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
            assert np.array_equal(np.abs(s).view(np.uint32), np.abs(os_).view(np.uint32)), mode
    finally:
        M.set_leaf_filter("auto")
