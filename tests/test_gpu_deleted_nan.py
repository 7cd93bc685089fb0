"""Meshes with deleted faces, and query points that are not numbers.

* A MeshLib triangulation lists faceSize() entries with {-1,-1,-1} for deleted faces
  (source/MRMesh/MRMeshTopology.cpp:430-441); the tree covers the valid faces and keeps their ids
  (source/MRMesh/MRAABBTree.cpp:35-49).  Signed distance, projections, winding numbers and
  self-intersections equal the oracle's on such a mesh, with either tree builder.
* NaN query points mixed into a findProjections call on a deep tree (where nearly every wave goes
  through the split form): every record, the NaN points' empty ones included, equals the oracle's.
"""
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


@pytest.fixture
def restore_builder():
    old = M.get_tree_builder()
    yield
    M.set_tree_builder(old)


def _deleted(f, frac, seed):
    rng = np.random.default_rng(seed)
    fd = f.copy()
    gone = rng.random(f.shape[0]) < frac
    fd[gone] = -1
    return fd, gone


@pytest.mark.parametrize("builder", ["device", "host"])
def test_deleted_faces_match_oracle(builder, restore_builder):
    M.set_tree_builder(builder)
    v, f = synth.bumpy_icosphere(5, 8, 12)
    fd, gone = _deleted(f, 0.05, 4)
    # a second, overlapping copy makes some faces self-intersecting
    v2 = np.concatenate([v, v * np.float32(0.97) + np.float32([0.12, 0.0, 0.0])])
    f2 = np.concatenate([fd, np.where(fd >= 0, fd + v.shape[0], -1)]).astype(np.int32)
    say(f"{builder}: {int(gone.sum())} of {f.shape[0]} faces deleted")
    for verts, tris in ((v, fd), (v2, f2)):
        gm, om = M.Mesh(verts, tris), O.Mesh(verts, tris)
        nodes, dips = gm.tree()
        assert np.array_equal(nodes.view(np.uint8), om.nodes().view(np.uint8))
        assert np.array_equal(dips.view(np.uint8), om.dipoles().view(np.uint8))
        fwn = M.FastWindingNumber(gm)
        dims, xf, vox = synth.grid_for_mesh(verts, (40, 36, 44))
        g = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
        o = om.sdf_grid(dims, xf)
        assert np.array_equal(np.abs(g).view(np.uint32), np.abs(o).view(np.uint32))
        far = np.abs(o) > 1e-3 * float(vox.min())
        assert np.array_equal(np.signbit(g[far]), np.signbit(o[far]))
        pts = synth.projection_queries(verts, np.where(tris >= 0, tris, 0), 20000, 8)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(gm)
        gp, op = pr.findProjections(pts), om.find_projections(pts)
        assert np.array_equal(gp.view(np.uint8), op.view(np.uint8))
        assert not np.isin(gp["face"], np.flatnonzero(tris[:, 0] < 0)).any()
        assert np.max(np.abs(fwn.calcFromVector(pts, 2.0) - om.fwn_vector(pts))) <= 1e-4
        skip = int(np.flatnonzero(tris[:, 0] >= 0)[7])
        assert np.max(np.abs(fwn.calcFromVector(pts[:2000], 2.0, skip) - om.fwn_vector(pts[:2000], skip_face=skip))) <= 1e-4
        gs, osi = fwn.calcSelfIntersections(2.0), om.self_intersections()
        assert gs.shape[0] == tris.shape[0] and not gs[tris[:, 0] < 0].any()
        assert np.array_equal(gs, osi)
    assert om.self_intersections().any()


def test_all_faces_deleted(restore_builder):
    v, f = synth.bumpy_icosphere(2, 4, 1)
    fd = np.full_like(f, -1)
    gm = M.Mesh(v, fd)
    assert gm.tree()[0].size == 0
    fwn = M.FastWindingNumber(gm)
    assert not fwn.calcSelfIntersections(2.0).any()
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    r = pr.findProjections(v[:10])
    assert (r["face"] == -1).all()


def test_nan_points_deep_tree():
    # k=8: 21 tree levels, projection waves hand over after 64 steps (split form on nearly every wave)
    v, f = synth.bumpy_icosphere(8, 32, 2)
    gm, om = M.Mesh(v, f), O.Mesh(v, f)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    n = 1 << 15
    pts = synth.projection_queries(v, f, n, 77)
    rng = np.random.default_rng(5)
    bad = rng.choice(n, 40, replace=False)
    pts[bad[:20]] = np.nan                       # whole point
    pts[bad[20:30], 1] = np.nan                  # one coordinate
    pts[bad[30:], 2] = np.inf
    # points that sort next to the NaN ones (Morton key 0: the bbox corner) share their waves
    corner = v.min(0) - np.float32(0.3)
    pts[rng.choice(n, 200, replace=False)] = corner + rng.normal(0, 0.01, (200, 3)).astype(np.float32)
    say(f"{n} points, {bad.size} non-finite, depth {gm.depth()}")
    got = pr.findProjections(pts)
    fin = np.isfinite(pts).all(1)
    ok = np.flatnonzero(fin)
    ref_ok = om.find_projections(pts[ok])
    assert np.array_equal(got[ok].view(np.uint8), ref_ok.view(np.uint8))
    # a non-finite query finds nothing below the limit: the reference's empty record
    assert (got["face"][~fin] == -1).all()
    ref_bad = om.find_projections(pts[~fin][:12])
    assert np.array_equal(got[~fin][:12].view(np.uint8), ref_bad.view(np.uint8))
