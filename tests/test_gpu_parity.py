"""GPU parity of the HIP path against the CPU oracle, through the C-ABI.

Bar: projection records bit-identical (face ids, points, barycentrics,
distSq -- ties included, since the per-lane visit order is the reference's); signed-distance
magnitudes bit-identical; winding numbers within 1e-4 absolute (they differ only by atan2f
ulps); inside/outside sign identical on every voxel farther than 1e-3 voxel from the surface.
"""
import sys

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FWN_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def k5():
    v, f = synth.bumpy_icosphere(5, 8, 1)
    return v, f, M.Mesh(v, f), O.Mesh(v, f)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def test_projection_bit_identical_k5(k5):
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (64, 64, 64))
    pts = np.concatenate([O.grid_points(dims, xf), synth.projection_queries(v, f, 65536, 11)])
    say("projection k=5:", pts.shape[0], "queries")
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    g = pr.findProjections(pts)
    o = om.find_projections(pts)
    mism = np.nonzero(np.any(_bits(g).reshape(-1, 32) != _bits(o).reshape(-1, 32), axis=1))[0]
    assert mism.size == 0, (mism[:10], g[mism[:3]], o[mism[:3]])


def test_projection_limits_and_xf(k5):
    v, f, gm, om = k5
    pts = synth.projection_queries(v, f, 20000, 12)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    for up, lo in ((0.01, 0.0), (M.FLT_MAX, 0.02), (0.05, 0.001)):
        say("limits", up, lo)
        g = pr.findProjections(pts, upDistLimitSq=up, loDistLimitSq=lo)
        o = om.find_projections(pts, up=up, lo=lo)
        assert np.array_equal(_bits(g), _bits(o))
    # point transform, and a non-rigid reference transform applied to boxes and triangles
    objXf = np.array([0.9, 0.1, 0, -0.1, 0.9, 0, 0, 0, 1.0, 0.05, -0.02, 0.01], np.float32)
    refXf = np.array([1.3, 0, 0.1, 0, 0.8, 0, 0.05, 0, 1.1, 0.1, 0.0, -0.1], np.float32)
    say("xf")
    g = pr.findProjections(pts, objXf=objXf)
    o = om.find_projections(pts, pt_xf=objXf)
    assert np.array_equal(_bits(g), _bits(o))
    g = pr.findProjections(pts, objXf=objXf, refObjXf=refXf)
    o = om.find_projections(pts, pt_xf=objXf, tree_xf=refXf)
    assert np.array_equal(_bits(g), _bits(o))
    # rigid reference transform is folded into the point transform: same distances
    c, s = np.cos(0.3), np.sin(0.3)
    rig = np.array([c, -s, 0, s, c, 0, 0, 0, 1, 0.2, 0.1, 0.0], np.float32)
    g = pr.findProjections(pts, refObjXf=rig)
    A = rig[:9].reshape(3, 3).astype(np.float64)
    local = (pts.astype(np.float64) - rig[9:]) @ A   # inverse of a rotation = transpose
    d64, _ = om.find_projections_f64(local.astype(np.float32))
    # the folded float transform moves points by ~1 ulp of their coordinates
    assert np.max(np.abs(np.sqrt(g["distSq"].astype(np.float64)) - np.sqrt(d64))) < 2e-6


def test_sdf_k5_64(k5):
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (64, 64, 64))
    say("sdf 64^3 k=5")
    fwn = M.FastWindingNumber(gm)
    g = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    o = om.sdf_grid(dims, xf)
    assert not np.isnan(g).any()
    assert np.array_equal(np.abs(g), np.abs(o))            # distances bit-identical
    far = np.abs(o) > 1e-3 * float(vox.min())
    assert np.array_equal(np.sign(g[far]), np.sign(o[far]))
    # nullOutsideMinMax band
    opts = M.DistanceToMeshOptions(minDistSq=0.001, maxDistSq=0.04)
    g2 = fwn.calcFromGridWithDistances(dims, xf, opts)
    o2 = om.sdf_grid(dims, xf, opts=O.default_opts(minDistSq=0.001, maxDistSq=0.04))
    assert np.array_equal(np.isnan(g2), np.isnan(o2))
    ok = ~np.isnan(o2)
    assert np.array_equal(np.abs(g2[ok]), np.abs(o2[ok]))


def test_fwn_grid_vector_selfint(k5):
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (48, 40, 36))
    fwn = M.FastWindingNumber(gm)
    say("fwn grid")
    g = fwn.calcFromGrid(dims, xf, 2.0)
    o = om.fwn_grid(dims, xf, beta=2.0)
    err = np.abs(g.astype(np.float64) - o)
    say("fwn grid max abs err", err.max())
    assert err.max() <= FWN_TOL
    pts = synth.projection_queries(v, f, 30000, 3)
    say("fwn vector")
    g = fwn.calcFromVector(pts, 2.0)
    o = om.fwn_vector(pts, beta=2.0)
    assert np.abs(g - o).max() <= FWN_TOL
    g = fwn.calcFromVector(pts, 2.0, skipFace=17)
    o = om.fwn_vector(pts, beta=2.0, skip_face=17)
    assert np.abs(g - o).max() <= FWN_TOL
    say("self intersections")
    assert np.array_equal(fwn.calcSelfIntersections(2.0), om.self_intersections(2.0))
    # a self-intersecting union of two copies
    v2 = np.concatenate([v, v + np.float32([0.6, 0, 0])])
    f2 = np.concatenate([f, f + v.shape[0]])
    g2 = M.FastWindingNumber(M.Mesh(v2, f2)).calcSelfIntersections(2.0)
    o2 = O.Mesh(v2, f2).self_intersections(2.0)
    assert g2.any() and np.array_equal(g2, o2)


@pytest.fixture
def per_lane_then_restore():
    prev = M.get_traversal()
    yield
    M.set_traversal(prev)


def test_traversal_forms_identical(k5, per_lane_then_restore):
    # the wave-cooperative kernels keep each lane's node set (winding number: also its order), so
    # every output is bit-identical to the per-lane (reference-form) kernels
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (50, 44, 38))
    pts = synth.projection_queries(v, f, 30000, 31)
    fwn = M.FastWindingNumber(gm)
    outs = {}
    for mode in ("per_lane", "wave_coop"):
        say("traversal", mode)
        M.set_traversal(mode)
        outs[mode] = (
            fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions()),
            fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions(maxDistSq=0.01,
                                                                            nullOutsideMinMax=False)),
            fwn.calcFromGrid(dims, xf, 2.0),
            fwn.calcFromVector(pts, 2.0, skipFace=5),
            fwn.calcSelfIntersections(2.0),
        )
    for a, b in zip(outs["per_lane"], outs["wave_coop"]):
        assert np.array_equal(_bits(a), _bits(b))
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    refXf = np.array([1.3, 0, 0.1, 0, 0.8, 0, 0.05, 0, 1.1, 0.1, 0.0, -0.1], np.float32)
    proj = {}
    for mode in ("per_lane", "wave_coop"):
        M.set_traversal(mode)
        proj[mode] = (pr.findProjections(pts), pr.findProjections(pts, upDistLimitSq=0.02),
                      pr.findProjections(pts, loDistLimitSq=0.01), pr.findProjections(pts, refObjXf=refXf))
    for a, b in zip(proj["per_lane"], proj["wave_coop"]):
        assert np.array_equal(_bits(a), _bits(b))


def test_projection_ties(k5):
    # points just outside the vertices (along the radial direction): the vertex is the closest
    # feature of every incident face, so several faces tie at exactly the same distSq and the
    # reference's depth-first order picks the winner
    v, f, gm, om = k5
    vv = v.astype(np.float64)
    n = vv / np.linalg.norm(vv, axis=1, keepdims=True)
    pts = np.concatenate([(vv + 0.003 * n), vv, 0.5 * (vv[f[:, 0]] + vv[f[:, 1]])]).astype(np.float32)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    o = om.find_projections(pts)
    for mode in ("wave_coop", "per_lane"):
        M.set_traversal(mode)
        g = pr.findProjections(pts)
        assert np.array_equal(_bits(g), _bits(o)), mode
    M.set_traversal("wave_coop")
    # the ties are real: many queries have another face at exactly the winning distance
    d64 = O.bf_nearest_f64(v, f, pts[:2000])
    assert np.mean(o["distSq"][:2000] > 0) > 0.9 and d64.shape[0] == 2000


def test_chunked_and_by_parts(k5):
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (32, 24, 40))
    fwn = M.FastWindingNumber(gm)
    full = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    pts = synth.projection_queries(v, f, 5000, 4)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    pfull = pr.findProjections(pts)
    wfull = fwn.calcFromVector(pts, 2.0)
    try:
        M.set_memory_limit(32 * 24 * 4 * 7)                 # several z-layers per chunk
        say("chunked grid")
        assert np.array_equal(_bits(fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())),
                              _bits(full))
        M.set_memory_limit(700 * 64)                        # ~700 points per chunk
        say("chunked points")
        assert np.array_equal(_bits(pr.findProjections(pts)), _bits(pfull))
        assert np.array_equal(_bits(fwn.calcFromVector(pts, 2.0)), _bits(wfull))
        M.set_memory_limit(32 * 24 * 10 * 2 * 3)            # 3 layers per buffer (10 B per voxel), 2 buffers
        for overlap in (0, 1, 2):
            parts = []
            say("by parts overlap", overlap)
            fwn.calcFromGridWithDistancesByParts(
                lambda data, d, z0: parts.append((z0, d, data)), dims, xf,
                M.DistanceToMeshOptions(), overlap)
            assert len(parts) > 2
            layer = 32 * 24
            for z0, d, data in parts:
                assert d[0] == 32 and d[1] == 24
                assert np.array_equal(_bits(data), _bits(full[z0 * layer:(z0 + d[2]) * layer]))
            covered = np.zeros(40, bool)
            for z0, d, _ in parts:
                covered[z0:z0 + d[2]] = True
            assert covered.all()
            if overlap:
                assert all(parts[i + 1][0] == parts[i][0] + parts[i][1][2] - overlap for i in range(len(parts) - 1))
        say("winding numbers by parts, overlap 1")
        wfull = fwn.calcFromGrid(dims, xf, 2.0)
        parts = []
        fwn.calcFromGridByParts(lambda data, d, z0: parts.append((z0, d, data)), dims, xf, 2.0, 1)
        assert len(parts) > 2
        for z0, d, data in parts:
            assert np.array_equal(_bits(data), _bits(wfull[z0 * layer:(z0 + d[2]) * layer]))
        assert parts[-1][0] + parts[-1][1][2] == 40
        # canceling through the progress callback
        with pytest.raises(M.MeshLibAmdError) as ei:
            fwn.calcFromGrid(dims, xf, 2.0, cb=lambda p: p < 0.7)
        assert ei.value.code == M.MI_E_CANCELED
    finally:
        M.set_memory_limit(0)


def test_part_pipelines_at_full_budget(k5):
    """Host-pointer and by-parts calls larger than one pipeline part (64 MiB of voxels, 2^21 points):
    several parts in flight on two streams equal the device entry points bit for bit."""
    import torch
    v, f, gm, om = k5
    fwn = M.FastWindingNumber(gm)
    opts = M.DistanceToMeshOptions()
    dims, xf, vox = synth.grid_for_mesh(v, (256, 256, 300))   # 256 + 44 layers
    nvox, layer = 256 * 256 * 300, 256 * 256
    dev = torch.empty(nvox, dtype=torch.float32, device="cuda")
    for name, host, fill in (
            ("sdf", lambda: fwn.calcFromGridWithDistances(dims, xf, opts),
             lambda: fwn.sdf_grid_device(dims, xf, opts, dev.data_ptr())),
            ("fwn", lambda: fwn.calcFromGrid(dims, xf, 2.0),
             lambda: fwn.fwn_grid_device(dims, xf, 2.0, dev.data_ptr()))):
        say("host-pointer grid", name)
        fill()
        torch.cuda.synchronize()
        assert np.array_equal(_bits(host()), _bits(dev.cpu().numpy())), name
    say("by parts, overlap 1")
    fwn.sdf_grid_device(dims, xf, opts, dev.data_ptr())
    full = dev.cpu().numpy()
    parts = []
    fwn.calcFromGridWithDistancesByParts(lambda data, d, z0: parts.append((z0, d, data)), dims, xf, opts, 1)
    assert len(parts) == 2 and parts[1][0] == parts[0][1][2] - 1
    for z0, d, data in parts:
        assert np.array_equal(_bits(data), _bits(full[z0 * layer:(z0 + d[2]) * layer]))
    assert parts[-1][0] + parts[-1][1][2] == 300
    say("winding numbers by parts, overlap 3")
    fwn.fwn_grid_device(dims, xf, 2.0, dev.data_ptr())
    wfull = dev.cpu().numpy()
    wparts = []
    fwn.calcFromGridByParts(lambda data, d, z0: wparts.append((z0, d, data)), dims, xf, 2.0, 3)
    assert len(wparts) == 2 and wparts[1][0] == wparts[0][1][2] - 3
    for z0, d, data in wparts:
        assert np.array_equal(_bits(data), _bits(wfull[z0 * layer:(z0 + d[2]) * layer]))
    assert wparts[-1][0] + wparts[-1][1][2] == 300

    n = (2 << 21) + 4321                                      # two full parts and a ragged one
    pts = synth.projection_queries(v, f, n, 12)
    dp = torch.from_numpy(pts).cuda()
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    say("host-pointer projections", n)
    dr = torch.empty(n * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(dp.data_ptr(), n, dr.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(_bits(pr.findProjections(pts)), _bits(dr.cpu().numpy()))
    say("host-pointer winding numbers", n)
    dw = torch.empty(n, dtype=torch.float32, device="cuda")
    fwn.fwn_vector_device(dp.data_ptr(), n, 2.0, dw.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(_bits(fwn.calcFromVector(pts, 2.0)), _bits(dw.cpu().numpy()))
    sel = np.random.default_rng(3).choice(n, 3000, replace=False)
    assert np.array_equal(_bits(dr.cpu().numpy().view(M.PROJ_DTYPE)[sel]), _bits(om.find_projections(pts[sel])))

    say("cancellation with parts in flight, then the same calls again")
    with pytest.raises(M.MeshLibAmdError) as ei:
        fwn.calcFromGridWithDistancesByParts(lambda data, d, z0: False, dims, xf, opts, 1)
    assert ei.value.code == M.MI_E_CANCELED
    with pytest.raises(M.MeshLibAmdError) as ei:
        fwn.calcFromVector(pts, 2.0, cb=lambda p: p < 0.7)
    assert ei.value.code == M.MI_E_CANCELED
    with pytest.raises(M.MeshLibAmdError) as ei:
        fwn.calcFromGrid(dims, xf, 2.0, cb=lambda p: p < 0.65)
    assert ei.value.code == M.MI_E_CANCELED
    parts.clear()
    fwn.calcFromGridWithDistancesByParts(lambda data, d, z0: parts.append((z0, d, data)), dims, xf, opts, 1)
    assert [p[0] for p in parts] == [0, parts[0][1][2] - 1]
    for z0, d, data in parts:
        assert np.array_equal(_bits(data), _bits(full[z0 * layer:(z0 + d[2]) * layer]))
    assert np.array_equal(_bits(fwn.calcFromVector(pts, 2.0)), _bits(dw.cpu().numpy()))


def test_device_entry_points_match_host(k5):
    import torch
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (40, 40, 40))
    fwn = M.FastWindingNumber(gm)
    host = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    out = torch.empty(40 * 40 * 40, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    fwn.sdf_grid_device(dims, xf, M.DistanceToMeshOptions(), out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(host))
    # slab of layers
    slab = torch.empty(40 * 40 * 7, dtype=torch.float32, device="cuda")
    fwn.sdf_grid_device(dims, xf, M.DistanceToMeshOptions(), slab.data_ptr(), zBegin=11, zEnd=18, stream=s)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(slab.cpu().numpy()), _bits(host[11 * 1600:18 * 1600]))
    pts = synth.projection_queries(v, f, 10000, 8)
    dp = torch.from_numpy(pts).cuda()
    dr = torch.empty(10000 * 8, dtype=torch.int32, device="cuda")
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    pr.find_projections_device(dp.data_ptr(), 10000, dr.data_ptr(), stream=s)
    torch.cuda.synchronize()
    g = dr.cpu().numpy().view(M.PROJ_DTYPE)
    assert np.array_equal(_bits(g), _bits(om.find_projections(pts)))


def test_degenerate_meshes():
    say("empty mesh")
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    em = M.Mesh(v, np.zeros((0, 3), np.int32))
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(em)
    r = pr.findProjections(np.ones((5, 3), np.float32), upDistLimitSq=4.0)
    assert np.all(r["face"] == -1) and np.all(r["distSq"] == 4.0)
    w = M.FastWindingNumber(em).calcFromVector(np.ones((5, 3), np.float32), 2.0)
    assert np.all(w == 0)
    say("single triangle")
    sm = M.Mesh(v, np.array([[0, 1, 2]], np.int32))
    om = O.Mesh(v, np.array([[0, 1, 2]], np.int32))
    pts = np.random.default_rng(0).normal(size=(1000, 3)).astype(np.float32)
    pr.updateMeshData(sm)
    assert np.array_equal(_bits(pr.findProjections(pts)), _bits(om.find_projections(pts)))
    g = M.FastWindingNumber(sm).calcFromVector(pts, 2.0)
    assert np.abs(g - om.fwn_vector(pts)).max() <= FWN_TOL


def test_k8_scale_subsample():
    # production mesh (k=8, 1.31M triangles): seeded subsamples against the oracle
    say("k=8 build")
    v, f = synth.bumpy_icosphere(8, 32, 2)
    gm, om = M.Mesh(v, f), O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 1 << 16, 21)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    say("k=8 projections")
    g = pr.findProjections(pts)
    sel = np.random.default_rng(1).choice(len(pts), 20000, replace=False)
    o = om.find_projections(pts[sel])
    assert np.array_equal(_bits(g[sel]), _bits(o))
    say("k=8 sdf 96^3 sampled")
    dims, xf, vox = synth.grid_for_mesh(v, (96, 96, 96))
    gs = M.FastWindingNumber(gm).calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    ids = np.random.default_rng(2).choice(gs.size, 20000, replace=False)
    os_ = om.sdf_grid_sample(dims, xf, ids)
    assert np.array_equal(np.abs(gs[ids]), np.abs(os_))
    far = np.abs(os_) > 1e-3 * float(vox.min())
    assert np.array_equal(np.sign(gs[ids][far]), np.sign(os_[far]))


def test_argument_and_budget_errors(k5):
    """Invalid arguments and budgets come back as status codes with a message (no launch)."""
    import ctypes
    import torch
    v, f, gm, om = k5
    fwn = M.FastWindingNumber(gm)
    opts = M.DistanceToMeshOptions()
    dims, xf, vox = synth.grid_for_mesh(v, (32, 24, 40))

    def code(fn):
        with pytest.raises(M.MeshLibAmdError) as ei:
            fn()
        assert str(ei.value)
        return ei.value.code
    assert code(lambda: fwn.calcFromGridWithDistancesByParts(lambda *a: None, dims, xf, opts, -1)) == M.MI_E_INVALID
    d = torch.empty(32 * 24 * 40, dtype=torch.float32, device="cuda")
    assert code(lambda: fwn.sdf_grid_device(dims, xf, opts, d.data_ptr(), zBegin=5, zEnd=41)) == M.MI_E_INVALID
    assert code(lambda: fwn.sdf_grid_device(dims, xf, opts, 0)) == M.MI_E_INVALID
    caps = torch.ones(32 * 24 * 40, dtype=torch.float32, device="cuda")
    band = M.DistanceToMeshOptions(minDistSq=0.001)
    assert code(lambda: fwn.sdf_grid_capped_device(dims, xf, band, caps.data_ptr(), d.data_ptr())) == M.MI_E_INVALID
    L = M.lib()
    xf12 = M._xf(xf)
    out = np.empty(32 * 24 * 40, np.float32)
    o = opts._c()
    neg = np.array([32, -1, 40], np.int32)
    no_cb = M._progress(None)[0]
    assert L.mi_fwn_calc_from_grid_with_distances(gm.handle, M._p(neg), M._p(xf12), ctypes.byref(o), M._p(out),
                                                  no_cb, None) == M.MI_E_INVALID
    good = np.array([32, 24, 40], np.int32)
    assert L.mi_fwn_calc_from_grid(gm.handle, M._p(good), M._p(xf12), 2.0, None, no_cb, None) == M.MI_E_INVALID
    try:
        M.set_memory_limit(32 * 24 * 4 * 2 * 3)       # 3 layers per by-parts buffer of winding numbers
        assert code(lambda: fwn.calcFromGridByParts(lambda *a: None, dims, xf, 2.0, 3)) == M.MI_E_LIMIT
        parts = []
        fwn.calcFromGridByParts(lambda data, dd, z0: parts.append(z0), dims, xf, 2.0, 2)
        assert parts[:3] == [0, 1, 2]                 # step = 3 layers - 2 overlapping
        M.set_memory_limit(100)                       # not one z-layer
        assert code(lambda: fwn.calcFromGrid(dims, xf, 2.0)) == M.MI_E_LIMIT
        assert code(lambda: M.PointsToMeshProjector().findProjections(v[:4])) == M.MI_E_INVALID
    finally:
        M.set_memory_limit(0)
    # the mesh still works after the errors
    g = fwn.calcFromGridWithDistances(dims, xf, opts)
    assert np.array_equal(np.abs(g).view(np.uint32), np.abs(om.sdf_grid(dims, xf)).view(np.uint32))
