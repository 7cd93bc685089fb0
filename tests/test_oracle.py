"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (dropped term, wrong sign or
index, swapped operand, wrong traversal order or pruning) fails one of them:
* the reference's own known-answer tests (source/MRTest/MRDipoleTests.cpp:25-51,
  source/MRTest/MRTriMathTests.cpp:168-182, source/MRTest/MRAABBTreeTests.cpp:10-27,
  source/MRTestC2/MRPointsToMeshProjector.c:10-36);
* closed forms (closest-point Voronoi regions, solid angle of an octant, winding number
  1 / 0 inside / outside a closed surface);
* brute force over every triangle in FP64 on small meshes;
* an independent re-implementation (numpy) of the closest point and of the tree invariants.
"""
import numpy as np
import pytest

from meshlib_amd import synth
from oracle import oracle as O

F32MAX = float(np.finfo(np.float32).max)


# ---------------------------------------------------------------- primitives

def test_solid_angle_kat():
    # source/MRTest/MRDipoleTests.cpp:25-51
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    c = (tri[0] + tri[1] + tri[2]) / np.float32(3)
    x = O.triangle_solid_angle(c + np.float32([0, 0, 1e-5]), *tri)
    y = O.triangle_solid_angle(c - np.float32([0, 0, 1e-5]), *tri)
    assert abs(x - (-2 * np.pi)) < 1e-3
    assert abs(y - (2 * np.pi)) < 1e-3
    for i in range(3):
        assert O.triangle_solid_angle(tri[i], *tri) == 0.0
    assert O.triangle_solid_angle(tri[1] + tri[2], *tri) == 0.0
    assert O.triangle_solid_angle(-tri[1], *tri) == 0.0
    assert O.triangle_solid_angle(-tri[2], *tri) == 0.0


def test_solid_angle_octant():
    # the octant triangle (e_x, e_y, e_z) seen from the origin spans 4*pi/8, with the sign
    # of det[a, b, c] = +1
    tri = np.eye(3, dtype=np.float32)
    w = O.triangle_solid_angle(np.zeros(3, np.float32), *tri)
    assert abs(w - 4 * np.pi / 8) < 1e-6
    w2 = O.triangle_solid_angle(np.zeros(3, np.float32), tri[0], tri[2], tri[1])
    assert abs(w2 + 4 * np.pi / 8) < 1e-6


def test_closest_point_kat():
    # source/MRTest/MRTriMathTests.cpp:168-182
    a = np.zeros(3)
    b = np.array([0, 0, 1.6487993830814958])
    c = np.array([0, 0, -0.0025229454040527344])
    x = np.array([0, 0, 0.29875588417053223])
    p = np.array([-0.17274856567382812, -0.092119634151458740, 0]) + x
    pr, ba = O.closest_point_in_triangle_d(p, a, b, c)
    assert np.array_equal(pr, x)
    assert ba[0] >= 0 and ba[1] >= 0 and ba[0] + ba[1] <= 1
    interp = a + ba[0] * (b - a) + ba[1] * (c - a)
    assert np.linalg.norm(interp - x) < 1e-15


@pytest.mark.parametrize("p,proj,bary", [
    ((-1, -1, 0.5), (0, 0, 0), (0, 0)),           # region #1: vertex a
    ((2, -0.5, 0.3), (1, 0, 0), (1, 0)),          # region #2: vertex b
    ((-0.5, 2, 0.1), (0, 1, 0), (0, 1)),          # region #3: vertex c
    ((0.5, -1, 0.2), (0.5, 0, 0), (0.5, 0)),      # region #4: edge ab
    ((-1, 0.25, 0), (0, 0.25, 0), (0, 0.25)),     # region #5: edge ac
    ((1, 1, 1), (0.5, 0.5, 0), (0.5, 0.5)),       # region #6: edge bc
    ((0.25, 0.25, 3), (0.25, 0.25, 0), (0.25, 0.25)),  # region #0: interior
])
def test_closest_point_regions(p, proj, bary):
    a, b, c = np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])
    pr, ba = O.closest_point_in_triangle_d(np.array(p, float), a, b, c)
    assert np.allclose(pr, proj, atol=1e-15)
    assert np.allclose(ba, bary, atol=1e-15)


def _closest_independent(p, a, b, c):
    """closest point by plane projection + segment clamps (independent formulation)"""
    n = np.cross(b - a, c - a)
    nn = n @ n
    q = p - ((p - a) @ n) / nn * n
    # barycentric inside test
    v0, v1, v2 = b - a, c - a, q - a
    d00, d01, d11, d20, d21 = v0 @ v0, v0 @ v1, v1 @ v1, v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    if v >= 0 and w >= 0 and v + w <= 1:
        return q
    best, bd = None, np.inf
    for s, e in ((a, b), (b, c), (c, a)):
        t = np.clip(((p - s) @ (e - s)) / ((e - s) @ (e - s)), 0, 1)
        x = s + t * (e - s)
        d = (x - p) @ (x - p)
        if d < bd:
            best, bd = x, d
    return best


def test_closest_point_random_vs_independent():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        pr, ba = O.closest_point_in_triangle_d(p, a, b, c)
        ind = _closest_independent(p, a, b, c)
        assert abs((pr - p) @ (pr - p) - (ind - p) @ (ind - p)) < 1e-10
        # barycentrics reproduce the point
        assert np.linalg.norm(a + ba[0] * (b - a) + ba[1] * (c - a) - pr) < 1e-9


# ---------------------------------------------------------------- tree + dipoles

@pytest.fixture(scope="module", params=[(0, 4, 1), (2, 8, 3), (4, 8, 5)])
def small_mesh(request):
    k, bumps, seed = request.param
    v, f = synth.bumpy_icosphere(k, bumps, seed)
    return v, f, O.Mesh(v, f)


def test_tree_invariants(small_mesh):
    v, f, m = small_mesh
    nodes = m.nodes()
    F = f.shape[0]
    assert len(nodes) == 2 * F - 1                     # MRAABBTreeTests.cpp:14
    # root box = mesh bbox insignificantly expanded (MRAABBTreeTests.cpp:15)
    lo = np.nextafter(v.min(0), np.float32(-np.inf))
    hi = np.nextafter(v.max(0), np.float32(np.inf))
    assert np.array_equal(nodes[0]["min"], lo) and np.array_equal(nodes[0]["max"], hi)
    leaves = nodes[nodes["r"] < 0]
    assert np.array_equal(np.sort(leaves["l"]), np.arange(F))       # every face once
    # leaf box = triangle bbox widened by one ulp each side
    tv = v[f[leaves["l"]]]
    assert np.array_equal(leaves["min"], np.nextafter(tv.min(1), np.float32(-np.inf)))
    assert np.array_equal(leaves["max"], np.nextafter(tv.max(1), np.float32(np.inf)))
    # pre-order layout, children inside parents, internal box = union of children
    for i in np.nonzero(nodes["r"] >= 0)[0]:
        l, r = nodes[i]["l"], nodes[i]["r"]
        assert l == i + 1 and r > l
        for ch in (l, r):
            assert np.all(nodes[ch]["min"] >= nodes[i]["min"]) and np.all(nodes[ch]["max"] <= nodes[i]["max"])
        assert np.array_equal(np.minimum(nodes[l]["min"], nodes[r]["min"]), nodes[i]["min"])
        assert np.array_equal(np.maximum(nodes[l]["max"], nodes[r]["max"]), nodes[i]["max"])
        # balanced median split: left subtree holds floor(n/2) leaves
        nl = (r - l + 1) // 2
        n_total = (len(nodes) if i == 0 else None)
        if n_total:
            assert nl == F // 2


def _median_tree_reference_python(v, f):
    """Independent Python build of the same deterministic tree (full sort instead of a
    selection): pins the split rule and the layout."""
    tv = v[f]
    bmin = np.nextafter(tv.min(1), np.float32(-np.inf))
    bmax = np.nextafter(tv.max(1), np.float32(np.inf))
    F = f.shape[0]
    out = []

    def rec(ids, root):
        while len(out) <= root:
            out.append(None)
        if len(ids) == 1:
            out[root] = (bmin[ids[0]], bmax[ids[0]], int(ids[0]), -1)
            return
        mn = bmin[ids].min(0)
        mx = bmax[ids].max(0)
        ext = (mx - mn).astype(np.float32)
        dims = sorted(range(3), key=lambda d: (ext[d], d))   # smallest first
        keys = [(bmin[ids, d] + bmax[ids, d]).astype(np.float32) for d in reversed(dims)]
        order = np.lexsort([ids] + keys[::-1])
        srt = ids[order]
        half = len(ids) // 2
        l, r = root + 1, root + 1 + (2 * half - 1)
        out[root] = (mn, mx, l, r)
        rec(srt[:half], l)
        rec(srt[half:], r)

    rec(np.arange(F), 0)
    return out


def test_tree_matches_independent_build():
    v, f = synth.bumpy_icosphere(3, 8, 11)
    nodes = O.Mesh(v, f).nodes()
    ref = _median_tree_reference_python(v, f)
    assert len(ref) == len(nodes)
    for i, (mn, mx, l, r) in enumerate(ref):
        assert np.array_equal(nodes[i]["min"], mn) and np.array_equal(nodes[i]["max"], mx)
        assert nodes[i]["l"] == l and nodes[i]["r"] == r


def test_dipoles_invariants(small_mesh):
    v, f, m = small_mesh
    nodes, dips = m.nodes(), m.dipoles()
    a, b, c = (v[f[:, i]].astype(np.float64) for i in range(3))
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    assert abs(dips[0]["area"] - area.sum()) < 1e-5 * area.sum()
    # closed mesh: total directed area vanishes
    assert np.linalg.norm(dips[0]["dirArea"]) < 1e-5 * area.sum()
    # leaves: recomputed in FP64
    leaf = nodes["r"] < 0
    fl = nodes["l"][leaf]
    assert np.allclose(dips["area"][leaf], area[fl], rtol=1e-6)
    cen = (a + b + c) / 3
    assert np.allclose(dips["pos"][leaf], cen[fl], atol=1e-6)
    # internal nodes (source/MRMesh/MRDipole.cpp:47-70): area = the sum of the areas of the leaves
    # below, pos = their area-weighted centre, dirArea = the sum of their directed areas.  The leaves
    # below node i are the leaf nodes of its pre-order range [i, end_i), end_i = end of its right
    # child's range (found from the r links alone), summed in FP64 from the triangles themselves.
    n = len(nodes)
    end = np.zeros(n, np.int64)
    for i in range(n - 1, -1, -1):
        end[i] = i + 1 if nodes["r"][i] < 0 else end[nodes["r"][i]]
    assert end[0] == n
    dA = 0.5 * cr
    w_area = np.zeros(n + 1)
    w_pos = np.zeros((n + 1, 3))
    w_dir = np.zeros((n + 1, 3))
    w_area[1:][leaf] = area[fl]
    w_pos[1:][leaf] = area[fl, None] * cen[fl]
    w_dir[1:][leaf] = dA[fl]
    c_area, c_pos, c_dir = np.cumsum(w_area), np.cumsum(w_pos, 0), np.cumsum(w_dir, 0)
    inner = np.nonzero(~leaf)[0]
    s_area = c_area[end[inner]] - c_area[inner]
    s_pos = (c_pos[end[inner]] - c_pos[inner]) / s_area[:, None]
    s_dir = c_dir[end[inner]] - c_dir[inner]
    scale = float(np.abs(v).max())
    assert np.allclose(dips["area"][inner], s_area, rtol=2e-5)
    assert np.all(np.abs(dips["pos"][inner] - s_pos) <= 2e-5 * scale)
    assert np.all(np.abs(dips["dirArea"][inner] - s_dir) <= 2e-5 * s_area[:, None] + 1e-7)
    # rr reaches the farthest box corner
    corners_max = np.maximum((dips["pos"] - nodes["min"]) ** 2, (dips["pos"] - nodes["max"]) ** 2).sum(1)
    assert np.allclose(dips["rr"], corners_max, rtol=1e-6)


def test_dipole_check_catches_planted_mistakes(small_mesh):
    # the internal-node check above must fail for an unweighted centre or a child dropped from a sum
    v, f, m = small_mesh
    nodes, dips = m.nodes(), m.dipoles()
    leaf = nodes["r"] < 0
    inner = np.nonzero(~leaf)[0]
    l, r = nodes["l"][inner], nodes["r"][inner]
    al, ar = dips["area"][l].astype(np.float64), dips["area"][r].astype(np.float64)
    pl, pr = dips["pos"][l].astype(np.float64), dips["pos"][r].astype(np.float64)
    right_pos = (al[:, None] * pl + ar[:, None] * pr) / (al + ar)[:, None]
    wrong_pos = 0.5 * (pl + pr)                                                       # not area-weighted
    assert np.max(np.abs(wrong_pos - right_pos)) > 2e-5 * float(np.abs(v).max())
    wrong_area = al                                                                   # right child dropped
    assert np.all(np.abs(wrong_area - dips["area"][inner]) > 2e-5 * dips["area"][inner])


def _delete_faces(f, frac, seed):
    rng = np.random.default_rng(seed)
    fd = f.copy()
    gone = rng.random(f.shape[0]) < frac
    fd[gone] = -1
    return fd, np.nonzero(~gone)[0].astype(np.int32)


def test_oracle_deleted_faces_equal_packed_mesh():
    # getTriangulation keeps {-1,-1,-1} for deleted faces (source/MRMesh/MRMeshTopology.cpp:430-441); the
    # reference's tree covers the valid faces, keeping their ids (source/MRMesh/MRAABBTree.cpp:35-49):
    # every result equals that of the packed mesh with the face ids mapped back
    v, f = synth.bumpy_icosphere(3, 8, 5)
    fd, ids = _delete_faces(f, 0.05, 3)
    md, mp = O.Mesh(v, fd), O.Mesh(v, f[ids])
    nd, npk = md.nodes(), mp.nodes()
    assert len(nd) == 2 * len(ids) - 1
    lf = npk["r"] < 0
    npk["l"][lf] = ids[npk["l"][lf]]
    assert np.array_equal(nd.view(np.uint8), npk.view(np.uint8))
    assert np.array_equal(md.dipoles().view(np.uint8), mp.dipoles().view(np.uint8))
    pts = synth.projection_queries(v, f, 4000, 12)
    rd, rp = md.find_projections(pts), mp.find_projections(pts)
    hit = rp["face"] >= 0
    rp["face"][hit] = ids[rp["face"][hit]]
    assert np.array_equal(rd.view(np.uint8), rp.view(np.uint8))
    assert np.array_equal(md.fwn_vector(pts).view(np.uint32), mp.fwn_vector(pts).view(np.uint32))
    dims, xf, _ = synth.grid_for_mesh(v, (12, 12, 12))
    assert np.array_equal(md.sdf_grid(dims, xf).view(np.uint32), mp.sdf_grid(dims, xf).view(np.uint32))
    sd, sp = md.self_intersections(), mp.self_intersections()
    assert sd.shape[0] == f.shape[0] and not sd[fd[:, 0] < 0].any()
    assert np.array_equal(sd[ids], sp)
    # the brute-force checkers skip deleted faces too
    assert np.array_equal(O.bf_nearest_f64(v, fd, pts[:300]), O.bf_nearest_f64(v, f[ids], pts[:300]))


def test_product_host_tree_deleted_faces():
    import meshlib_amd as M
    v, f = synth.bumpy_icosphere(4, 8, 6)
    fd, ids = _delete_faces(f, 0.05, 9)
    nodes, dips, depth = M.build_tree_host(v, fd)
    om = O.Mesh(v, fd)
    assert len(nodes) == 2 * len(ids) - 1
    assert np.array_equal(nodes.view(np.uint8), om.nodes().view(np.uint8))
    assert np.array_equal(dips.view(np.uint8), om.dipoles().view(np.uint8))
    # every face deleted: an empty tree
    nodes, dips, depth = M.build_tree_host(v, np.full_like(f, -1))
    assert len(nodes) == 0


def test_product_host_tree_matches_oracle():
    import meshlib_amd as M
    for k, bumps, seed in ((0, 4, 1), (3, 8, 2), (5, 8, 1)):
        v, f = synth.bumpy_icosphere(k, bumps, seed)
        nodes, dips, depth = M.build_tree_host(v, f)
        om = O.Mesh(v, f)
        assert np.array_equal(nodes.view(np.uint8), om.nodes().view(np.uint8))
        assert np.array_equal(dips.view(np.uint8), om.dipoles().view(np.uint8))
        assert depth == int(np.ceil(np.log2(f.shape[0])))


# ---------------------------------------------------------------- projection

def test_projection_vs_bruteforce():
    v, f = synth.bumpy_icosphere(3, 8, 4)
    m = O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 3000, 9)
    r = m.find_projections(pts)
    bf = O.bf_nearest_f64(v, f, pts)
    d64, f64 = m.find_projections_f64(pts)
    assert np.allclose(d64, bf, rtol=1e-12, atol=1e-15)          # same algorithm in FP64
    pn = np.linalg.norm(pts.astype(np.float64), axis=1) + 1.0
    tol = 4e-7 * np.sqrt(bf) * pn + 1e-12 * pn * pn
    assert np.all(np.abs(r["distSq"] - bf) <= tol)
    # the winning face attains the minimum in FP64
    vv = v.astype(np.float64)
    for i in range(0, len(pts), 7):
        tri = vv[f[r["face"][i]]]
        pr, _ = O.closest_point_in_triangle_d(pts[i].astype(np.float64), *tri)
        assert (pr - pts[i]) @ (pr - pts[i]) <= bf[i] * (1 + 1e-9) + 1e-15
    # bit-level self-consistency of the record: point = float(closest point on face)
    assert np.all(r["e"] == -1)


def test_projection_limits():
    v, f = synth.bumpy_icosphere(2, 8, 4)
    m = O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 500, 3)
    full = m.find_projections(pts)
    up = np.float32(0.01)
    r = m.find_projections(pts, up=up)
    found = full["distSq"] < up
    assert np.array_equal(r["face"] >= 0, found)
    assert np.array_equal(r[found].view(np.uint8), full[found].view(np.uint8))
    assert np.all(r["distSq"][~found] == up)
    # loDistLimitSq: stops at the first hit within lo, which is still within lo
    lo = np.float32(0.05)
    rl = m.find_projections(pts, lo=lo)
    assert np.all((rl["distSq"] <= lo) | (rl["distSq"] == full["distSq"]))


def _cube():
    v = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], np.float32)
    # outward-oriented 12 triangles
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    return v, np.array(t, np.int32)


def test_translated_cube_kat():
    # source/MRTestC2/MRPointsToMeshProjector.c:10-36: distances of a cube's vertices moved by
    # `shift` to the cube reach |shift| at most (the far corner)
    v, f = _cube()
    m = O.Mesh(v, f)
    shift = np.array([1, 2, 3], np.float32)
    xf = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, *shift], np.float32)
    r = m.find_projections(v, pt_xf=xf)
    d = np.sqrt(r["distSq"].astype(np.float64))
    assert abs(d.max() - np.linalg.norm(shift)) < 1e-6


def test_nonrigid_tree_xf_equals_transformed_mesh():
    # a non-rigid xf applied to the tree equals projecting onto the transformed mesh up to the
    # looser transformed boxes (same minimum distance)
    v, f = synth.bumpy_icosphere(2, 8, 4)
    m = O.Mesh(v, f)
    A = np.array([[1.5, 0.2, 0], [0, 0.7, 0.1], [0.1, 0, 1.2]], np.float32)
    b = np.array([0.3, -0.2, 0.1], np.float32)
    xf = np.concatenate([A.reshape(-1), b]).astype(np.float32)
    vt = (v.astype(np.float64) @ A.T.astype(np.float64) + b).astype(np.float32)
    pts = synth.projection_queries(vt, f, 400, 1)
    r = m.find_projections(pts, tree_xf=xf)
    bf = O.bf_nearest_f64(vt, f, pts)
    assert np.allclose(r["distSq"], bf, rtol=2e-5, atol=1e-9)


# ---------------------------------------------------------------- winding number

def test_fwn_exact_with_huge_beta_matches_bruteforce():
    v, f = synth.bumpy_icosphere(3, 8, 4)
    m = O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 300, 2)
    w = m.fwn_vector(pts, beta=1e6)        # no dipole is ever accepted: exact leaf sum
    bf = O.bf_winding_f64(v, f, pts)
    assert np.max(np.abs(w - bf)) < 2e-5


def test_fwn_float_vs_fp64_same_algorithm():
    v, f = synth.bumpy_icosphere(4, 8, 4)
    m = O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 3000, 2)
    w = m.fwn_vector(pts, beta=2.0)
    w64 = m.fwn_vector_f64(pts, beta=2.0)
    assert np.max(np.abs(w - w64)) < 1e-4


def test_fwn_inside_outside():
    v, f = synth.bumpy_icosphere(4, 8, 6)
    m = O.Mesh(v, f)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    inside = (d * 0.5).astype(np.float32)
    outside = (d * 2.0).astype(np.float32)
    # exact leaf sum: 1 inside, 0 outside a closed outward-oriented surface
    assert np.all(np.abs(m.fwn_vector(inside, beta=1e6) - 1) < 1e-4)
    assert np.all(np.abs(m.fwn_vector(outside, beta=1e6)) < 1e-4)
    # beta = 2 far-field approximation: first-order dipoles, error of a few 1e-3 here
    assert np.all(np.abs(m.fwn_vector(inside) - 1) < 2e-2)
    assert np.all(np.abs(m.fwn_vector(outside)) < 2e-2)


def test_self_intersections_closed_and_folded():
    v, f = synth.bumpy_icosphere(3, 8, 4)
    assert not O.Mesh(v, f).self_intersections().any()
    # two overlapping copies: faces of one copy inside the other see wn ~ 1.5 > 1
    shift = np.array([0.6, 0, 0], np.float32)
    v2 = np.concatenate([v, v + shift])
    f2 = np.concatenate([f, f + v.shape[0]])
    flags = O.Mesh(v2, f2).self_intersections()
    c = v2[f2].mean(1)
    other = np.where(np.arange(len(f2)) < len(f), 1, 0)  # copy index of the *other* surface
    cc = c - shift * other[:, None]                       # centroid seen from the other copy
    r = np.linalg.norm(cc, axis=1)
    assert np.all(flags[r < 0.8]) and not np.any(flags[r > 1.4])


def test_sdf_grid_matches_pointwise_definition():
    v, f = synth.bumpy_icosphere(3, 8, 4)
    m = O.Mesh(v, f)
    dims, xf, vox = synth.grid_for_mesh(v, (12, 10, 9))
    s = m.sdf_grid(dims, xf)
    pts = O.grid_points(dims, xf)
    r = m.find_projections(pts)
    w = m.fwn_vector(pts)
    exp = np.where(w > 0.5, -1.0, 1.0).astype(np.float32) * np.sqrt(r["distSq"])
    assert np.array_equal(s, exp)
    # FWN grid driver = vector driver on the same points
    assert np.array_equal(m.fwn_grid(dims, xf), w)
    # sample driver picks the same voxels
    ids = np.array([0, 5, 17, s.size - 1], np.int64)
    assert np.array_equal(m.sdf_grid_sample(dims, xf, ids), s[ids])
    # nullOutsideMinMax: NaN outside [min, max)
    o = O.default_opts(minDistSq=0.01, maxDistSq=0.09)
    sn = m.sdf_grid(dims, xf, opts=o)
    d2 = r["distSq"]
    out = (d2 < 0.01) | (d2 >= 0.09)
    assert np.all(np.isnan(sn[out]))
    assert np.array_equal(np.abs(sn[~out]), np.sqrt(d2[~out]))


def test_learned_bound_identity_on_oracle():
    # seeding upDistLimitSq with (d*(1+m))^2 for any m > 0 returns bit-identical records
    v, f = synth.bumpy_icosphere(4, 8, 4)
    m = O.Mesh(v, f)
    pts = synth.projection_queries(v, f, 4000, 5)
    r = m.find_projections(pts)
    d = np.sqrt(r["distSq"].astype(np.float64))
    for mm in (0.0001, 0.05, 0.25):
        ups = ((d * (1 + mm)) ** 2).astype(np.float32)
        ups = np.maximum(ups, np.nextafter(r["distSq"], np.float32(np.inf)))
        rb = m.find_projections_bounded(pts, ups)
        assert np.array_equal(rb.view(np.uint8), r.view(np.uint8))
    # a too-small bound finds nothing (face -1), signalling a rerun
    rb = m.find_projections_bounded(pts, r["distSq"])
    assert np.all(rb["face"] == -1)
