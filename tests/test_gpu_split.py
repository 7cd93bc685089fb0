"""Split form of the signed-distance walk: waves that run past the step budget hand their voxels to a
second launch that walks the tree's top subtrees in parallel.  The volume must not depend on the
budget: budget 1 sends almost every wave through the split path, 0 none, -1 picks automatically."""
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
def restore_budget():
    old = (M.get_split_budget("sdf"), M.get_split_budget("projections"))
    yield
    M.set_split_budget(old[0], "sdf")
    M.set_split_budget(old[1], "projections")
    M.set_leaf_filter("auto")


@pytest.mark.parametrize("level,bumps,grid", [(5, 8, 64), (8, 32, 96)])
def test_split_budgets_identical(level, bumps, grid, restore_budget):
    import torch
    v, f = synth.bumpy_icosphere(level, bumps, 3)
    dims, xf, vox = synth.grid_for_mesh(v, (grid, grid - 5, grid + 3))
    gm = M.Mesh(v, f)
    fwn = M.FastWindingNumber(gm)
    opts = M.DistanceToMeshOptions()
    M.set_split_budget(0)
    ref = fwn.calcFromGridWithDistances(dims, xf, opts)
    n = ref.size
    d = torch.empty(n, dtype=torch.float32, device="cuda")
    for budget in (1, 16, 4096, -1):
        say(f"k={level} grid {grid} budget {budget}")
        M.set_split_budget(budget)
        got = fwn.calcFromGridWithDistances(dims, xf, opts)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), budget
        M.set_leaf_filter("on" if M.get_leaf_filter() != "on" else "off")
        got = fwn.calcFromGridWithDistances(dims, xf, opts)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (budget, M.get_leaf_filter())
        d.fill_(-3.0)
        fwn.sdf_grid_device(dims, xf, opts, d.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(d.cpu().numpy().view(np.uint32), ref.view(np.uint32)), budget
    # by parts through the split path, and the oracle on the small mesh
    M.set_split_budget(1)
    parts = np.full(n, -5.0, np.float32)
    layer = int(dims[0]) * int(dims[1])

    def take(data, pdims, z0):
        parts[z0 * layer:(z0 + int(pdims[2])) * layer] = data
    M.set_memory_limit(layer * 4 * 2 * 2 * 7)
    try:
        fwn.calcFromGridWithDistancesByParts(take, dims, xf, opts, 0)
    finally:
        M.set_memory_limit(0)
    assert np.array_equal(parts.view(np.uint32), ref.view(np.uint32))
    if level <= 5:
        o = O.Mesh(v, f).sdf_grid(dims, xf)
        assert np.array_equal(np.abs(ref), np.abs(o))


@pytest.mark.parametrize("level,bumps,nq", [(4, 8, 30000), (7, 32, 60000)])
def test_split_projections_identical(level, bumps, nq, restore_budget):
    import torch
    v, f = synth.bumpy_icosphere(level, bumps, 4)
    gm = M.Mesh(v, f)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    # the query distribution plus exact ties (vertices, edge midpoints) and points near the centre,
    # whose walks are the longest
    rng = np.random.default_rng(level)
    vv = v.astype(np.float32)
    centre = (vv.mean(0) + rng.normal(0, 0.05, (2000, 3))).astype(np.float32)
    pts = np.concatenate([synth.projection_queries(v, f, nq, 8), vv[:3000],
                          (0.5 * (vv[f[:1000, 0]] + vv[f[:1000, 1]])).astype(np.float32), centre]).astype(np.float32)
    M.set_split_budget(0)
    ref = pr.findProjections(pts)
    n = pts.shape[0]
    dp = torch.from_numpy(pts).cuda()
    dr = torch.empty(n * 8, dtype=torch.int32, device="cuda")
    bound = torch.full((n,), 0.3, dtype=torch.float32, device="cuda")
    M.set_split_budget(0)
    pr.find_projections_bounded_device(dp.data_ptr(), n, bound.data_ptr(), dr.data_ptr())
    torch.cuda.synchronize()
    ref_b = dr.cpu().numpy().copy()
    for budget in (1, 16, 4096, -1):
        say(f"k={level} projections budget {budget}")
        M.set_split_budget(budget)
        got = pr.findProjections(pts)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), budget
        dr.fill_(7)
        pr.find_projections_device(dp.data_ptr(), n, dr.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(dr.cpu().numpy().view(np.uint8), ref.view(np.uint8)), budget
        dr.fill_(7)
        pr.find_projections_bounded_device(dp.data_ptr(), n, bound.data_ptr(), dr.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(dr.cpu().numpy(), ref_b), budget
    if level <= 4:
        o = O.Mesh(v, f).find_projections(pts)
        assert np.array_equal(ref.view(np.uint8), o.view(np.uint8))


def test_infinite_limits(restore_budget):
    """upDistLimitSq / maxDistSq = +inf: every pruning limit stays finite-or-infinite as in the
    reference, for the plain walks and with every wave split (budget 1)."""
    v, f = synth.bumpy_icosphere(4, 8, 5)
    gm = M.Mesh(v, f)
    om = O.Mesh(v, f)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    pts = synth.projection_queries(v, f, 20000, 13)
    dims, xf, vox = synth.grid_for_mesh(v, (40, 36, 44))
    fwn = M.FastWindingNumber(gm)
    o_proj = om.find_projections(pts, up=float("inf"))
    o_sdf = om.sdf_grid(dims, xf, O.default_opts(maxDistSq=float("inf")))
    for budget in (0, 1, 64):
        say("infinite limits, budget", budget)
        M.set_split_budget(budget)
        g = pr.findProjections(pts, upDistLimitSq=float("inf"))
        assert np.array_equal(g.view(np.uint8), o_proj.view(np.uint8)), budget
        s = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions(maxDistSq=float("inf")))
        assert np.array_equal(np.abs(s), np.abs(o_sdf)), budget
