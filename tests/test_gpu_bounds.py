"""Per-point search bounds on the GPU: the ICP-style bounded projection (per-point upDistLimitSq) and
the capped exact forms that the learned bound feeds (first pass under the cap, uncapped rerun of the
points that found nothing) -- records must equal the uncapped reference records."""
import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.fixture(scope="module")
def k5():
    v, f = synth.bumpy_icosphere(5, 8, 3)
    return v, f, M.Mesh(v, f), O.Mesh(v, f)


def test_bounded_projection_matches_oracle(k5):
    import torch
    v, f, gm, om = k5
    pts = synth.projection_queries(v, f, 40000, 5)
    ups = np.random.default_rng(1).uniform(0.0, 0.05, size=pts.shape[0]).astype(np.float32)
    o = om.find_projections_bounded(pts, ups)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    dp, du = torch.from_numpy(pts).cuda(), torch.from_numpy(ups).cuda()
    dr = torch.empty(pts.shape[0] * 8, dtype=torch.int32, device="cuda")
    for mode in ("wave_coop", "per_lane"):
        M.set_traversal(mode)
        pr.find_projections_bounded_device(dp.data_ptr(), pts.shape[0], du.data_ptr(), dr.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(_bits(dr.cpu().numpy()), _bits(o)), mode
    M.set_traversal("wave_coop")


def test_capped_projection_is_exact(k5):
    import torch
    v, f, gm, om = k5
    pts = synth.projection_queries(v, f, 60000, 6)
    o = om.find_projections(pts)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    dp = torch.from_numpy(pts).cuda()
    dr = torch.empty(pts.shape[0] * 8, dtype=torch.int32, device="cuda")
    d = np.sqrt(o["distSq"].astype(np.float64))
    rng = np.random.default_rng(2)
    for name, caps in (("generous", (d * 1.1) ** 2 + 1e-12),
                       ("mixed", (d * rng.uniform(0.5, 1.5, size=d.shape)) ** 2),
                       ("zero", np.zeros_like(d))):
        caps = caps.astype(np.float32)
        dc = torch.from_numpy(caps).cuda()
        dr.fill_(0)
        reruns = pr.find_projections_capped_device(dp.data_ptr(), pts.shape[0], dc.data_ptr(), dr.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(_bits(dr.cpu().numpy()), _bits(o)), name
        expect = int(np.sum(~(o["distSq"] < caps)))
        assert reruns == expect, (name, reruns, expect)


def test_capped_sdf_is_exact(k5):
    import torch
    v, f, gm, om = k5
    dims, xf, vox = synth.grid_for_mesh(v, (40, 36, 44))
    fwn = M.FastWindingNumber(gm)
    plain = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    n = plain.size
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(4)
    for name, caps in (("generous", (np.abs(plain) * 1.2) ** 2 + 1e-12),
                       ("mixed", (np.abs(plain) * rng.uniform(0.6, 1.4, size=n)) ** 2),
                       ("zero", np.zeros(n))):
        dc = torch.from_numpy(caps.astype(np.float32)).cuda()
        out.fill_(0)
        reruns = fwn.sdf_grid_capped_device(dims, xf, M.DistanceToMeshOptions(), dc.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(_bits(out.cpu().numpy()), _bits(plain)), name
        if name == "zero":
            assert reruns == n
    # a slab of layers with caps
    caps = (np.abs(plain) * 1.05) ** 2
    slab = slice(9 * 40 * 36, 23 * 40 * 36)
    dc = torch.from_numpy(caps[slab].astype(np.float32)).cuda()
    o2 = torch.empty(14 * 40 * 36, dtype=torch.float32, device="cuda")
    fwn.sdf_grid_capped_device(dims, xf, M.DistanceToMeshOptions(), dc.data_ptr(), o2.data_ptr(), zBegin=9, zEnd=23)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(o2.cpu().numpy()), _bits(plain[slab]))
