"""Persistent batched launch over uneven items: outputs equal the per-item calls bit-for-bit for
every autotuner configuration (and the per-item calls equal the oracle, tested elsewhere)."""
import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth

pytestmark = pytest.mark.gpu


def _bits(t):
    return t.cpu().numpy().view(np.uint8)


def test_batch_matches_per_item_calls():
    import torch
    rng = np.random.default_rng(0)
    specs = [(3, 8, 1, (20, 17, 9), 3000), (5, 8, 2, (33, 40, 21), 0), (4, 8, 3, None, 7777),
             (2, 8, 4, (8, 8, 8), 64), (5, 16, 5, (24, 24, 24), 20000)]
    items, expect = [], []
    keep = []
    for k, b, seed, gdims, npts in specs:
        v, f = synth.bumpy_icosphere(k, b, seed)
        mesh = M.Mesh(v, f)
        fwn = M.FastWindingNumber(mesh)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(mesh)
        it = {"mesh": mesh}
        exp = {}
        if gdims:
            dims, xf, vox = synth.grid_for_mesh(v, gdims)
            sdf = torch.full((int(np.prod(dims)),), -7.0, dtype=torch.float32, device="cuda")
            it.update(dims=dims, xf=xf, d_sdf=sdf.data_ptr())
            exp["sdf"] = (sdf, fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions()))
        if npts:
            pts = synth.projection_queries(v, f, npts, seed + 10)
            dp = torch.from_numpy(pts).cuda()
            dr = torch.zeros(npts * 8, dtype=torch.int32, device="cuda")
            it.update(d_pts=dp.data_ptr(), n=npts, d_proj=dr.data_ptr())
            exp["proj"] = (dr, pr.findProjections(pts))
            keep.append(dp)
        items.append(it)
        expect.append(exp)
        keep.append(mesh)
    batch = M.Batch()

    def check():
        torch.cuda.synchronize()
        for exp in expect:
            if "sdf" in exp:
                got, ref = exp["sdf"]
                assert np.array_equal(_bits(got), ref.view(np.uint8))
            if "proj" in exp:
                got, ref = exp["proj"]
                assert np.array_equal(_bits(got), ref.view(np.uint8))

    ms = batch.autotune(items)
    assert ms > 0
    check()
    for bpc in (1, 3, 6):
        for split in (0, 1):
            for exp in expect:
                for key in exp:
                    exp[key][0].fill_(0)
            batch.set_config(bpc, split)
            batch.run(items)
            check()
