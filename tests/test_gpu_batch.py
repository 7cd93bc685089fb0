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


def _uneven_items(torch):
    specs = [(3, 8, 1, (20, 17, 9), 3000), (5, 8, 2, (33, 40, 21), 0), (4, 8, 3, None, 7777),
             (5, 16, 5, (24, 24, 24), 20000)]
    items, keep = [], []
    for k, b, seed, gdims, npts in specs:
        v, f = synth.bumpy_icosphere(k, b, seed)
        mesh = M.Mesh(v, f)
        it = {"mesh": mesh, "v": v, "f": f}
        if gdims:
            dims, xf, vox = synth.grid_for_mesh(v, gdims)
            sdf = torch.full((int(np.prod(dims)),), -7.0, dtype=torch.float32, device="cuda")
            it.update(dims=dims, xf=xf, d_sdf=sdf.data_ptr(), sdf=sdf)
        if npts:
            dp = torch.from_numpy(synth.projection_queries(v, f, npts, seed + 10)).cuda()
            dr = torch.zeros(npts * 8, dtype=torch.int32, device="cuda")
            it.update(d_pts=dp.data_ptr(), n=npts, d_proj=dr.data_ptr(), pts=dp, proj=dr)
        items.append(it)
    return items


def _snapshot(items):
    return [(it["sdf"].clone() if "sdf" in it else None, it["proj"].clone() if "proj" in it else None)
            for it in items]


def _same(items, ref):
    import torch
    torch.cuda.synchronize()
    for it, (s, p) in zip(items, ref):
        if s is not None:
            assert torch.equal(it["sdf"].view(torch.int32), s.view(torch.int32))
        if p is not None:
            assert torch.equal(it["proj"], p)


def _clear(items):
    for it in items:
        if "sdf" in it:
            it["sdf"].fill_(-7.0)
        if "proj" in it:
            it["proj"].fill_(0)


def test_batch_capped_and_learned_match_plain():
    import torch
    items = _uneven_items(torch)
    batch = M.Batch()
    batch.run(items)
    ref = _snapshot(items)
    # caps scattered on both sides of the true distances: factor in [0.5, 1.5]
    gen = torch.Generator(device="cuda").manual_seed(5)
    caps_s, caps_p, keep = [], [], []
    for it, (s, p) in zip(items, ref):
        cs = cp = None
        if s is not None:
            fac = 0.5 + torch.rand(s.shape, generator=gen, device="cuda")
            c = (s.abs() * fac) ** 2
            keep.append(c)
            cs = c.data_ptr()
        if p is not None:
            fac = 0.5 + torch.rand((p.shape[0] // 8,), generator=gen, device="cuda")
            c = p.view(torch.float32).view(-1, 8)[:, 7] * fac * fac
            c = c.contiguous()
            keep.append(c)
            cp = c.data_ptr()
        caps_s.append(cs)
        caps_p.append(cp)
    _clear(items)
    reruns = batch.run_capped(items, caps_s, caps_p)
    _same(items, ref)
    total = sum(int(np.prod(it["dims"])) if "dims" in it else 0 for it in items) + sum(it.get("n", 0) for it in items)
    assert 0 < reruns < total
    # caps on some items only
    _clear(items)
    batch.run_capped(items, [caps_s[0], None, None, caps_s[3]], None)
    _same(items, ref)
    # learned caps from per-item networks (briefly trained; exactness does not depend on quality)
    nets = []
    for i, it in enumerate(items):
        if i == 1:
            nets.append(None)
            continue
        v, f = it["v"], it["f"]
        lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
        net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=i)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(it["mesh"])
        tp = torch.from_numpy(synth.projection_queries(v, f, 1 << 14, 40 + i)).cuda()
        tr = torch.empty((1 << 14) * 8, dtype=torch.int32, device="cuda")
        pr.find_projections_device(tp.data_ptr(), 1 << 14, tr.data_ptr())
        lab = tr.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
        net.train(tp, lab, steps=100, batch=4096, lr=1e-3)
        nets.append(net)
    for m in (0.05, 0.25):
        _clear(items)
        r = batch.run_learned(items, nets, m)
        _same(items, ref)
        assert r >= 0
