"""The learned-bound and batched workloads at their full sizes, against the oracle.

* Learned bound (k=8 bumpy sphere, 1,310,720 triangles): a 2,000-step Adam run of batch 8,192 on 2^20
  labels from the GPU projector, then 2^22 projections and a 256^3 signed-distance volume with the
  capped search at m = 0.05, 0.1, 0.25, with the bare cap and with the calibrated slack.  Every learned
  output equals the plain one bit-for-bit, and the plain outputs equal the oracle on seeded samples.
  The test prints what the network gains (or costs) against the plain walk: time ratio, rerun rate,
  and the walks' node visits.
* Batched uneven workload (64 bumpy spheres, split levels 4..8, grids 32..192, 2^18..2^21 queries):
  one persistent launch; per item a seeded sample of 2,000 voxels and 2,000 points equals the oracle
  (magnitudes and records bit-exact, sign identical beyond 1e-3 voxel), plain and learned batches.
"""
import sys
import time

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_sdf_sample(om, dims, xf, vox, got, ids):
    o = om.sdf_grid_sample(dims, xf, ids)
    g = got[ids]
    assert np.array_equal(np.abs(g).view(np.uint32), np.abs(o).view(np.uint32))
    far = np.abs(o) > 1e-3 * float(np.min(vox))
    assert np.array_equal(np.signbit(g[far]), np.signbit(o[far]))


@pytest.mark.timeout(600)
def test_learned_bound_k8_full_size():
    import torch
    v, f = synth.bumpy_icosphere(8, 32, 2)
    dims, xf, vox = synth.grid_for_mesh(v, (256, 256, 256))
    nq, nvox = 1 << 22, 256 ** 3
    pts = synth.projection_queries(v, f, nq, 7)
    gm = M.Mesh(v, f)
    fwn = M.FastWindingNumber(gm)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    opts = M.DistanceToMeshOptions()
    dp = torch.from_numpy(pts).cuda()
    d_proj = torch.empty(nq * 8, dtype=torch.int32, device="cuda")
    d_sdf = torch.empty(nvox, dtype=torch.float32, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        ev[0].record()
        r = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]), r

    for _ in range(2):     # the second run is the timed one (code objects loaded, caches warm)
        t_pp, _ = timed(lambda: pr.find_projections_device(dp.data_ptr(), nq, d_proj.data_ptr()))
        t_ps, _ = timed(lambda: fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr()))
    plain_proj, plain_sdf = d_proj.cpu().numpy(), d_sdf.cpu().numpy()
    say(f"plain: projections {t_pp:.1f} ms, sdf {t_ps:.1f} ms")

    # plain against the oracle on seeded samples
    om = O.Mesh(v, f)
    rng = np.random.default_rng(31)
    sel = rng.choice(nq, 2000, replace=False)
    ref = om.find_projections(pts[sel])
    assert np.array_equal(plain_proj.reshape(-1, 8)[sel].view(np.uint8).ravel(), ref.view(np.uint8).ravel())
    _check_sdf_sample(om, dims, xf, vox, plain_sdf, rng.choice(nvox, 2000, replace=False).astype(np.int64))

    # the network: 2^20 labels from the GPU projector, 2,000 Adam steps of 8,192
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=11)
    nt = 1 << 20
    tp = torch.from_numpy(synth.projection_queries(v, f, nt, 11)).cuda()
    trec = torch.empty(nt * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(tp.data_ptr(), nt, trec.data_ptr())
    tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    torch.cuda.synchronize()
    loss0 = net.loss(tp[:65536], tlab[:65536])
    t_train, _ = timed(lambda: net.train(tp, tlab, steps=2000, batch=8192, lr=1e-3))
    loss1 = net.loss(tp[:65536], tlab[:65536])
    say(f"training: {t_train:.0f} ms, loss {loss0:.4g} -> {loss1:.4g}")
    assert loss1 < 0.1 * loss0

    # node visits of the walks (their own counters), plain against learned
    def visits(fn):
        buf = torch.zeros(M.STAT_LAUNCHES * M.STAT_STRIDE, dtype=torch.int64, device="cuda")
        gm.set_stats(buf.data_ptr())
        try:
            fn()
            torch.cuda.synchronize()
        finally:
            gm.set_stats(0)
        c = buf.view(M.STAT_LAUNCHES, M.STAT_STRIDE).cpu().numpy()
        return int(c[:, M.STAT_NAMES.index("keyed_steps")].sum() + c[:, M.STAT_NAMES.index("quad_steps")].sum())
    v_plain = visits(lambda: (pr.find_projections_device(dp.data_ptr(), nq, d_proj.data_ptr()),
                              fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr())))

    # the bare cap (pred (1 + m))^2 and the calibrated one (slack: the 0.99 quantile of held-out
    # under-predictions); the bare cap reruns many queries, which exercises the rerun passes
    cp = torch.from_numpy(synth.projection_queries(v, f, 1 << 16, 99)).cuda()
    crec = torch.empty((1 << 16) * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(cp.data_ptr(), 1 << 16, crec.data_ptr())
    clab = crec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    slack = net.calibrate(cp, clab, 0.99)
    assert slack > 0
    for sl in (0.0, slack):
        net.set_slack(sl)
        for m in (0.05, 0.1, 0.25):
            d_proj.fill_(0)
            t_lp, rp = timed(lambda: pr.find_projections_learned_device(net, m, dp.data_ptr(), nq, d_proj.data_ptr()))
            d_sdf.fill_(0)
            t_ls, rs = timed(lambda: fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr()))
            assert np.array_equal(d_proj.cpu().numpy(), plain_proj), (sl, m)
            assert np.array_equal(d_sdf.cpu().numpy().view(np.uint32), plain_sdf.view(np.uint32)), (sl, m)
            if sl == 0.0 and m == 0.05:
                assert rp > 0 and rs > 0      # the rerun passes ran
            v_l = visits(lambda: (pr.find_projections_learned_device(net, m, dp.data_ptr(), nq, d_proj.data_ptr()),
                                  fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr())))
            say(f"slack={sl:.4f} m={m}: projections {t_lp:.1f} ms ({t_pp / t_lp:.2f}x plain, {rp / nq:.4f} rerun), "
                f"sdf {t_ls:.1f} ms ({t_ps / t_ls:.2f}x plain, {rs / nvox:.4f} rerun), cooperative node visits "
                f"{v_l / v_plain:.3f} of plain (reruns walk per lane, uncounted)")


@pytest.fixture(scope="module")
def batched_workload():
    import torch
    t0 = time.time()
    items, geoms = [], []
    for level, bumps, mseed, side, nq, qseed in synth.batched_workload_specs():
        v, f = synth.bumpy_icosphere(level, bumps, mseed)
        dims, xf, vox = synth.grid_for_mesh(v, (side,) * 3)
        pts = synth.projection_queries(v, f, nq, qseed)
        mesh = M.Mesh(v, f)
        it = {"mesh": mesh, "dims": dims, "xf": xf, "n": nq,
              "sdf": torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda"),
              "pts": torch.from_numpy(pts).cuda(),
              "proj": torch.empty(nq * 8, dtype=torch.int32, device="cuda")}
        it.update(d_sdf=it["sdf"].data_ptr(), d_pts=it["pts"].data_ptr(), d_proj=it["proj"].data_ptr())
        items.append(it)
        geoms.append((level, v, f, vox, pts))
    say(f"batched inputs: {len(items)} meshes in {time.time() - t0:.1f} s")
    return items, geoms


@pytest.mark.timeout(900)
def test_batched_64_meshes_vs_oracle(batched_workload):
    import torch
    items, geoms = batched_workload
    batch = M.Batch()
    ms = batch.autotune(items)
    say(f"autotuned: {batch.config()} {ms:.1f} ms")
    for it in items:
        it["sdf"].fill_(-7.0)
        it["proj"].fill_(0)
    batch.run(items)
    torch.cuda.synchronize()
    plain = [(it["sdf"].cpu().numpy(), it["proj"].cpu().numpy()) for it in items]
    rng = np.random.default_rng(17)
    t0 = time.time()
    for i, ((level, v, f, vox, pts), it, (sdf, proj)) in enumerate(zip(geoms, items, plain)):
        om = O.Mesh(v, f)
        nvox = sdf.size
        _check_sdf_sample(om, it["dims"], it["xf"], vox, sdf,
                          rng.choice(nvox, min(2000, nvox), replace=False).astype(np.int64))
        sel = rng.choice(it["n"], 2000, replace=False)
        ref = om.find_projections(pts[sel])
        got = proj.reshape(-1, 8)[sel]
        assert np.array_equal(got.view(np.uint8).ravel(), ref.view(np.uint8).ravel()), i
        del om
    say(f"oracle samples of {len(items)} items checked in {time.time() - t0:.1f} s")

    # learned batch: one small network per mesh (2^16 labels, 300 Adam steps), caps at m = 0.1
    t0 = time.time()
    nets = []
    for i, ((level, v, f, vox, pts), it) in enumerate(zip(geoms, items)):
        lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
        net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=100 + i)
        nt = 1 << 16
        tp = torch.from_numpy(synth.projection_queries(v, f, nt, 3000 + i)).cuda()
        trec = torch.empty(nt * 8, dtype=torch.int32, device="cuda")
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(it["mesh"])
        pr.find_projections_device(tp.data_ptr(), nt, trec.data_ptr())
        tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
        net.train(tp, tlab, steps=300, batch=8192, lr=1e-3)
        nets.append(net)
    torch.cuda.synchronize()
    say(f"64 networks trained in {time.time() - t0:.1f} s")
    for it in items:
        it["sdf"].fill_(-7.0)
        it["proj"].fill_(0)
    reruns = batch.run_learned(items, nets, 0.1)
    torch.cuda.synchronize()
    total = sum(it["sdf"].numel() + it["n"] for it in items)
    say(f"learned batch: {reruns} reruns ({reruns / total:.3f})")
    assert 0 < reruns < total
    for i, (it, (sdf, proj)) in enumerate(zip(items, plain)):
        assert np.array_equal(it["sdf"].cpu().numpy().view(np.uint32), sdf.view(np.uint32)), i
        assert np.array_equal(it["proj"].cpu().numpy(), proj), i
