"""The learned bound's network: the CPU oracle (oracle/mlp_oracle.py) pinned by finite differences,
then the GPU kernels (BF16 MFMA forward/backward, Adam) checked against it, trained on labels from
the GPU projector, and used as an exact cap for projections and signed distances."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _random_theta(seed, scale=0.2):
    rng = np.random.default_rng(seed)
    th = np.zeros(MO.NPARAMS, np.float32)
    W = rng.uniform(-scale, scale, size=(MO.LAYERS, MO.HID, MO.HID)).astype(np.float32)
    W[0][:, MO.NFEAT:] = 0
    th[MO.OFF_W:MO.OFF_B] = W.reshape(-1)
    th[MO.OFF_B:MO.OFF_W5] = rng.uniform(-0.1, 0.1, size=MO.LAYERS * MO.HID)
    th[MO.OFF_W5:MO.OFF_B5] = rng.uniform(-scale, scale, size=MO.HID)
    th[MO.OFF_B5] = 0.1
    return th


def test_oracle_gradient_matches_finite_differences():
    th = _random_theta(1).astype(np.float64)
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1, 1, size=(7, 3)).astype(np.float32)
    lab = rng.uniform(0, 0.5, size=7)
    c, s = np.zeros(3, np.float32), 1.0
    g = MO.gradient(th, pts, lab, c, s, exact=True)
    idx = list(rng.choice(MO.OFF_B, 12, replace=False)) + [MO.OFF_B + 5, MO.OFF_B + 300, MO.OFF_W5 + 7, MO.OFF_B5]
    idx = [i for i in idx if not (i < MO.HID * MO.HID and (i % MO.HID) >= MO.NFEAT)]
    for i in idx:
        e = 1e-4
        tp, tm = th.copy(), th.copy()
        tp[i] += e
        tm[i] -= e
        fd = (MO.loss(tp, pts, lab, c, s, exact=True) - MO.loss(tm, pts, lab, c, s, exact=True)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_oracle_bf16_rounding():
    # BF16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie (to even: 1), 1 + 3*2^-8 a tie (to even: 1 + 2^-6)
    x = np.array([1.0, 1.00390625, 1.01171875, 1.001953125, 1.005859375, -2.5], np.float32)
    r = MO.bf16(x)
    assert list(r) == [1.0, 1.0, 1.015625, 1.0, 1.0078125, -2.5]
    assert np.all(MO.bf16(r) == r)


# ---------------------------------------------------------------- GPU

def _mlp_module():
    import meshlib_amd as M
    return M


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    import torch
    M = _mlp_module()
    center, scale = np.array([0.1, -0.2, 0.05], np.float32), 1.3
    net = M.DistanceMLP(center, scale, seed=5)
    th = _random_theta(3)
    net.set_params(th)
    rng = np.random.default_rng(4)
    n = 300                                     # not a multiple of the 64-row tile
    pts = rng.uniform(-1.5, 1.5, size=(n, 3)).astype(np.float32)
    lab = rng.uniform(0.0, 0.8, size=n).astype(np.float32)
    dp, dl = torch.from_numpy(pts).cuda(), torch.from_numpy(lab).cuda()
    pred = net.predict(dp)
    o_pred, _ = MO.forward(th, pts, center, scale)
    assert np.max(np.abs(pred - o_pred)) <= 2e-2 * np.max(np.abs(o_pred)) + 1e-4, np.max(np.abs(pred - o_pred))
    g = net.gradient(dp, dl)
    og = MO.gradient(th, pts, lab, center, scale)
    assert g.shape[0] == MO.NPARAMS
    # the padded input columns of layer 0 get exactly zero gradient
    W0 = g[:MO.HID * MO.HID].reshape(MO.HID, MO.HID)
    assert np.all(W0[:, MO.NFEAT:] == 0)
    for name, sl in (("W", slice(MO.OFF_W, MO.OFF_B)), ("b", slice(MO.OFF_B, MO.OFF_W5)),
                     ("w5", slice(MO.OFF_W5, MO.OFF_B5)), ("b5", slice(MO.OFF_B5, MO.NPARAMS))):
        a, b = g[sl].astype(np.float64), og[sl]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-2, (name, err)


@pytest.mark.gpu
def test_gpu_training_reduces_loss_and_bound_is_exact():
    import torch
    M = _mlp_module()
    from meshlib_amd import synth
    v, f = synth.bumpy_icosphere(5, 8, 2)
    mesh = M.Mesh(v, f)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(mesh)
    lo, hi = v.min(0), v.max(0)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=1)
    pts = synth.projection_queries(v, f, 1 << 15, 11)
    dp = torch.from_numpy(pts).cuda()
    rec = torch.empty(pts.shape[0] * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(dp.data_ptr(), pts.shape[0], rec.data_ptr())
    lab = rec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    l0 = net.loss(dp, lab)
    net.train(dp, lab, steps=300, batch=4096, lr=2e-3)
    l1 = net.loss(dp, lab)
    assert l1 < 0.2 * l0, (l0, l1)
    # the learned cap keeps the projections exact, with some points rerun
    out = torch.empty_like(rec)
    reruns = pr.find_projections_learned_device(net, 0.1, dp.data_ptr(), pts.shape[0], out.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out, rec)
    assert 0 <= reruns < pts.shape[0] // 2
    dims, xf, vox = synth.grid_for_mesh(v, (24, 24, 24))
    fwn = M.FastWindingNumber(mesh)
    plain = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
    dsdf = torch.empty(plain.size, dtype=torch.float32, device="cuda")
    fwn.sdf_grid_learned_device(net, 0.1, dims, xf, M.DistanceToMeshOptions(), dsdf.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dsdf.cpu().numpy().view(np.uint32), plain.view(np.uint32))


@pytest.mark.gpu
def test_gpu_calibration_quantile_and_capped_reruns():
    """The slack is the requested quantile of label - prediction (never below 0), computed from the
    network's own predictions; with caps (pred (1 + m) + slack)^2 the projections that exceed their cap
    are exactly those the call reports as reruns, and the output stays exact."""
    import torch
    M = _mlp_module()
    from meshlib_amd import synth
    v, f = synth.bumpy_icosphere(4, 6, 3)
    mesh = M.Mesh(v, f)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(mesh)
    lo, hi = v.min(0), v.max(0)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=3)
    pts = synth.projection_queries(v, f, 20000, 5)
    dp = torch.from_numpy(pts).cuda()
    pred = net.predict(dp).astype(np.float64)
    # labels = prediction + a known excess per point: the slack is that excess's quantile
    rng = np.random.default_rng(9)
    excess = rng.uniform(-0.05, 0.2, size=pred.shape[0])
    lab = torch.from_numpy((pred + excess).astype(np.float32)).cuda()
    for q in (0.5, 0.9, 0.99):
        got = net.calibrate(dp, lab, q)
        under = np.sort((lab.cpu().numpy().astype(np.float64) - pred).astype(np.float32))
        want = max(0.0, float(under[min(len(under) - 1, int(np.floor(q * (len(under) - 1) + 0.5)))]))
        assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (q, got, want)
    # all labels under the prediction: slack 0
    assert net.calibrate(dp, torch.from_numpy((0.5 * pred).astype(np.float32)).cuda(), 0.99) == 0.0
    # caps with a slack: exact records, reruns = points whose true distance reaches the cap
    rec = torch.empty(pts.shape[0] * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(dp.data_ptr(), pts.shape[0], rec.data_ptr())
    true_d = rec.view(torch.float32).view(-1, 8)[:, 7].cpu().numpy()
    for slack in (0.0, 0.02):
        net.set_slack(slack)
        out = torch.empty_like(rec)
        reruns = pr.find_projections_learned_device(net, 0.1, dp.data_ptr(), pts.shape[0], out.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(out, rec), slack
        cap = (pred.astype(np.float32) * np.float32(1.1) + np.float32(slack)) ** 2
        expect = int(np.sum(~(true_d < cap)))
        assert abs(reruns - expect) <= max(2, expect // 1000), (slack, reruns, expect)
