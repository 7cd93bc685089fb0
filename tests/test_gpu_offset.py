"""Offset-scale volume: bumpy icosphere split 9 times (5,242,880 triangles, 64 bumps), 512^3 signed
distances with FWN beta 2, plain and learned-bound traversals.

* the learned-bound run (MLP-predicted caps, uncapped rerun of the queries whose cap was too small)
  returns the plain run's volume bit-for-bit;
* against the CPU oracle on a seeded 1 % voxel sample plus every voxel within 2 voxels of the
  surface: magnitudes bit-identical and the inside/outside sign identical wherever the distance
  exceeds 1e-3 voxel.
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


@pytest.mark.timeout(600)
def test_offset_k9_512_plain_learned_and_oracle():
    import torch
    v, f = synth.bumpy_icosphere(9, 64, 5)
    dims, xf, vox = synth.grid_for_mesh(v, (512, 512, 512))
    nvox = 512 ** 3
    say(f"k=9: {f.shape[0]} tris, 512^3")
    gm = M.Mesh(v, f)
    fwn = M.FastWindingNumber(gm)
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    opts = M.DistanceToMeshOptions()
    d_sdf = torch.empty(nvox, dtype=torch.float32, device="cuda")
    fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr())
    torch.cuda.synchronize()
    plain = d_sdf.cpu().numpy()
    say("plain volume done")

    # learned bound: a briefly trained net is enough to exercise caps on both sides of the truth
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=3)
    ntrain = 1 << 18
    tp = torch.from_numpy(synth.projection_queries(v, f, ntrain, 21)).cuda()
    trec = torch.empty(ntrain * 8, dtype=torch.int32, device="cuda")
    pr.find_projections_device(tp.data_ptr(), ntrain, trec.data_ptr())
    tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    net.train(tp, tlab, steps=300, batch=8192, lr=1e-3)
    d_l = torch.full((nvox,), -1.0, dtype=torch.float32, device="cuda")
    reruns = fwn.sdf_grid_learned_device(net, 0.1, dims, xf, opts, d_l.data_ptr())
    torch.cuda.synchronize()
    say(f"learned volume done, {reruns} reruns")
    assert 0 < reruns < nvox
    assert np.array_equal(d_l.cpu().numpy().view(np.uint32), plain.view(np.uint32))
    del d_l, d_sdf

    # oracle on the seeded 1 % sample and the 2-voxel band
    assert not np.isnan(plain).any()
    rng = np.random.default_rng(5)
    sample = rng.choice(nvox, nvox // 100, replace=False)
    band = np.flatnonzero(np.abs(plain) <= 2.0 * float(np.max(vox)))
    ids = np.union1d(sample, band).astype(np.int64)
    say(f"oracle check on {ids.size} voxels ({band.size} in the band)")
    om = O.Mesh(v, f)
    o = om.sdf_grid_sample(dims, xf, ids)
    g = plain[ids]
    assert np.array_equal(np.abs(g).view(np.uint32), np.abs(o).view(np.uint32))
    far = np.abs(o) > 1e-3 * float(np.min(vox))
    assert np.array_equal(np.signbit(g[far]), np.signbit(o[far]))
