"""Per-mesh settings (mi_mesh_set_mode): a mesh's own traversal, leaf filter and split budgets take
precedence over the process-wide ones, for that mesh only; results never depend on them.

Which form ran is read from the walks' own record counters (mi_mesh_set_stats): the wave-cooperative
walks count their walks, the per-lane form counts nothing, and a split pass counts in the call's
second launch block."""
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
def restore_process_modes():
    old = (M.get_traversal(), M.get_leaf_filter(), M.get_split_budget("sdf"), M.get_split_budget("projections"))
    yield
    M.set_traversal(old[0])
    M.set_leaf_filter(old[1])
    M.set_split_budget(old[2], "sdf")
    M.set_split_budget(old[3], "projections")


def _counted(torch, mesh, fn):
    buf = torch.zeros(M.STAT_LAUNCHES * M.STAT_STRIDE, dtype=torch.int64, device="cuda")
    mesh.set_stats(buf.data_ptr())
    try:
        r = fn()
        torch.cuda.synchronize()
    finally:
        mesh.set_stats(0)
    c = buf.view(M.STAT_LAUNCHES, M.STAT_STRIDE).cpu().numpy()
    return r, [{n: int(c[k][i]) for i, n in enumerate(M.STAT_NAMES)} for k in range(M.STAT_LAUNCHES)]


def test_mesh_modes_override_process_modes(restore_process_modes):
    import torch
    v, f = synth.bumpy_icosphere(5, 8, 4)
    dims, xf, vox = synth.grid_for_mesh(v, (40, 36, 44))
    pts = synth.projection_queries(v, f, 20000, 5)
    a, b = M.Mesh(v, f), M.Mesh(v, f)
    fa, fb = M.FastWindingNumber(a), M.FastWindingNumber(b)
    pa, pb = M.PointsToMeshProjector(), M.PointsToMeshProjector()
    pa.updateMeshData(a)
    pb.updateMeshData(b)
    opts = M.DistanceToMeshOptions()
    om = O.Mesh(v, f)
    ref_sdf = om.sdf_grid(dims, xf)
    ref_proj = om.find_projections(pts)

    assert a.get_mode("traversal") is None and a.get_mode("split_sdf") is None
    # process-wide per-lane; mesh a overrides it with the wave-cooperative form
    M.set_traversal("per_lane")
    a.set_mode("traversal", "wave_coop")
    assert a.get_mode("traversal") == "wave_coop" and b.get_mode("traversal") is None
    ga, ca = _counted(torch, a, lambda: fa.calcFromGridWithDistances(dims, xf, opts))
    gb, cb = _counted(torch, b, lambda: fb.calcFromGridWithDistances(dims, xf, opts))
    say(f"walks: mesh a {ca[0]['walks']}, mesh b {cb[0]['walks']}")
    assert ca[0]["walks"] > 0 and cb[0]["walks"] == 0
    for g in (ga, gb):
        assert np.array_equal(np.abs(g).view(np.uint32), np.abs(ref_sdf).view(np.uint32))
        far = np.abs(ref_sdf) > 1e-3 * float(vox.min())
        assert np.array_equal(np.signbit(g[far]), np.signbit(ref_sdf[far]))
    ra, _ = _counted(torch, a, lambda: pa.findProjections(pts))
    rb = pb.findProjections(pts)
    for r in (ra, rb):
        assert np.array_equal(r.view(np.uint8), ref_proj.view(np.uint8))

    # split budgets: process-wide off, mesh a hands every wave over after one step
    M.set_traversal("wave_coop")
    M.set_split_budget(0)
    a.set_mode("split_sdf", 1)
    a.set_mode("split_projections", 1)
    ga, ca = _counted(torch, a, lambda: fa.calcFromGridWithDistances(dims, xf, opts))
    gb, cb = _counted(torch, b, lambda: fb.calcFromGridWithDistances(dims, xf, opts))
    say(f"split-pass walks: mesh a {ca[1]['walks']}, mesh b {cb[1]['walks']}")
    assert ca[1]["walks"] > 0 and cb[1]["walks"] == 0
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    ra, ca = _counted(torch, a, lambda: pa.findProjections(pts))
    rb, cb = _counted(torch, b, lambda: pb.findProjections(pts))
    assert ca[1]["walks"] > 0 and cb[1]["walks"] == 0
    assert np.array_equal(ra.view(np.uint8), ref_proj.view(np.uint8))
    assert np.array_equal(rb.view(np.uint8), ref_proj.view(np.uint8))

    # leaf filter: mesh a forces it off while the process forces it on; results identical
    M.set_leaf_filter("on")
    a.set_mode("leaf_filter", "off")
    ga = fa.calcFromGridWithDistances(dims, xf, opts)
    gb = fb.calcFromGridWithDistances(dims, xf, opts)
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))

    # back to following the process-wide settings
    for mode in ("traversal", "leaf_filter", "split_sdf", "split_projections"):
        a.set_mode(mode, None)
        assert a.get_mode(mode) is None
    M.set_traversal("per_lane")
    _, ca = _counted(torch, a, lambda: fa.calcFromGridWithDistances(dims, xf, opts))
    assert ca[0]["walks"] == 0


def test_mesh_mode_rejects_bad_values():
    v, f = synth.bumpy_icosphere(2, 2, 1)
    m = M.Mesh(v, f)
    with pytest.raises(M.MeshLibAmdError):
        m.set_mode("split_sdf", -5)
    with pytest.raises(KeyError):
        m.set_mode("traversal", "sideways")
    rc = M.lib().mi_mesh_set_mode(m.handle, 9, 0)
    assert rc == M.MI_E_INVALID


def test_threads_on_meshes_with_their_own_settings():
    """Four host threads, each on its own mesh with its own settings (per-lane / wave-cooperative,
    split form on / off, leaf filter on / off), run host-pointer signed distance, projections and
    by-parts calls at the same time (the library releases no lock across meshes, and by-parts
    callers share the pool of page-locked buffers): every result equals the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    v, f = synth.bumpy_icosphere(5, 8, 4)
    dims, xf, vox = synth.grid_for_mesh(v, (48, 40, 44))
    pts = synth.projection_queries(v, f, 30000, 9)
    om = O.Mesh(v, f)
    ref_sdf = om.sdf_grid(dims, xf)
    ref_proj = om.find_projections(pts)
    far = np.abs(ref_sdf) > 1e-3 * float(vox.min())
    settings = [{"traversal": "per_lane"},
                {"traversal": "wave_coop", "split_sdf": 0, "split_projections": 0},
                {"traversal": "wave_coop", "split_sdf": 1, "split_projections": 1, "leaf_filter": "off"},
                {"traversal": "wave_coop", "leaf_filter": "on"}]
    meshes = []
    for s in settings:
        m = M.Mesh(v, f)
        for k, val in s.items():
            m.set_mode(k, val)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(m)
        meshes.append((m, M.FastWindingNumber(m), pr))
    layer = dims[0] * dims[1]

    def work(i):
        _, fwn, pr = meshes[i]
        errs = []
        for rep in range(3):
            g = fwn.calcFromGridWithDistances(dims, xf, M.DistanceToMeshOptions())
            if not (np.array_equal(np.abs(g).view(np.uint32), np.abs(ref_sdf).view(np.uint32))
                    and np.array_equal(np.signbit(g[far]), np.signbit(ref_sdf[far]))):
                errs.append(("sdf", rep))
            if not np.array_equal(pr.findProjections(pts).view(np.uint8), ref_proj.view(np.uint8)):
                errs.append(("projections", rep))
            got = np.full(ref_sdf.size, np.nan, np.float32)

            def take(data, d, z0):
                got[z0 * layer:(z0 + d[2]) * layer] = data
            fwn.calcFromGridWithDistancesByParts(take, dims, xf, M.DistanceToMeshOptions(), 1)
            if not np.array_equal(got.view(np.uint32), g.view(np.uint32)):
                errs.append(("by parts", rep))
        return errs

    with ThreadPoolExecutor(len(meshes)) as ex:
        results = list(ex.map(work, range(len(meshes))))
    say("per-thread mismatches:", results)
    assert all(not r for r in results)
