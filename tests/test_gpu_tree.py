"""The GPU tree + dipole builder against the host builder (which the oracle pins against the
reference's tree rules): nodes and dipoles bit-identical, same depth, and the traversals over the
device-built layout return the host-built layout's results bit-for-bit.

Cases cover the tie rules of the split order: a planar grid (equal keys along whole rows), every
triangle duplicated (ties down to the face id), zero-area triangles, and the 1- and 2-face trees."""
import sys

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grid(g, dup):
    xs, ys = np.meshgrid(np.arange(g + 1, dtype=np.float32), np.arange(g + 1, dtype=np.float32))
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size, np.float32)], 1)
    t = []
    for i in range(g):
        for j in range(g):
            a = i * (g + 1) + j
            t += [[a, a + 1, a + g + 2], [a, a + g + 2, a + g + 1]]
    return v, np.array(t * dup, np.int32)


def _cases():
    out = []
    for k, b, s in [(2, 8, 1), (3, 8, 2), (5, 8, 4), (6, 32, 5)]:
        out.append((f"ico{k}",) + synth.bumpy_icosphere(k, b, s))
    out.append(("grid",) + _grid(24, 1))
    out.append(("grid-dup",) + _grid(16, 3))
    v, f = synth.bumpy_icosphere(3, 8, 7)
    f2 = f.copy()
    f2[::7, 2] = f2[::7, 1]          # zero-area triangles
    out.append(("degenerate", v, f2))
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float32)
    out.append(("one", tri, np.array([[0, 1, 2]], np.int32)))
    out.append(("two", tri, np.array([[0, 1, 2], [1, 3, 2]], np.int32)))
    return out


@pytest.fixture
def restore_builder():
    yield
    M.set_tree_builder("device")


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_device_tree_equals_host_tree(case, restore_builder):
    name, v, f = case
    say(f"tree case {name}: {f.shape[0]} faces")
    M.set_tree_builder("device")
    gd = M.Mesh(v, f)
    nd, dd = gd.tree()
    hn, hd, hdepth = M.build_tree_host(v, f)
    assert gd.depth() == hdepth
    assert np.array_equal(nd.view(np.uint8), hn.view(np.uint8))
    assert np.array_equal(dd.view(np.uint8), hd.view(np.uint8))
    # the flattened layouts: identical results from meshes built both ways
    M.set_tree_builder("host")
    gh = M.Mesh(v, f)
    rng = np.random.default_rng(1)
    lo, hi = v.min(0), v.max(0)
    pts = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, size=(3000, 3))).astype(np.float32)
    ph, pd = M.PointsToMeshProjector(), M.PointsToMeshProjector()
    ph.updateMeshData(gh)
    pd.updateMeshData(gd)
    assert np.array_equal(ph.findProjections(pts).view(np.uint8), pd.findProjections(pts).view(np.uint8))
    wh = M.FastWindingNumber(gh).calcFromVector(pts, 2.0)
    wd = M.FastWindingNumber(gd).calcFromVector(pts, 2.0)
    assert np.array_equal(wh.view(np.uint32), wd.view(np.uint32))


def test_device_tree_large_and_invalid(restore_builder):
    import time
    v, f = synth.bumpy_icosphere(8, 32, 2)
    M.set_tree_builder("device")
    t = time.perf_counter()
    gd = M.Mesh(v, f)
    t_dev = time.perf_counter() - t
    nd, dd = gd.tree()
    t = time.perf_counter()
    hn, hd, hdepth = M.build_tree_host(v, f)
    t_host = time.perf_counter() - t
    say(f"k=8 tree: device {t_dev:.3f}s (incl. upload), host {t_host:.3f}s")
    assert gd.depth() == hdepth
    assert np.array_equal(nd.view(np.uint8), hn.view(np.uint8))
    assert np.array_equal(dd.view(np.uint8), hd.view(np.uint8))
    bad = f.copy()
    bad[5, 1] = v.shape[0]
    with pytest.raises(M.MeshLibAmdError):
        M.Mesh(v, bad)
