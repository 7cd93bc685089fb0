"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle must
reproduce them bit-for-bit; the brute-force FP64 columns pin the oracle independently; the GPU
path must reproduce the projection records and distance magnitudes bit-for-bit."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.fixture(params=FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def fx(request):
    with np.load(request.param, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 2


def test_oracle_reproduces_fixture(fx):
    m = O.Mesh(fx["verts"], fx["tris"])
    assert np.array_equal(m.nodes().view(np.uint8).reshape(-1, 32), fx["nodes"])
    proj = m.find_projections(fx["pts"])
    assert np.array_equal(proj.view(np.uint8).reshape(-1, 32), fx["proj"])
    assert np.array_equal(m.sdf_grid(fx["dims"], fx["xf"]).view(np.uint32), fx["sdf"].view(np.uint32))
    assert np.array_equal(m.fwn_vector(fx["pts"]), fx["fwn_pts"])
    assert np.array_equal(m.fwn_grid(fx["dims"], fx["xf"]), fx["fwn_grid"])
    assert np.array_equal(m.self_intersections(), fx["selfint"])


def test_fixture_against_bruteforce(fx):
    proj = fx["proj"].reshape(-1).view(O.PROJ_DTYPE)
    bf = fx["bf_dist_sq"]
    pn = np.linalg.norm(fx["pts"].astype(np.float64), axis=1) + 1
    assert np.all(np.abs(proj["distSq"] - bf) <= 4e-7 * np.sqrt(bf) * pn + 1e-12 * pn * pn)
    # beta = 2 dipole approximation error stays small away from the surface
    far = bf > 0.01
    assert np.max(np.abs(fx["fwn_pts"][far] - fx["bf_winding"][far])) < 5e-2


@pytest.mark.gpu
def test_gpu_reproduces_fixture(fx):
    import meshlib_amd as M
    gm = M.Mesh(fx["verts"], fx["tris"])
    pr = M.PointsToMeshProjector()
    pr.updateMeshData(gm)
    g = pr.findProjections(fx["pts"])
    assert np.array_equal(g.view(np.uint8).reshape(-1, 32), fx["proj"])
    fwn = M.FastWindingNumber(gm)
    s = fwn.calcFromGridWithDistances(fx["dims"], fx["xf"], M.DistanceToMeshOptions())
    assert np.array_equal(np.abs(s), np.abs(fx["sdf"]))
    w = fwn.calcFromVector(fx["pts"], 2.0)
    assert np.max(np.abs(w - fx["fwn_pts"])) <= 1e-4
    assert np.array_equal(fwn.calcSelfIntersections(2.0), fx["selfint"])
