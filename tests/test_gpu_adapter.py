"""The C++ drop-in classes (include/meshlib_amd/MRAmdDistanceField.h), driven by
tests/cpp/adapter_main.cpp the way MeshLib's callers use MR::Cuda::FastWindingNumber and
MR::Cuda::PointsToMeshProjector, checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from meshlib_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "adapter_main")


def test_adapter_driver_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_cpp_adapter_matches_oracle(tmp_path):
    v, f = synth.bumpy_icosphere(4, 8, 9)
    dims, xf, vox = synth.grid_for_mesh(v, (20, 18, 22))
    pts = synth.projection_queries(v, f, 5000, 2)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([v.shape[0], f.shape[0], pts.shape[0], *dims], np.int32).tofile(fh)
        v.astype(np.float32).tofile(fh)
        f.astype(np.int32).tofile(fh)
        pts.astype(np.float32).tofile(fh)
        xf.astype(np.float32).tofile(fh)
    env = dict(os.environ)
    r = subprocess.run([BIN, str(inp), str(out)], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    nvox = int(dims.prod())
    raw = open(out, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    sdf = take(np.float32, nvox)
    sdf_parts = take(np.float32, nvox)
    wn = take(np.float32, pts.shape[0])
    si = take(np.uint8, f.shape[0])
    proj = take(np.uint8, pts.shape[0] * 32).reshape(-1, 32)
    canceled = take(np.int32, 1)[0]
    om = O.Mesh(v, f)
    o_sdf = om.sdf_grid(dims, xf)
    assert np.array_equal(np.abs(sdf), np.abs(o_sdf))
    far = np.abs(o_sdf) > 1e-3 * float(vox.min())
    assert np.array_equal(np.sign(sdf[far]), np.sign(o_sdf[far]))
    assert np.array_equal(sdf_parts.view(np.uint32), sdf.view(np.uint32))
    assert np.max(np.abs(wn - om.fwn_vector(pts))) <= 1e-4
    assert np.array_equal(si.astype(bool), om.self_intersections())
    assert np.array_equal(proj, om.find_projections(pts).view(np.uint8).reshape(-1, 32))
    assert canceled == 1
