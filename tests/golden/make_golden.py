"""Writes the golden fixtures of the distance-field path (run from the repo root:
``python tests/golden/make_golden.py``).

Each fixture holds its inputs (mesh, queries, grid) and the CPU oracle's outputs: projection
records (32 B MeshProjectionResult), a signed-distance grid, winding numbers at points and on a
grid, self-intersection flags, plus brute-force FP64 nearest distances and exact winding numbers
that pin the oracle independently.  The reference ships no golden vectors for this path (its
asset-based regression tests need test_data/, which is empty), so these fixtures are the
project's own regression pins.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from meshlib_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, level, bumps, seed, grid dims, n queries
    ("ico2", 2, 8, 101, (16, 16, 16), 2048),
    ("ico3", 3, 8, 102, (16, 12, 20), 4096),
]


def make(name, level, bumps, seed, dims, nq):
    v, f = synth.bumpy_icosphere(level, bumps, seed)
    m = O.Mesh(v, f)
    dims, xf, vox = synth.grid_for_mesh(v, dims)
    pts = synth.projection_queries(v, f, nq, seed + 1)
    proj = m.find_projections(pts)
    sdf = m.sdf_grid(dims, xf)
    fwn_pts = m.fwn_vector(pts, beta=2.0)
    fwn_grid = m.fwn_grid(dims, xf, beta=2.0)
    selfint = m.self_intersections(2.0)
    bf_d = O.bf_nearest_f64(v, f, pts)
    bf_w = O.bf_winding_f64(v, f, pts)
    nodes = m.nodes()
    out = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(out, verts=v, tris=f, dims=dims, xf=xf, pts=pts,
                        proj=proj.view(np.uint8).reshape(-1, 32), sdf=sdf, fwn_pts=fwn_pts,
                        fwn_grid=fwn_grid, selfint=selfint, bf_dist_sq=bf_d, bf_winding=bf_w,
                        nodes=nodes.view(np.uint8).reshape(-1, 32))
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    for c in CASES:
        make(*c)
