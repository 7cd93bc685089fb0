"""The projection entry point that allocates nothing (mi_proj_find_projections_noalloc_device), called
through the C-ABI: its records equal mi_proj_find_projections_device's byte for byte, on two calls in a
row (the second smaller, so the fixed scratch holds the first call's contents), on a shallow and a deep
tree (the deep one runs the split form out of the same scratch), with and without the band and point
transform options; the oracle checks a sample; sizes beyond MI_NOALLOC_MAX_POINTS are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bind():
    L = M.lib()
    fn = L.mi_proj_find_projections_noalloc_device
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p,
                   C.c_void_p]
    return fn


def _xf_arg(xf):
    return None if xf is None else xf.ctypes.data


@pytest.mark.parametrize("level,bumps,sizes", [(5, 8, (300000, 4097)), (8, 32, (MAX_POINTS, 777))])
def test_noalloc_equals_device_entry(level, bumps, sizes):
    import torch
    noalloc = _bind()
    L = M.lib()
    v, f = synth.bumpy_icosphere(level, bumps, 6)
    mesh = M.Mesh(v, f)
    om = O.Mesh(v, f) if level <= 5 else None
    stream = torch.cuda.current_stream().cuda_stream
    rot = np.array([0.96, -0.28, 0, 0, 0.28, 0.96, 0, 0, 0, 0, 1, 0.01], np.float32)
    for call, n in enumerate(sizes):
        pts = synth.projection_queries(v, f, n, 40 + call)
        dp = torch.from_numpy(pts).cuda()
        for up, lo, xf in ((M.FLT_MAX, 0.0, None), (0.05, 1e-4, rot)):
            say(f"k={level} call {call} n={n} up={up} lo={lo} xf={'yes' if xf is not None else 'no'}")
            want = torch.full((n * 8,), -9, dtype=torch.int32, device="cuda")
            got = torch.full((n * 8,), -7, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            rc = L.mi_proj_find_projections_device(mesh.handle, dp.data_ptr(), n, _xf_arg(xf), None, up, lo,
                                                   want.data_ptr(), C.c_void_p(stream))
            assert rc == M.MI_OK
            rc = noalloc(mesh.handle, dp.data_ptr(), n, _xf_arg(xf), None, up, lo, got.data_ptr(), C.c_void_p(stream))
            assert rc == M.MI_OK
            torch.cuda.synchronize()
            w, g = want.cpu().numpy(), got.cpu().numpy()
            assert np.array_equal(g, w)
            if om is not None and xf is None:
                sel = np.random.default_rng(call).choice(n, 3000, replace=False)
                ref = om.find_projections(pts[sel], up=up, lo=lo)
                assert np.array_equal(g.reshape(-1, 8)[sel].view(np.uint8).ravel(), ref.view(np.uint8).ravel())


def test_noalloc_limits():
    import torch
    noalloc = _bind()
    v, f = synth.bumpy_icosphere(2, 2, 1)
    mesh = M.Mesh(v, f)
    # refused before any device work: more points than the fixed scratch holds, null arguments
    assert noalloc(mesh.handle, 16, MAX_POINTS + 1, None, None, M.FLT_MAX, 0.0, 16, None) == M.MI_E_LIMIT
    assert noalloc(mesh.handle, None, 5, None, None, M.FLT_MAX, 0.0, None, None) == M.MI_E_INVALID
    assert noalloc(None, None, 0, None, None, M.FLT_MAX, 0.0, None, None) == M.MI_E_INVALID
    assert noalloc(mesh.handle, None, 0, None, None, M.FLT_MAX, 0.0, None, None) == M.MI_OK
    torch.cuda.synchronize()
