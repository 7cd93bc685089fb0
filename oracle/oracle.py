"""ctypes front end of the CPU oracle (``liboracle.so``, built from ``oracle.c``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module.  See ``oracle.h`` for the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NODE_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("l", "<i4"), ("r", "<i4")])
DIPOLE_DTYPE = np.dtype([("pos", "<f4", 3), ("area", "<f4"), ("dirArea", "<f4", 3), ("rr", "<f4")])
PROJ_DTYPE = np.dtype([("face", "<i4"), ("proj", "<f4", 3), ("e", "<i4"), ("a", "<f4"),
                       ("b", "<f4"), ("distSq", "<f4")])
assert NODE_DTYPE.itemsize == 32 and DIPOLE_DTYPE.itemsize == 32 and PROJ_DTYPE.itemsize == 32


class DistOpts(C.Structure):
    _fields_ = [("minDistSq", C.c_float), ("maxDistSq", C.c_float),
                ("nullOutsideMinMax", C.c_int32), ("threshold", C.c_float), ("beta", C.c_float)]


class Counters(C.Structure):
    _fields_ = [("queries", C.c_int64), ("dist_probes", C.c_int64), ("dist_internal", C.c_int64),
                ("dist_leaves", C.c_int64), ("fwn_visits", C.c_int64), ("fwn_leaves", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.or_mesh_create.restype = P
        L.or_mesh_create.argtypes = [P, C.c_int32, P, C.c_int32]
        L.or_mesh_destroy.argtypes = [P]
        L.or_mesh_num_nodes.restype = C.c_int32
        L.or_mesh_num_nodes.argtypes = [P]
        L.or_mesh_get_nodes.argtypes = [P, P]
        L.or_mesh_get_dipoles.argtypes = [P, P]
        L.or_triangle_solid_angle.restype = C.c_float
        L.or_triangle_solid_angle.argtypes = [P, P, P, P]
        L.or_closest_point_in_triangle_d.argtypes = [P, P, P, P, P, P]
        L.or_find_projections.argtypes = [P, P, C.c_int64, P, P, C.c_float, C.c_float, P, C.c_int32, P]
        L.or_find_projections_bounded.argtypes = [P, P, C.c_int64, P, C.c_float, P, C.c_int32, P]
        L.or_fwn_vector.argtypes = [P, P, C.c_int64, C.c_float, C.c_int32, P, C.c_int32, P]
        L.or_fwn_grid.argtypes = [P, P, P, C.c_float, P, C.c_int32, P]
        L.or_sdf_grid.argtypes = [P, P, P, P, P, C.c_int32, P]
        L.or_sdf_grid_sample.argtypes = [P, P, P, P, P, C.c_int64, P, C.c_int32, P]
        L.or_self_intersections.argtypes = [P, C.c_float, P, C.c_int32]
        L.or_find_projections_f64.argtypes = [P, P, C.c_int64, P, P, C.c_int32]
        L.or_fwn_vector_f64.argtypes = [P, P, C.c_int64, C.c_double, P, C.c_int32]
        L.or_bf_nearest_f64.argtypes = [P, P, C.c_int32, P, C.c_int64, P, C.c_int32]
        L.or_bf_winding_f64.argtypes = [P, P, C.c_int32, P, C.c_int64, P, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape_last=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape_last is not None:
        a = a.reshape(-1, shape_last)
    return a


def _xf(xf):
    return None if xf is None else np.ascontiguousarray(xf, dtype=np.float32).reshape(12)


def triangle_solid_angle(p, a, b, c) -> float:
    arr = [np.ascontiguousarray(x, dtype=np.float32).reshape(3) for x in (p, a, b, c)]
    return float(lib().or_triangle_solid_angle(*[_p(x) for x in arr]))


def closest_point_in_triangle_d(p, a, b, c):
    arr = [np.ascontiguousarray(x, dtype=np.float64).reshape(3) for x in (p, a, b, c)]
    proj = np.zeros(3, np.float64)
    bary = np.zeros(2, np.float64)
    lib().or_closest_point_in_triangle_d(*[_p(x) for x in arr], _p(proj), _p(bary))
    return proj, bary


def default_opts(beta=2.0, minDistSq=0.0, maxDistSq=float(np.finfo(np.float32).max),
                 nullOutsideMinMax=True, threshold=0.5):
    return DistOpts(minDistSq, maxDistSq, 1 if nullOutsideMinMax else 0, threshold, beta)


class Mesh:
    """Oracle mesh: restated AABB tree + dipoles over (verts, tris)."""

    def __init__(self, verts, tris):
        self.verts = _f32(verts, 3)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        self._h = lib().or_mesh_create(_p(self.verts), self.verts.shape[0], _p(self.tris),
                                       self.tris.shape[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_mesh_destroy(h)
            self._h = None

    def nodes(self):
        n = lib().or_mesh_num_nodes(self._h)
        out = np.zeros(n, NODE_DTYPE)
        lib().or_mesh_get_nodes(self._h, _p(out))
        return out

    def dipoles(self):
        n = lib().or_mesh_num_nodes(self._h)
        out = np.zeros(n, DIPOLE_DTYPE)
        lib().or_mesh_get_dipoles(self._h, _p(out))
        return out

    def find_projections(self, pts, pt_xf=None, tree_xf=None, up=float(np.finfo(np.float32).max),
                         lo=0.0, nthreads=0, counters=None):
        pts = _f32(pts, 3)
        out = np.zeros(pts.shape[0], PROJ_DTYPE)
        cnt = counters if counters is not None else None
        lib().or_find_projections(self._h, _p(pts), pts.shape[0], _p(_xf(pt_xf)), _p(_xf(tree_xf)),
                                  up, lo, _p(out), nthreads, C.byref(cnt) if cnt is not None else None)
        return out

    def find_projections_bounded(self, pts, up_sq, lo=0.0, nthreads=0, counters=None):
        pts = _f32(pts, 3)
        ups = np.ascontiguousarray(up_sq, dtype=np.float32).reshape(-1)
        out = np.zeros(pts.shape[0], PROJ_DTYPE)
        lib().or_find_projections_bounded(self._h, _p(pts), pts.shape[0], _p(ups), lo, _p(out),
                                          nthreads, C.byref(counters) if counters is not None else None)
        return out

    def fwn_vector(self, pts, beta=2.0, skip_face=-1, nthreads=0, counters=None):
        pts = _f32(pts, 3)
        out = np.zeros(pts.shape[0], np.float32)
        lib().or_fwn_vector(self._h, _p(pts), pts.shape[0], beta, skip_face, _p(out), nthreads,
                            C.byref(counters) if counters is not None else None)
        return out

    def fwn_grid(self, dims, xf, beta=2.0, nthreads=0, counters=None):
        d = np.ascontiguousarray(dims, dtype=np.int32).reshape(3)
        out = np.zeros(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
        lib().or_fwn_grid(self._h, _p(d), _p(_xf(xf)), beta, _p(out), nthreads,
                          C.byref(counters) if counters is not None else None)
        return out

    def sdf_grid(self, dims, xf, opts=None, nthreads=0, counters=None):
        d = np.ascontiguousarray(dims, dtype=np.int32).reshape(3)
        o = opts if opts is not None else default_opts()
        out = np.zeros(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
        lib().or_sdf_grid(self._h, _p(d), _p(_xf(xf)), C.byref(o), _p(out), nthreads,
                          C.byref(counters) if counters is not None else None)
        return out

    def sdf_grid_sample(self, dims, xf, ids, opts=None, nthreads=0, counters=None):
        d = np.ascontiguousarray(dims, dtype=np.int32).reshape(3)
        o = opts if opts is not None else default_opts()
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        out = np.zeros(ids.shape[0], np.float32)
        lib().or_sdf_grid_sample(self._h, _p(d), _p(_xf(xf)), C.byref(o), _p(ids), ids.shape[0],
                                 _p(out), nthreads, C.byref(counters) if counters is not None else None)
        return out

    def self_intersections(self, beta=2.0, nthreads=0):
        out = np.zeros(self.tris.shape[0], np.uint8)
        lib().or_self_intersections(self._h, beta, _p(out), nthreads)
        return out.astype(bool)

    def find_projections_f64(self, pts, nthreads=0):
        pts = _f32(pts, 3)
        d = np.zeros(pts.shape[0], np.float64)
        f = np.zeros(pts.shape[0], np.int32)
        lib().or_find_projections_f64(self._h, _p(pts), pts.shape[0], _p(d), _p(f), nthreads)
        return d, f

    def fwn_vector_f64(self, pts, beta=2.0, nthreads=0):
        pts = _f32(pts, 3)
        out = np.zeros(pts.shape[0], np.float64)
        lib().or_fwn_vector_f64(self._h, _p(pts), pts.shape[0], beta, _p(out), nthreads)
        return out


def bf_nearest_f64(verts, tris, pts, nthreads=0):
    v = _f32(verts, 3)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    p = _f32(pts, 3)
    out = np.zeros(p.shape[0], np.float64)
    lib().or_bf_nearest_f64(_p(v), _p(t), t.shape[0], _p(p), p.shape[0], _p(out), nthreads)
    return out


def bf_winding_f64(verts, tris, pts, nthreads=0):
    v = _f32(verts, 3)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    p = _f32(pts, 3)
    out = np.zeros(p.shape[0], np.float64)
    lib().or_bf_winding_f64(_p(v), _p(t), t.shape[0], _p(p), p.shape[0], _p(out), nthreads)
    return out


def grid_points(dims, xf):
    """Voxel-centre points of the grid in x-fastest order, computed in float32 with the
    reference's operation order (VolumeIndexer::toPos then AffineXf3f)."""
    dx, dy, dz = (int(v) for v in dims)
    xf = np.asarray(xf, np.float32)
    z, y, x = np.meshgrid(np.arange(dz, dtype=np.float32), np.arange(dy, dtype=np.float32),
                          np.arange(dx, dtype=np.float32), indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    A = xf[:9].reshape(3, 3)
    out = np.empty_like(p)
    for i in range(3):
        r = (A[i, 0] * p[:, 0] + A[i, 1] * p[:, 1]) + A[i, 2] * p[:, 2]
        out[:, i] = r + xf[9 + i]
    return out
