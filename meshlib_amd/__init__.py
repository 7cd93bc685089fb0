"""MI355X (gfx950) distance-field path: host mirror of MeshLib's GPU interfaces.

Python surface of the C-ABI in ``include/meshlib_amd.h`` (``libmeshlib_amd.so``).  The
class and method names mirror the reference interfaces this path drops in behind:

* ``FastWindingNumber`` -- ``MR::IFastWindingNumber`` + ``MR::IFastWindingNumberByParts``
  (source/MRMesh/MRFastWindingNumber.h:11-107), GPU class ``MR::Cuda::FastWindingNumber``
  (source/MRCuda/MRCudaFastWindingNumber.h:17-38);
* ``PointsToMeshProjector`` -- ``MR::IPointsToMeshProjector``
  (source/MRMesh/MRPointsToMeshProjector.h:12-30), GPU class
  ``MR::Cuda::PointsToMeshProjector`` (source/MRCuda/MRCudaPointsToMeshProjector.h:9-31);
* ``DistanceToMeshOptions`` -- source/MRMesh/MRDistanceToMeshOptions.h:10-33.

Errors surface as ``MeshLibAmdError`` carrying the library's message (the reference returns
``Expected<void>`` strings).  There is no CPU fallback: without the built native library
every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

__all__ = ["Mesh", "FastWindingNumber", "PointsToMeshProjector", "DistanceToMeshOptions", "DistanceMLP",
           "MeshLibAmdError", "PROJ_DTYPE", "NODE_DTYPE", "DIPOLE_DTYPE", "lib", "lib_path",
           "set_memory_limit", "FLT_MAX"]

FLT_MAX = float(np.finfo(np.float32).max)
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmeshlib_amd.so")

PROJ_DTYPE = np.dtype([("face", "<i4"), ("proj", "<f4", 3), ("e", "<i4"), ("a", "<f4"),
                       ("b", "<f4"), ("distSq", "<f4")])
NODE_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("l", "<i4"), ("r", "<i4")])
DIPOLE_DTYPE = np.dtype([("pos", "<f4", 3), ("area", "<f4"), ("dirArea", "<f4", 3), ("rr", "<f4")])

MI_OK, MI_E_INVALID, MI_E_HIP, MI_E_CANCELED, MI_E_NOMEM, MI_E_LIMIT = 0, -1, -2, -3, -4, -5

# measurement counters of mi_mesh_set_stats: names of the slots and the bytes each event reads
STAT_NAMES = ("walks", "quad_steps", "prism_reads", "leaf_tests", "keyed_steps", "keyed_tests", "fwn_visits",
              "fwn_descents", "fwn_leaves", "seeds", "seed_levels")
STAT_STRIDE, STAT_LAUNCHES = 16, 3


class MeshLibAmdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class _DistOpts(C.Structure):
    _fields_ = [("minDistSq", C.c_float), ("maxDistSq", C.c_float), ("nullOutsideMinMax", C.c_int32),
                ("windingNumberThreshold", C.c_float), ("windingNumberBeta", C.c_float)]


_PROGRESS = C.CFUNCTYPE(C.c_int, C.c_float, C.c_void_p)
_PART = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32, C.c_void_p)

_lib = None


def lib_path() -> str:
    return _LIB_PATH


def _share_hip_runtime_with_torch():
    """One HIP runtime per process.  PyTorch-ROCm wheels bundle their own libamdhip64 (SONAME
    libamdhip64.so.7, loaded by file name from torch/lib); this library links libamdhip64.so.7
    from /opt/rocm.  Two copies in one process each initialise the HSA runtime and the second
    sees no GPU.  Loading torch's copy first (RTLD_GLOBAL) makes both resolve to it: this
    library's NEEDED entry matches the SONAME, and torch's own lookup finds the same file."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for base in spec.submodule_search_locations:
        cand = os.path.join(base, "lib", "libamdhip64.so")
        if os.path.exists(cand):
            C.CDLL(cand, mode=C.RTLD_GLOBAL)
            return


def lib():
    """Loads the native library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise MeshLibAmdError(MI_E_INVALID, f"native library missing: {_LIB_PATH} "
                              "(build it with `python -m meshlib_amd.build`)")
    _share_hip_runtime_with_torch()
    L = C.CDLL(_LIB_PATH)
    P, I32, I64, F, SZ = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_size_t
    sig = {
        "mi_last_error": (SZ, [C.c_char_p, SZ]),
        "mi_set_memory_limit": (C.c_int, [SZ]),
        "mi_set_traversal": (C.c_int, [C.c_int]),
        "mi_get_traversal": (C.c_int, []),
        "mi_set_leaf_filter": (C.c_int, [C.c_int]),
        "mi_get_leaf_filter": (C.c_int, []),
        "mi_set_tree_builder": (C.c_int, [C.c_int]),
        "mi_get_tree_builder": (C.c_int, []),
        "mi_set_split_budget": (C.c_int, [I32, I32]),
        "mi_get_split_budget": (I32, [I32]),
        "mi_mesh_create": (C.c_int, [P, I32, P, I32, C.POINTER(P)]),
        "mi_mesh_destroy": (None, [P]),
        "mi_mesh_num_nodes": (C.c_int, [P]),
        "mi_mesh_get_tree": (C.c_int, [P, P, P]),
        "mi_mesh_depth": (C.c_int, [P]),
        "mi_mesh_set_stats": (C.c_int, [P, P]),
        "mi_mesh_set_mode": (C.c_int, [P, I32, I32]),
        "mi_mesh_get_mode": (I32, [P, I32]),
        "mi_build_tree_host": (C.c_int, [P, I32, P, I32, P, P, P]),
        "mi_fwn_calc_from_vector": (C.c_int, [P, P, I64, F, I32, P, _PROGRESS, P]),
        "mi_fwn_calc_self_intersections": (C.c_int, [P, F, P, _PROGRESS, P]),
        "mi_fwn_calc_from_grid": (C.c_int, [P, P, P, F, P, _PROGRESS, P]),
        "mi_fwn_calc_from_grid_with_distances": (C.c_int, [P, P, P, P, P, _PROGRESS, P]),
        "mi_fwn_calc_from_grid_by_parts": (C.c_int, [P, _PART, P, P, P, F, I32, _PROGRESS, P]),
        "mi_fwn_calc_from_grid_with_distances_by_parts": (C.c_int, [P, _PART, P, P, P, P, I32, _PROGRESS, P]),
        "mi_fwn_vector_device": (C.c_int, [P, P, I64, F, I32, P, P]),
        "mi_fwn_grid_device": (C.c_int, [P, P, P, F, I32, I32, P, P]),
        "mi_sdf_grid_device": (C.c_int, [P, P, P, P, I32, I32, P, P]),
        "mi_proj_find_projections": (C.c_int, [P, P, I64, P, P, F, F, P]),
        "mi_proj_heap_bytes": (SZ, [SZ]),
        "mi_proj_find_projections_device": (C.c_int, [P, P, I64, P, P, F, F, P, P]),
        "mi_proj_find_projections_bounded_device": (C.c_int, [P, P, I64, P, P, P, F, F, P, P]),
        "mi_proj_find_projections_capped_device": (C.c_int, [P, P, I64, P, F, P, C.POINTER(I64), P]),
        "mi_sdf_grid_capped_device": (C.c_int, [P, P, P, P, P, I32, I32, P, C.POINTER(I64), P]),
        "mi_batch_create": (C.c_int, [C.POINTER(P)]),
        "mi_batch_destroy": (None, [P]),
        "mi_batch_run_device": (C.c_int, [P, P, I32, P, F, P]),
        "mi_batch_autotune": (C.c_int, [P, P, I32, P, F, P, C.POINTER(F)]),
        "mi_batch_set_config": (C.c_int, [P, I32, I32]),
        "mi_batch_get_config": (C.c_int, [P, C.POINTER(I32), C.POINTER(I32)]),
        "mi_batch_run_capped_device": (C.c_int, [P, P, I32, P, P, P, F, P, C.POINTER(I64)]),
        "mi_batch_run_learned_device": (C.c_int, [P, P, I32, P, F, P, F, P, C.POINTER(I64)]),
        "mi_mlp_create": (C.c_int, [P, F, C.c_uint64, C.POINTER(P)]),
        "mi_mlp_destroy": (None, [P]),
        "mi_mlp_num_params": (C.c_int, []),
        "mi_mlp_get_params": (C.c_int, [P, P]),
        "mi_mlp_set_params": (C.c_int, [P, P]),
        "mi_mlp_gradient_device": (C.c_int, [P, P, P, I64, P, P]),
        "mi_mlp_train_device": (C.c_int, [P, P, P, I64, I32, I32, F, P]),
        "mi_mlp_predict_device": (C.c_int, [P, P, I64, P, P]),
        "mi_mlp_calibrate": (C.c_int, [P, P, P, I64, F, C.POINTER(F), P]),
        "mi_mlp_set_slack": (C.c_int, [P, F]),
        "mi_proj_find_projections_learned_device": (C.c_int, [P, P, F, P, I64, F, P, C.POINTER(I64), P]),
        "mi_sdf_grid_learned_device": (C.c_int, [P, P, F, P, P, P, I32, I32, P, C.POINTER(I64), P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != MI_OK:
        buf = C.create_string_buffer(4096)
        lib().mi_last_error(buf, 4096)
        raise MeshLibAmdError(rc, buf.value.decode(errors="replace") or f"status {rc}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _xf(xf):
    if xf is None:
        return None
    a = np.ascontiguousarray(xf, dtype=np.float32).reshape(-1)
    if a.size == 16:      # 4x4 homogeneous matrix -> A rows + b
        m = a.reshape(4, 4)
        a = np.concatenate([m[:3, :3].reshape(-1), m[:3, 3]]).astype(np.float32)
    if a.size != 12:
        raise ValueError("affine transform must be float[12] (row-major A then b) or 4x4")
    return a


def _dims(d):
    a = np.ascontiguousarray(d, dtype=np.int32).reshape(3)
    return a


def build_tree_host(verts, tris):
    """Host-only tree + dipole build (no device): (nodes, dipoles, depth) in the reference layout."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    valid = int(np.count_nonzero((t >= 0).all(axis=1)))     # deleted faces ({-1,-1,-1}) get no leaf
    n = max(2 * valid - 1, 0)
    nodes = np.zeros(n, NODE_DTYPE)
    dips = np.zeros(n, DIPOLE_DTYPE)
    depth = C.c_int32(0)
    _check(lib().mi_build_tree_host(_p(v), v.shape[0], _p(t), t.shape[0], _p(nodes), _p(dips),
                                    C.byref(depth)))
    return nodes, dips, int(depth.value)


TRAVERSALS = {"per_lane": 0, "wave_coop": 1}


def set_traversal(mode: str):
    """Kernel traversal form: 'per_lane' (one walk per query, the reference's form) or
    'wave_coop' (one walk per 64-query wave; default).  Results are identical."""
    _check(lib().mi_set_traversal(TRAVERSALS[mode]))


def get_traversal() -> str:
    v = int(lib().mi_get_traversal())
    return {b: a for a, b in TRAVERSALS.items()}[v]


LEAF_FILTERS = {"off": 0, "on": 1, "auto": -1}


def set_leaf_filter(mode: str):
    """Leaf filter of the signed-distance walks: 'on', 'off' or 'auto' (default: grids of more than
    2^(depth - 8) waves of 64 voxels).  Results are identical."""
    _check(lib().mi_set_leaf_filter(LEAF_FILTERS[mode]))


def get_leaf_filter() -> str:
    v = int(lib().mi_get_leaf_filter())
    return {b: a for a, b in LEAF_FILTERS.items()}[v]


TREE_BUILDERS = {"device": 0, "host": 1}


def set_tree_builder(builder: str):
    """Where Mesh() builds the AABB tree and dipoles: 'device' (GPU, default) or 'host'.  Both give
    the reference's tree node-for-node and bit-for-bit."""
    _check(lib().mi_set_tree_builder(TREE_BUILDERS[builder]))


SPLIT_WALKS = {"sdf": 0, "projections": 1}


def set_split_budget(steps: int, walk: str = "both"):
    """Steps after which a wave hands its walk to the split pass (0 = never, -1 = automatic, the
    default), for the signed-distance walk ('sdf'), the projection walk ('projections') or both;
    results do not depend on it."""
    for w in (SPLIT_WALKS if walk == "both" else [walk]):
        _check(lib().mi_set_split_budget(SPLIT_WALKS[w], int(steps)))


def get_split_budget(walk: str = "sdf") -> int:
    return int(lib().mi_get_split_budget(SPLIT_WALKS[walk]))


def get_tree_builder() -> str:
    v = int(lib().mi_get_tree_builder())
    return {b: a for a, b in TREE_BUILDERS.items()}[v]


MESH_MODES = {"traversal": 0, "leaf_filter": 1, "split_sdf": 2, "split_projections": 3}
MODE_DEFAULT = -2


def _mode_value(mode: str, value) -> int:
    if value is None:
        return MODE_DEFAULT
    if mode == "traversal":
        return TRAVERSALS[value]
    if mode == "leaf_filter":
        return LEAF_FILTERS[value]
    return int(value)


def set_memory_limit(nbytes: int):
    """Caps the device-memory budget of the chunked entry points (0 = default policy)."""
    _check(lib().mi_set_memory_limit(int(nbytes)))


@dataclass
class DistanceToMeshOptions:
    minDistSq: float = 0.0
    maxDistSq: float = FLT_MAX
    nullOutsideMinMax: bool = True
    windingNumberThreshold: float = 0.5
    windingNumberBeta: float = 2.0

    def _c(self):
        return _DistOpts(self.minDistSq, self.maxDistSq, 1 if self.nullOutsideMinMax else 0,
                         self.windingNumberThreshold, self.windingNumberBeta)


def _progress(cb: Optional[Callable[[float], bool]]):
    if cb is None:
        return _PROGRESS(0), None
    def f(v, _u):
        try:
            return 1 if cb(float(v)) else 0
        except Exception:   # a raising callback cancels
            return 0
    c = _PROGRESS(f)
    return c, c


class Mesh:
    """Triangle mesh (float32 vertices, int32 triangles in getTriangulation order) with its
    device-resident AABB tree + dipoles (built on construction, like Mesh::getAABBTree())."""

    def __init__(self, verts, tris):
        self.verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        h = C.c_void_p()
        _check(lib().mi_mesh_create(_p(self.verts), self.verts.shape[0], _p(self.tris),
                                    self.tris.shape[0], C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().mi_mesh_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_faces(self) -> int:
        return int(self.tris.shape[0])

    def tree(self):
        """AABB tree nodes + dipoles in the reference layout (AABBTree::nodes(), getDipoles())."""
        n = lib().mi_mesh_num_nodes(self._h)
        nodes = np.zeros(n, NODE_DTYPE)
        dips = np.zeros(n, DIPOLE_DTYPE)
        _check(lib().mi_mesh_get_tree(self._h, _p(nodes), _p(dips)))
        return nodes, dips

    def depth(self) -> int:
        return int(lib().mi_mesh_depth(self._h))

    def set_mode(self, mode: str, value):
        """This mesh's own setting, overriding the process-wide one for calls on this mesh only:
        mode 'traversal' (value as set_traversal), 'leaf_filter' (as set_leaf_filter), 'split_sdf' or
        'split_projections' (steps, -1 automatic); value None follows the process-wide setting."""
        _check(lib().mi_mesh_set_mode(self._h, MESH_MODES[mode], _mode_value(mode, value)))

    def get_mode(self, mode: str):
        """This mesh's own setting (None when it follows the process-wide one)."""
        v = int(lib().mi_mesh_get_mode(self._h, MESH_MODES[mode]))
        if v == MODE_DEFAULT:
            return None
        names = {"traversal": TRAVERSALS, "leaf_filter": LEAF_FILTERS}.get(mode)
        return {b: a for a, b in names.items()}[v] if names else v

    def set_stats(self, d_counters: int):
        """Measurement hook: device pointer to MI_STAT_LAUNCHES * MI_STAT_STRIDE zeroed uint64 counters
        that later calls on this mesh add their walks' record reads to (0 = off)."""
        _check(lib().mi_mesh_set_stats(self._h, C.c_void_p(d_counters or None)))


class _BatchItem(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("dims", C.c_int32 * 3), ("xf", C.c_float * 12), ("d_sdf", C.c_void_p),
                ("d_pts", C.c_void_p), ("numPts", C.c_int64), ("d_proj", C.c_void_p)]


class Batch:
    """One persistent launch over many (mesh, grid, point set) items of uneven size, with the
    project's autotuner for resident workgroups per CU and per-XCD work partitioning."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().mi_batch_create(C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mi_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def _items(items):
        arr = (_BatchItem * len(items))()
        for i, it in enumerate(items):
            arr[i].mesh = it["mesh"].handle
            d = it.get("dims")
            if d is None:
                d = (0, 0, 0)
            for k in range(3):
                arr[i].dims[k] = int(d[k])
            x = _xf(it["xf"]) if it.get("xf") is not None else np.zeros(12, np.float32)
            for k in range(12):
                arr[i].xf[k] = float(x[k])
            arr[i].d_sdf = it.get("d_sdf") or None
            arr[i].d_pts = it.get("d_pts") or None
            arr[i].numPts = int(it.get("n", 0))
            arr[i].d_proj = it.get("d_proj") or None
        return arr

    def run(self, items, options: DistanceToMeshOptions = None, upDistLimitSq: float = FLT_MAX, stream: int = 0):
        arr = self._items(items)
        o = (options or DistanceToMeshOptions())._c()
        _check(lib().mi_batch_run_device(self._h, arr, len(items), C.byref(o), upDistLimitSq,
                                         C.c_void_p(stream or None)))

    def autotune(self, items, options: DistanceToMeshOptions = None, upDistLimitSq: float = FLT_MAX,
                 stream: int = 0) -> float:
        arr = self._items(items)
        o = (options or DistanceToMeshOptions())._c()
        ms = C.c_float(0)
        _check(lib().mi_batch_autotune(self._h, arr, len(items), C.byref(o), upDistLimitSq,
                                       C.c_void_p(stream or None), C.byref(ms)))
        return float(ms.value)

    def run_capped(self, items, cap_sdf=None, cap_proj=None, options: DistanceToMeshOptions = None,
                   upDistLimitSq: float = FLT_MAX, stream: int = 0) -> int:
        """Two persistent launches with per-query caps (device pointers per item, 0/None = no cap);
        returns the number of rerun queries.  Outputs equal run()'s."""
        arr = self._items(items)
        n = len(items)

        def ptrs(lst):
            if lst is None:
                return None
            a = (C.c_void_p * n)()
            for i, x in enumerate(lst):
                a[i] = x or None
            return a
        cs, cp = ptrs(cap_sdf), ptrs(cap_proj)
        o = (options or DistanceToMeshOptions())._c()
        r = C.c_int64(0)
        _check(lib().mi_batch_run_capped_device(self._h, arr, n, cs, cp, C.byref(o), upDistLimitSq,
                                                C.c_void_p(stream or None), C.byref(r)))
        return int(r.value)

    def run_learned(self, items, nets, margin: float, options: DistanceToMeshOptions = None,
                    upDistLimitSq: float = FLT_MAX, stream: int = 0) -> int:
        """The batch with caps (pred * (1 + margin))^2 from each item's DistanceMLP (None = no cap);
        returns the number of rerun queries.  Outputs equal run()'s."""
        arr = self._items(items)
        n = len(items)
        hs = (C.c_void_p * n)()
        for i, net in enumerate(nets):
            hs[i] = net.handle.value if net is not None else None
        o = (options or DistanceToMeshOptions())._c()
        r = C.c_int64(0)
        _check(lib().mi_batch_run_learned_device(self._h, arr, n, hs, margin, C.byref(o), upDistLimitSq,
                                                 C.c_void_p(stream or None), C.byref(r)))
        return int(r.value)

    def config(self):
        a, b = C.c_int32(0), C.c_int32(0)
        _check(lib().mi_batch_get_config(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def set_config(self, blocks_per_cu: int, xcd_split: int):
        _check(lib().mi_batch_set_config(self._h, blocks_per_cu, xcd_split))


class DistanceMLP:
    """The learned distance bound: Fourier-feature MLP (99 -> 4 x 128 SiLU -> softplus) trained and
    evaluated on the GPU (BF16 MFMA kernels, FP32 accumulation).  Device arrays are torch tensors."""

    def __init__(self, center, scale: float, seed: int = 0):
        c = np.ascontiguousarray(center, dtype=np.float32).reshape(3)
        h = C.c_void_p()
        _check(lib().mi_mlp_create(_p(c), float(scale), int(seed), C.byref(h)))
        self._h = h
        self.center, self.scale = c, float(scale)
        self.num_params = int(lib().mi_mlp_num_params())

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().mi_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _check(lib().mi_mlp_get_params(self._h, _p(out)))
        return out

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32).reshape(-1)
        assert th.size == self.num_params
        _check(lib().mi_mlp_set_params(self._h, _p(th)))

    def predict_device(self, d_pts, d_out, stream: int = 0):
        """Predicted distances of the device points (float32 [n, 3]) into d_out (float32 [n], device),
        asynchronous on `stream`."""
        _check(lib().mi_mlp_predict_device(self._h, C.c_void_p(d_pts.data_ptr()), int(d_pts.shape[0]),
                                           C.c_void_p(d_out.data_ptr()), C.c_void_p(stream or None)))

    def predict(self, d_pts, stream: int = 0):
        import torch
        out = torch.empty(d_pts.shape[0], dtype=torch.float32, device=d_pts.device)
        self.predict_device(d_pts, out, stream)
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def gradient(self, d_pts, d_labels, stream: int = 0):
        import torch
        g = torch.zeros(self.num_params + 256, dtype=torch.float32, device=d_pts.device)
        _check(lib().mi_mlp_gradient_device(self._h, C.c_void_p(d_pts.data_ptr()), C.c_void_p(d_labels.data_ptr()),
                                            d_pts.shape[0], C.c_void_p(g.data_ptr()), C.c_void_p(stream or None)))
        torch.cuda.synchronize()
        return g.cpu().numpy()[:self.num_params]

    def loss(self, d_pts, d_labels) -> float:
        pred = self.predict(d_pts)
        lab = d_labels.cpu().numpy()
        d = (pred - lab) / self.scale
        return float(0.5 * np.mean(d.astype(np.float64) ** 2))

    def calibrate(self, d_pts, d_labels, quantile: float = 0.99, stream: int = 0) -> float:
        """Sets the slack added to every capped radius: the `quantile` of (label - prediction) over these
        labelled device points (float32 [n, 3] and [n]); returns it."""
        sl = C.c_float()
        _check(lib().mi_mlp_calibrate(self._h, C.c_void_p(d_pts.data_ptr()), C.c_void_p(d_labels.data_ptr()),
                                      int(d_labels.shape[0]), float(quantile), C.byref(sl), C.c_void_p(stream or None)))
        return float(sl.value)

    def set_slack(self, slack: float):
        _check(lib().mi_mlp_set_slack(self._h, float(slack)))

    def train(self, d_pts, d_labels, steps: int, batch: int, lr: float, stream: int = 0):
        _check(lib().mi_mlp_train_device(self._h, C.c_void_p(d_pts.data_ptr()), C.c_void_p(d_labels.data_ptr()),
                                         d_pts.shape[0], steps, batch, lr, C.c_void_p(stream or None)))


class FastWindingNumber:
    """``MR::IFastWindingNumber`` / ``IFastWindingNumberByParts`` on the GPU."""

    def __init__(self, mesh: Mesh):
        self.mesh = mesh

    def calcFromVector(self, points, beta: float, skipFace: int = -1, cb=None) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        res = np.empty(pts.shape[0], np.float32)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_from_vector(self.mesh.handle, _p(pts), pts.shape[0], beta, skipFace,
                                             _p(res), c, None))
        return res

    def calcSelfIntersections(self, beta: float, cb=None) -> np.ndarray:
        res = np.zeros(self.mesh.num_faces(), np.uint8)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_self_intersections(self.mesh.handle, beta, _p(res), c, None))
        return res.astype(bool)

    def calcFromGrid(self, dims, gridToMeshXf, beta: float, cb=None) -> np.ndarray:
        d = _dims(dims)
        res = np.empty(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_from_grid(self.mesh.handle, _p(d), _p(_xf(gridToMeshXf)), beta,
                                           _p(res), c, None))
        return res

    def calcFromGridWithDistances(self, dims, gridToMeshXf, options: DistanceToMeshOptions = None,
                                  cb=None) -> np.ndarray:
        d = _dims(dims)
        o = (options or DistanceToMeshOptions())._c()
        res = np.empty(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_from_grid_with_distances(self.mesh.handle, _p(d), _p(_xf(gridToMeshXf)),
                                                          C.byref(o), _p(res), c, None))
        return res

    @staticmethod
    def _part_cb(resFunc):
        def f(data, dims, zoff, _u):
            try:
                dx, dy, dz = dims[0], dims[1], dims[2]
                arr = np.ctypeslib.as_array(data, shape=(dx * dy * dz,)).copy()
                r = resFunc(arr, (dx, dy, dz), int(zoff))
                return MI_OK if (r is None or r is True) else MI_E_CANCELED
            except Exception:
                return MI_E_CANCELED
        return _PART(f)

    def calcFromGridByParts(self, resFunc, dims, gridToMeshXf, beta: float, layerOverlap: int,
                            cb=None):
        d = _dims(dims)
        part = self._part_cb(resFunc)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_from_grid_by_parts(self.mesh.handle, part, None, _p(d),
                                                    _p(_xf(gridToMeshXf)), beta, layerOverlap, c, None))

    def calcFromGridWithDistancesByParts(self, resFunc, dims, gridToMeshXf,
                                         options: DistanceToMeshOptions, layerOverlap: int, cb=None):
        d = _dims(dims)
        o = (options or DistanceToMeshOptions())._c()
        part = self._part_cb(resFunc)
        c, _keep = _progress(cb)
        _check(lib().mi_fwn_calc_from_grid_with_distances_by_parts(
            self.mesh.handle, part, None, _p(d), _p(_xf(gridToMeshXf)), C.byref(o), layerOverlap, c, None))

    # device-resident forms (pointers to HBM, e.g. torch tensors' data_ptr())
    def sdf_grid_device(self, dims, xf, options, d_out: int, zBegin=0, zEnd=None, stream: int = 0):
        d = _dims(dims)
        o = (options or DistanceToMeshOptions())._c()
        zEnd = int(d[2]) if zEnd is None else zEnd
        _check(lib().mi_sdf_grid_device(self.mesh.handle, _p(d), _p(_xf(xf)), C.byref(o), zBegin, zEnd,
                                        C.c_void_p(d_out), C.c_void_p(stream or None)))

    def fwn_grid_device(self, dims, xf, beta, d_out: int, zBegin=0, zEnd=None, stream: int = 0):
        d = _dims(dims)
        zEnd = int(d[2]) if zEnd is None else zEnd
        _check(lib().mi_fwn_grid_device(self.mesh.handle, _p(d), _p(_xf(xf)), beta, zBegin, zEnd,
                                        C.c_void_p(d_out), C.c_void_p(stream or None)))

    def sdf_grid_capped_device(self, dims, xf, options, d_capSq: int, d_out: int, zBegin=0, zEnd=None,
                               stream: int = 0) -> int:
        """Exact signed-distance slab with per-voxel caps on the search radius (squared); returns how
        many voxels needed the uncapped second pass."""
        d = _dims(dims)
        o = (options or DistanceToMeshOptions())._c()
        zEnd = int(d[2]) if zEnd is None else zEnd
        r = C.c_int64(0)
        _check(lib().mi_sdf_grid_capped_device(self.mesh.handle, _p(d), _p(_xf(xf)), C.byref(o),
                                               C.c_void_p(d_capSq), zBegin, zEnd, C.c_void_p(d_out),
                                               C.byref(r), C.c_void_p(stream or None)))
        return int(r.value)

    def sdf_grid_learned_device(self, net: "DistanceMLP", margin: float, dims, xf, options, d_out: int,
                                zBegin=0, zEnd=None, stream: int = 0) -> int:
        """Exact signed-distance slab with the learned cap (pred * (1 + margin))^2; returns reruns."""
        d = _dims(dims)
        o = (options or DistanceToMeshOptions())._c()
        zEnd = int(d[2]) if zEnd is None else zEnd
        r = C.c_int64(0)
        _check(lib().mi_sdf_grid_learned_device(self.mesh.handle, net.handle, margin, _p(d), _p(_xf(xf)),
                                                C.byref(o), zBegin, zEnd, C.c_void_p(d_out), C.byref(r),
                                                C.c_void_p(stream or None)))
        return int(r.value)

    def fwn_vector_device(self, d_pts: int, n: int, beta: float, d_out: int, skipFace: int = -1,
                          stream: int = 0):
        _check(lib().mi_fwn_vector_device(self.mesh.handle, C.c_void_p(d_pts), n, beta, skipFace,
                                          C.c_void_p(d_out), C.c_void_p(stream or None)))


class PointsToMeshProjector:
    """``MR::IPointsToMeshProjector`` on the GPU."""

    def __init__(self):
        self.mesh: Optional[Mesh] = None

    def updateMeshData(self, mesh: Optional[Mesh]):
        self.mesh = mesh

    def findProjections(self, points, objXf=None, refObjXf=None, upDistLimitSq: float = FLT_MAX,
                        loDistLimitSq: float = 0.0) -> np.ndarray:
        if self.mesh is None:
            raise MeshLibAmdError(MI_E_INVALID, "updateMeshData was not called")
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        res = np.zeros(pts.shape[0], PROJ_DTYPE)
        _check(lib().mi_proj_find_projections(self.mesh.handle, _p(pts), pts.shape[0], _p(_xf(objXf)),
                                              _p(_xf(refObjXf)), upDistLimitSq, loDistLimitSq, _p(res)))
        return res

    def find_projections_bounded_device(self, d_pts: int, n: int, d_upSq: int, d_out: int, ptXf=None,
                                        treeXf=None, upDistLimitSq: float = FLT_MAX,
                                        loDistLimitSq: float = 0.0, stream: int = 0):
        """findProjection with a per-point upDistLimitSq (ICP's limitedProjector pattern)."""
        _check(lib().mi_proj_find_projections_bounded_device(
            self.mesh.handle, C.c_void_p(d_pts), n, C.c_void_p(d_upSq), _p(_xf(ptXf)), _p(_xf(treeXf)),
            upDistLimitSq, loDistLimitSq, C.c_void_p(d_out), C.c_void_p(stream or None)))

    def find_projections_capped_device(self, d_pts: int, n: int, d_capSq: int, d_out: int,
                                       upDistLimitSq: float = FLT_MAX, stream: int = 0) -> int:
        """Exact projections with per-point caps on the search radius (squared); returns how many
        points needed the uncapped second pass."""
        r = C.c_int64(0)
        _check(lib().mi_proj_find_projections_capped_device(
            self.mesh.handle, C.c_void_p(d_pts), n, C.c_void_p(d_capSq), upDistLimitSq, C.c_void_p(d_out),
            C.byref(r), C.c_void_p(stream or None)))
        return int(r.value)

    def find_projections_learned_device(self, net: "DistanceMLP", margin: float, d_pts: int, n: int,
                                        d_out: int, upDistLimitSq: float = FLT_MAX, stream: int = 0) -> int:
        """Exact projections with the learned cap (pred * (1 + margin))^2; returns how many points
        were searched again without the cap."""
        r = C.c_int64(0)
        _check(lib().mi_proj_find_projections_learned_device(self.mesh.handle, net.handle, margin,
                                                             C.c_void_p(d_pts), n, upDistLimitSq,
                                                             C.c_void_p(d_out), C.byref(r),
                                                             C.c_void_p(stream or None)))
        return int(r.value)

    def projectionsHeapBytes(self, numProjections: int) -> int:
        return int(lib().mi_proj_heap_bytes(numProjections))

    def find_projections_device(self, d_pts: int, n: int, d_out: int, ptXf=None, treeXf=None,
                                upDistLimitSq: float = FLT_MAX, loDistLimitSq: float = 0.0,
                                stream: int = 0):
        if self.mesh is None:
            raise MeshLibAmdError(MI_E_INVALID, "updateMeshData was not called")
        _check(lib().mi_proj_find_projections_device(self.mesh.handle, C.c_void_p(d_pts), n,
                                                     _p(_xf(ptXf)), _p(_xf(treeXf)), upDistLimitSq,
                                                     loDistLimitSq, C.c_void_p(d_out),
                                                     C.c_void_p(stream or None)))
