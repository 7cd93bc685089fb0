"""The C-ABI library loads, exports every symbol include/meshlib_amd.h declares, and its host-only
paths (argument validation, host tree build) behave without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import meshlib_amd as M
from meshlib_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", txt))
    names -= {"mi_progress_fn", "mi_grid_part_fn"}
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(M.lib_path())
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_arguments_are_reported():
    L = M.lib()
    h = ctypes.c_void_p()
    rc = L.mi_mesh_create(None, -1, None, 0, ctypes.byref(h))
    assert rc == M.MI_E_INVALID
    buf = ctypes.create_string_buffer(256)
    assert L.mi_last_error(buf, 256) > 0
    assert b"invalid" in buf.value
    v = np.zeros((3, 3), np.float32)
    t = np.array([[0, 1, 5]], np.int32)          # index out of range
    with pytest.raises(M.MeshLibAmdError):
        M.build_tree_host(v, t)


def test_host_tree_empty_and_single():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nodes, dips, depth = M.build_tree_host(v, np.zeros((0, 3), np.int32))
    assert len(nodes) == 0 and depth == 0
    nodes, dips, depth = M.build_tree_host(v, np.array([[0, 1, 2]], np.int32))
    assert len(nodes) == 1 and nodes[0]["r"] == -1 and nodes[0]["l"] == 0 and depth == 0
    assert abs(dips[0]["area"] - 0.5) < 1e-7


def test_host_tree_depth_bound():
    v, f = synth.bumpy_icosphere(6, 8, 1)
    nodes, dips, depth = M.build_tree_host(v, f)
    assert depth == int(np.ceil(np.log2(f.shape[0])))
    assert M.lib().mi_proj_heap_bytes(1000) >= 1000 * 44


def test_tree_builder_setting():
    assert M.get_tree_builder() == "device"
    M.set_tree_builder("host")
    assert M.get_tree_builder() == "host"
    M.set_tree_builder("device")
    assert M.lib().mi_set_tree_builder(7) == M.MI_E_INVALID
    assert M.get_tree_builder() == "device"


def test_split_budget_setting():
    # automatic by default; explicit steps, 0 (off) and automatic again are accepted per walk
    assert M.get_split_budget("sdf") == -1 and M.get_split_budget("projections") == -1
    M.set_split_budget(4096, "sdf")
    M.set_split_budget(0, "projections")
    assert M.get_split_budget("sdf") == 4096 and M.get_split_budget("projections") == 0
    assert M.lib().mi_set_split_budget(0, -2) == M.MI_E_INVALID
    assert M.lib().mi_set_split_budget(5, 16) == M.MI_E_INVALID
    M.set_split_budget(-1)
    assert M.get_split_budget("sdf") == -1 and M.get_split_budget("projections") == -1


def test_leaf_filter_setting():
    # automatic by default; on, off and automatic again are accepted, anything else is rejected
    assert M.get_leaf_filter() == "auto"
    M.set_leaf_filter("off")
    assert M.get_leaf_filter() == "off"
    M.set_leaf_filter("on")
    assert M.get_leaf_filter() == "on"
    assert M.lib().mi_set_leaf_filter(2) == M.MI_E_INVALID
    assert M.get_leaf_filter() == "on"
    M.set_leaf_filter("auto")
    assert M.get_leaf_filter() == "auto"
