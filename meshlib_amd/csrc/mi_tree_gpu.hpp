// mi_tree_gpu.hpp -- the AABB tree, its dipoles and the flattened traversal layout built by gfx950
// kernels.  The result equals buildHostTree + flattenTree node-for-node and bit-for-bit (the split
// of every node is the unique median under the reference's total order, so it depends only on the
// set of leaves below the node, never on how the selection is carried out).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "mi_tree.hpp"

namespace mi {

struct DeviceTreeOut {
    TreeNode* nodes = nullptr;     // 2nf-1, reference pre-order
    Dipole* dipoles = nullptr;     // 2nf-1
    float4* dnode = nullptr;       // 4 per internal node
    float4* tri = nullptr;         // 3 per leaf slot
    float4* dint = nullptr;        // 2 per internal node
    float4* dleaf = nullptr;       // 2 per leaf slot
};

struct DeviceTreeInfo {
    int depth = 0;
    Box3 rootBox{};
    int32_t rootRef = 0;
};

// scratch bytes buildDeviceTree needs for nf faces
size_t deviceTreeScratchBytes(int32_t nf);

// d_verts float[nv*3], d_tris int32[nf*3] (device).  Returns hipErrorInvalidValue when a triangle
// names a vertex outside [0, nv).  Synchronises `s` (the level loop reads each level's size).
// d_faceIds (optional, device int32[nf]): the face id of each packed triangle, written to the leaf
// nodes and triangle records instead of the packed index (meshes with deleted faces).
hipError_t buildDeviceTree(const float* d_verts, int32_t nv, const int32_t* d_tris, int32_t nf, void* scratch,
                           size_t scratchBytes, const DeviceTreeOut& out, DeviceTreeInfo& info, hipStream_t s,
                           const int32_t* d_faceIds = nullptr);

} // namespace mi
