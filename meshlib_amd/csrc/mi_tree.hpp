// mi_tree.hpp -- host-side AABB tree + dipole builder and the flattened HBM layout.
//
// The tree is the reference's deterministic median-split BVH
// (source/MRMesh/MRAABBTreeMaker.hpp:50-165, source/MRMesh/MRAABBTree.cpp:15-60);
// because the split is a total order, any exact median selection produces it
// node-for-node.  The dipoles follow source/MRMesh/MRDipole.cpp:23-71 with the
// same float operation order.
//
// Device layout ("one stream per traversal need", every record 16-B aligned):
//   dnode  4 x float4 per internal node j (pre-order among internal nodes):
//          {lmin.xyz, lmax.x} {lmax.yz, rmin.xy} {rmin.z, rmax.xyz} {lref, rref, 0, 0}
//          a child ref >= 0 is an internal node index, < 0 is ~(leaf slot)
//   tri    3 x float4 per leaf slot k (left-to-right leaf order):
//          {a.xyz, b.x} {b.yz, c.xy} {c.z, face, 0, 0}
//   dint   2 x float4 per internal node: {pos.xyz, rr} {dirArea.xyz, area}
//   dleaf  2 x float4 per leaf slot:     {pos.xyz, rr} {dirArea.xyz, area}
#pragma once

#include <cstdint>
#include <vector>

namespace mi {

struct Box3 { float min[3], max[3]; };
struct TreeNode { Box3 box; int32_t l, r; };                 // reference AABBTreeNode layout
struct Dipole { float pos[3]; float area; float dirArea[3]; float rr; };

struct HostTree {
    std::vector<TreeNode> nodes;       // 2F-1 nodes, reference pre-order numbering
    std::vector<Dipole> dipoles;       // per node
    int depth = 0;                     // longest root-to-leaf path in edges
};

// Builds the tree and dipoles.  verts float[nv*3], tris int32[nf*3].
void buildHostTree(const float* verts, int32_t nv, const int32_t* tris, int32_t nf, HostTree& out);

struct FlatTree {
    std::vector<float> dnode;   // 16 floats per internal node
    std::vector<float> tri;     // 12 floats per leaf slot
    std::vector<float> dint;    // 8 floats per internal node
    std::vector<float> dleaf;   // 8 floats per leaf slot
    Box3 rootBox{};
    int32_t rootRef = 0;        // 0 (internal root), ~0 (single leaf) ; unused when empty
    int32_t numInternal = 0, numLeaves = 0;
};

void flattenTree(const HostTree& t, const float* verts, const int32_t* tris, FlatTree& out);

} // namespace mi
