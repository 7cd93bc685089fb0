// mi_tree.cpp -- host AABB tree + dipoles (see mi_tree.hpp).
//
// Compiled with -ffp-contract=off: every float expression below is evaluated
// with the same rounding and order as the reference CPU build.
#include "mi_tree.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

namespace mi {
namespace {

struct BoxedLeaf { int32_t face; Box3 box; };

inline void boxInvalid(Box3& b)
{
    for (int i = 0; i < 3; ++i) { b.min[i] = FLT_MAX; b.max[i] = -FLT_MAX; }
}
inline void boxInclude(Box3& b, const float* p)
{
    for (int i = 0; i < 3; ++i) {
        if (p[i] < b.min[i]) b.min[i] = p[i];
        if (p[i] > b.max[i]) b.max[i] = p[i];
    }
}
inline void boxInclude(Box3& b, const Box3& o)
{
    for (int i = 0; i < 3; ++i) {
        if (o.min[i] < b.min[i]) b.min[i] = o.min[i];
        if (o.max[i] > b.max[i]) b.max[i] = o.max[i];
    }
}

// dims ordered by (extent, index), smallest first -- source/MRMesh/MRBox.h:378-396
inline void sortedDims(const Box3& b, int out[3])
{
    std::pair<float, int> ps[3];
    for (int i = 0; i < 3; ++i) ps[i] = { b.max[i] - b.min[i], i };
    std::sort(ps, ps + 3);
    for (int i = 0; i < 3; ++i) out[i] = ps[i].second;
}

struct Builder {
    const float* verts;
    const int32_t* tris;
    std::vector<BoxedLeaf> leaves;
    std::vector<TreeNode>* nodes;

    // splits [first, first+num) at its median under the reference's total order and
    // fills node `root`; returns the left count (MRAABBTreeMaker.hpp:50-95)
    int32_t makeNode(int32_t root, int32_t first, int32_t num)
    {
        TreeNode& node = (*nodes)[root];
        boxInvalid(node.box);
        for (int32_t i = 0; i < num; ++i) boxInclude(node.box, leaves[first + i].box);
        int sd[3];
        sortedDims(node.box, sd);
        auto less = [sd](const BoxedLeaf& a, const BoxedLeaf& b) {
            for (int i = 2; i >= 0; --i) {
                const int d = sd[i];
                const float ad = a.box.min[d] + a.box.max[d];
                const float bd = b.box.min[d] + b.box.max[d];
                if (ad != bd) return ad < bd;
            }
            return a.face < b.face;
        };
        const int32_t half = num / 2;
        std::nth_element(leaves.begin() + first, leaves.begin() + first + half,
                         leaves.begin() + first + num, less);
        node.l = root + 1;
        node.r = root + 1 + (2 * half - 1);
        return half;
    }

    void makeLeaf(int32_t root, int32_t first)
    {
        TreeNode& node = (*nodes)[root];
        node.box = leaves[first].box;
        node.l = leaves[first].face;
        node.r = -1;
    }

    void build(int32_t root, int32_t first, int32_t num, int threads)
    {
        if (num == 1) { makeLeaf(root, first); return; }
        const int32_t half = makeNode(root, first, num);
        const int32_t l = root + 1, r = root + 1 + (2 * half - 1);
        if (threads >= 2 && num >= 4096) {
            std::thread t([&] { build(r, first + half, num - half, threads / 2); });
            build(l, first, half, threads - threads / 2);
            t.join();
        } else {
            build(l, first, half, 1);
            build(r, first + half, num - half, 1);
        }
    }
};

int subtreeDepth(const std::vector<TreeNode>& nodes)
{
    // pre-order layout: depth of node i is known when it is reached from its parent
    std::vector<int> d(nodes.size(), 0);
    int maxd = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
        const TreeNode& n = nodes[i];
        if (n.r >= 0) {
            d[n.l] = d[i] + 1;
            d[n.r] = d[i] + 1;
        }
        maxd = std::max(maxd, d[i]);
    }
    return maxd;
}

} // namespace

void buildHostTree(const float* verts, int32_t nv, const int32_t* tris, int32_t nf, HostTree& out)
{
    (void)nv;
    out.nodes.clear();
    out.dipoles.clear();
    out.depth = 0;
    if (nf <= 0) return;
    Builder b;
    b.verts = verts;
    b.tris = tris;
    b.leaves.resize(nf);
    // computeFaceBox + insignificantlyExpanded (MRAABBTree.cpp:15-29, MRBox.h:260-270)
    for (int32_t f = 0; f < nf; ++f) {
        BoxedLeaf& bl = b.leaves[f];
        bl.face = f;
        boxInvalid(bl.box);
        for (int k = 0; k < 3; ++k) boxInclude(bl.box, verts + 3 * (int64_t)tris[3 * (int64_t)f + k]);
        for (int i = 0; i < 3; ++i) {
            bl.box.min[i] = std::nextafter(bl.box.min[i], -FLT_MAX);
            bl.box.max[i] = std::nextafter(bl.box.max[i], FLT_MAX);
        }
    }
    out.nodes.resize(2 * (size_t)nf - 1);
    b.nodes = &out.nodes;
    unsigned hc = std::thread::hardware_concurrency();
    int threads = 1;
    while (threads < (int)(hc ? hc : 1) * 2 && threads < 64) threads *= 2;
    b.build(0, 0, nf, threads);
    out.depth = subtreeDepth(out.nodes);

    // dipoles (MRDipole.cpp:23-71)
    const size_t nn = out.nodes.size();
    out.dipoles.assign(nn, Dipole{});
    const float third = 1 / 3.0f;
    for (size_t i = 0; i < nn; ++i) {
        const TreeNode& node = out.nodes[i];
        if (node.r >= 0) continue;
        const int32_t* t = tris + 3 * (int64_t)node.l;
        const float* a = verts + 3 * (int64_t)t[0];
        const float* bb = verts + 3 * (int64_t)t[1];
        const float* c = verts + 3 * (int64_t)t[2];
        const float ab[3] = { bb[0] - a[0], bb[1] - a[1], bb[2] - a[2] };
        const float ac[3] = { c[0] - a[0], c[1] - a[1], c[2] - a[2] };
        const float cr[3] = { ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                              ab[0] * ac[1] - ab[1] * ac[0] };
        Dipole& d = out.dipoles[i];
        for (int k = 0; k < 3; ++k) d.dirArea[k] = 0.5f * cr[k];
        d.area = std::sqrt(d.dirArea[0] * d.dirArea[0] + d.dirArea[1] * d.dirArea[1]
                           + d.dirArea[2] * d.dirArea[2]);
        for (int k = 0; k < 3; ++k) d.pos[k] = d.area * (third * ((a[k] + bb[k]) + c[k]));
    }
    for (size_t ii = nn; ii-- > 0;) {
        const TreeNode& node = out.nodes[ii];
        if (node.r < 0) continue;
        const Dipole& dl = out.dipoles[node.l];
        const Dipole& dr = out.dipoles[node.r];
        Dipole& d = out.dipoles[ii];
        for (int k = 0; k < 3; ++k) {
            d.pos[k] = dl.pos[k] + dr.pos[k];
            d.dirArea[k] = dl.dirArea[k] + dr.dirArea[k];
        }
        d.area = dl.area + dr.area;
    }
    for (size_t i = 0; i < nn; ++i) {
        Dipole& d = out.dipoles[i];
        if (d.area > 0) {
            const float inv = 1 / d.area;       // Vector3::operator/= multiplies by 1/b
            for (int k = 0; k < 3; ++k) d.pos[k] *= inv;
        }
        const Box3& bx = out.nodes[i].box;
        float rr = 0;
        for (int k = 0; k < 3; ++k) {
            float tmin = d.pos[k] - bx.min[k]; tmin = tmin * tmin;
            float tmax = d.pos[k] - bx.max[k]; tmax = tmax * tmax;
            rr += std::max(tmin, tmax);
        }
        d.rr = rr;
    }
}

void flattenTree(const HostTree& t, const float* verts, const int32_t* tris, FlatTree& out)
{
    const size_t nn = t.nodes.size();
    out = FlatTree{};
    if (nn == 0) return;
    std::vector<int32_t> ref(nn);
    int32_t ni = 0, nl = 0;
    for (size_t i = 0; i < nn; ++i)
        ref[i] = t.nodes[i].r >= 0 ? ni++ : ~(nl++);
    out.numInternal = ni;
    out.numLeaves = nl;
    out.rootBox = t.nodes[0].box;
    out.rootRef = ref[0];
    out.dnode.assign((size_t)ni * 16, 0.f);
    out.dint.assign((size_t)ni * 8, 0.f);
    out.tri.assign((size_t)nl * 12, 0.f);
    out.dleaf.assign((size_t)nl * 8, 0.f);
    for (size_t i = 0; i < nn; ++i) {
        const TreeNode& n = t.nodes[i];
        const Dipole& d = t.dipoles[i];
        float* dp;
        if (n.r >= 0) {
            const int32_t j = ref[i];
            const Box3& L = t.nodes[n.l].box;
            const Box3& R = t.nodes[n.r].box;
            float* r = &out.dnode[(size_t)j * 16];
            r[0] = L.min[0]; r[1] = L.min[1]; r[2] = L.min[2]; r[3] = L.max[0];
            r[4] = L.max[1]; r[5] = L.max[2]; r[6] = R.min[0]; r[7] = R.min[1];
            r[8] = R.min[2]; r[9] = R.max[0]; r[10] = R.max[1]; r[11] = R.max[2];
            const int32_t lr = ref[n.l], rr = ref[n.r];
            std::memcpy(&r[12], &lr, 4);
            std::memcpy(&r[13], &rr, 4);
            dp = &out.dint[(size_t)j * 8];
        } else {
            const int32_t k = ~ref[i];
            const int32_t f = n.l;
            const int32_t* tv = tris + 3 * (int64_t)f;
            float* r = &out.tri[(size_t)k * 12];
            for (int v = 0; v < 3; ++v)
                for (int c = 0; c < 3; ++c) r[3 * v + c] = verts[3 * (int64_t)tv[v] + c];
            std::memcpy(&r[9], &f, 4);
            dp = &out.dleaf[(size_t)k * 8];
        }
        dp[0] = d.pos[0]; dp[1] = d.pos[1]; dp[2] = d.pos[2]; dp[3] = d.rr;
        dp[4] = d.dirArea[0]; dp[5] = d.dirArea[1]; dp[6] = d.dirArea[2]; dp[7] = d.area;
    }
}

} // namespace mi
