// mi_tree_gpu.hip -- AABB tree + dipoles + flattened layout built on the GPU.
//
// Tree (source/MRMesh/MRAABBTreeMaker.hpp:50-165): every internal node splits its leaves at the
// median (left gets num/2) under a total order -- key (min+max) along the node box's dims sorted by
// extent, largest first, ties to the next dim, last to the face id.  Because the order is total,
// the left set of each node is unique; the build below finds it for all nodes of one tree level at
// once, with no per-node selection:
//   * the leaves of each node stay contiguous, nodes of a level in pre-order (which is also their
//     array order, since left subtrees precede right subtrees in both);
//   * per level: segmented box reduction (wave-segmented shuffles + one atomic per run), per-leaf
//     keys from its node's dims, then a stable LSD radix chain face -> k2 -> k1 -> (node, k0), which
//     sorts every node's leaves by the full total order without moving them out of their node;
//   * the first num/2 leaves of a node form its left child; children are numbered root+1 and
//     root+2*half as in the reference's pre-order layout.
// Leaf boxes, dipoles (source/MRMesh/MRDipole.cpp:23-71) and the flattened arrays repeat the host
// builder's float expressions in the same order (the library is compiled with -ffp-contract=off and
// correctly rounded division and square root), so every record equals the host build bit-for-bit.
#include "mi_tree_gpu.hpp"

#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <vector>

namespace mi {
namespace {

constexpr int kB = 256;

// leaf record while building: {min.xyz, max.x} {max.y, max.z, face, node}
struct Elem { float4 a, b; };

__device__ __forceinline__ uint32_t ordU(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unordU(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
// sort key of a float under '<' (-0 and +0 compare equal)
__device__ __forceinline__ uint32_t keyU(float x) { return ordU(x + 0.0f); }

// std::nextafter(x, -FLT_MAX) / (x, FLT_MAX) for finite x
__device__ __forceinline__ float nextDown(float x)
{
    if (x == -FLT_MAX) return x;
    if (x == 0.0f) return __uint_as_float(0x80000001u);
    const uint32_t b = __float_as_uint(x);
    return __uint_as_float(x > 0.0f ? b - 1u : b + 1u);
}
__device__ __forceinline__ float nextUp(float x)
{
    if (x == FLT_MAX) return x;
    if (x == 0.0f) return __uint_as_float(0x00000001u);
    const uint32_t b = __float_as_uint(x);
    return __uint_as_float(x > 0.0f ? b + 1u : b - 1u);
}

// computeFaceBox + insignificantlyExpanded (source/MRMesh/MRAABBTree.cpp:15-29, MRBox.h:260-270)
__global__ void kLeafBoxes(const float* __restrict__ verts, int32_t nv, const int32_t* __restrict__ tris,
                           int32_t n, Elem* __restrict__ el, uint32_t* __restrict__ err)
{
    const int32_t f = blockIdx.x * kB + threadIdx.x;
    if (f >= n) return;
    float mn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, mx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
        const int32_t vi = tris[3 * (int64_t)f + k];
        if (vi < 0 || vi >= nv) { ok = false; continue; }
        const float* p = verts + 3 * (int64_t)vi;
        for (int i = 0; i < 3; ++i) {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    if (!ok) {
        atomicOr(err, 1u);
        for (int i = 0; i < 3; ++i) { mn[i] = 0; mx[i] = 0; }
    }
    for (int i = 0; i < 3; ++i) { mn[i] = nextDown(mn[i]); mx[i] = nextUp(mx[i]); }
    el[f].a = make_float4(mn[0], mn[1], mn[2], mx[0]);
    el[f].b = make_float4(mx[1], mx[2], __int_as_float(f), __int_as_float(0));
}

__global__ void kInit(int32_t n, int32_t* segFirst, int32_t* segNum, int32_t* list, uint32_t* iota)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    if (p == 0) { segFirst[0] = 0; segNum[0] = n; list[0] = 0; }
    if (p < n) iota[p] = (uint32_t)p;
}

__global__ void kSegInit(const int32_t* __restrict__ list, int32_t count, const int32_t* __restrict__ segNum,
                         uint32_t* __restrict__ acc)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const int32_t v = list[i];
    if (segNum[v] < 2) return;
    uint32_t* a = acc + 6 * (int64_t)v;
    a[0] = a[1] = a[2] = 0xffffffffu;      // min accumulators (ordered encoding)
    a[3] = a[4] = a[5] = 0u;               // max accumulators
}

// node box = union of its leaves' boxes: segmented min/max within the wave (a node's leaves are
// contiguous), then one atomic per node-run per wave
__global__ void kSegBox(const Elem* __restrict__ el, int32_t n, const int32_t* __restrict__ segNum,
                        uint32_t* __restrict__ acc)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    int v = -1;
    float b[6] = { FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX };
    if (p < n) {
        const Elem e = el[p];
        const int vv = __float_as_int(e.b.w);
        if (segNum[vv] >= 2) {
            v = vv;
            b[0] = e.a.x; b[1] = e.a.y; b[2] = e.a.z; b[3] = e.a.w; b[4] = e.b.x; b[5] = e.b.y;
        }
    }
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int ov = __shfl_up(v, d);
        float o[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) o[i] = __shfl_up(b[i], d);
        if (lane >= d && ov == v) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                b[i] = fminf(b[i], o[i]);
                b[i + 3] = fmaxf(b[i + 3], o[i + 3]);
            }
        }
    }
    const int nv = __shfl_down(v, 1);
    if (v >= 0 && (lane == 63 || nv != v)) {
        uint32_t* a = acc + 6 * (int64_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            atomicMin(a + i, ordU(b[i]));
            atomicMax(a + 3 + i, ordU(b[i + 3]));
        }
    }
}

__device__ __forceinline__ void nodeBox(const uint32_t* acc, int v, float mn[3], float mx[3])
{
    const uint32_t* a = acc + 6 * (int64_t)v;
    for (int i = 0; i < 3; ++i) { mn[i] = unordU(a[i]); mx[i] = unordU(a[3 + i]); }
}

// dims by (extent, index), smallest first -- findSortedBoxDims (source/MRMesh/MRBox.h:378-396)
__device__ __forceinline__ void sortedDims(const float mn[3], const float mx[3], int sd[3])
{
    const float e[3] = { mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2] };
    sd[0] = 0; sd[1] = 1; sd[2] = 2;
    auto lessP = [&](int a, int b) { return e[a] < e[b] || (!(e[b] < e[a]) && a < b); };
    int t;
    if (lessP(sd[1], sd[0])) { t = sd[0]; sd[0] = sd[1]; sd[1] = t; }
    if (lessP(sd[2], sd[1])) { t = sd[1]; sd[1] = sd[2]; sd[2] = t; }
    if (lessP(sd[1], sd[0])) { t = sd[0]; sd[0] = sd[1]; sd[1] = t; }
}

__global__ void kKeys(const Elem* __restrict__ el, int32_t n, const int32_t* __restrict__ segNum,
                      const uint32_t* __restrict__ acc, uint32_t* __restrict__ faceKey, uint32_t* __restrict__ k1,
                      uint32_t* __restrict__ k2, uint64_t* __restrict__ k64)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= n) return;
    const Elem e = el[p];
    const int v = __float_as_int(e.b.w);
    faceKey[p] = (uint32_t)__float_as_int(e.b.z);
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    if (segNum[v] >= 2) {
        float mn[3], mx[3];
        nodeBox(acc, v, mn, mx);
        int sd[3];
        sortedDims(mn, mx, sd);
        const float c[3] = { e.a.x + e.a.w, e.a.y + e.b.x, e.a.z + e.b.y };
        a0 = keyU(c[sd[2]]);
        a1 = keyU(c[sd[1]]);
        a2 = keyU(c[sd[0]]);
    }
    k1[p] = a1;
    k2[p] = a2;
    k64[p] = ((uint64_t)(uint32_t)v << 32) | a0;
}

template <typename T>
__global__ void kGather(const T* __restrict__ src, const uint32_t* __restrict__ perm, int32_t n, T* __restrict__ dst)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

__global__ void kGatherElems(const Elem* __restrict__ src, const uint32_t* __restrict__ perm, int32_t n,
                             Elem* __restrict__ dst)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    if (p < n) {
        const Elem e = src[perm[p]];
        dst[p].a = e.a;
        dst[p].b = e.b;
    }
}

__global__ void kZero(uint32_t* p) { if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0; }

// leaf dipole: dirArea = cross(b-a, c-a)/2, area = |dirArea|, pos = area * centroid
// (source/MRMesh/MRDipole.cpp:28-45, triCenter / leftDirDblArea in MRMeshMath.cpp:119-149)
__device__ __forceinline__ Dipole leafDipole(const float* __restrict__ verts, const int32_t* __restrict__ tris,
                                             int32_t face)
{
    const int32_t* t = tris + 3 * (int64_t)face;
    const float* a = verts + 3 * (int64_t)t[0];
    const float* b = verts + 3 * (int64_t)t[1];
    const float* c = verts + 3 * (int64_t)t[2];
    const float ab[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
    const float ac[3] = { c[0] - a[0], c[1] - a[1], c[2] - a[2] };
    const float cr[3] = { ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0] };
    const float third = 1 / 3.0f;
    Dipole d;
    for (int k = 0; k < 3; ++k) d.dirArea[k] = 0.5f * cr[k];
    d.area = sqrtf(d.dirArea[0] * d.dirArea[0] + d.dirArea[1] * d.dirArea[1] + d.dirArea[2] * d.dirArea[2]);
    for (int k = 0; k < 3; ++k) d.pos[k] = d.area * (third * ((a[k] + b[k]) + c[k]));
    d.rr = 0;
    return d;
}

// writes the nodes of one level; a leaf takes its element's box and face, an internal node its
// reduced box and children root+1 / root+2*half (MRAABBTreeMaker.hpp:88-89), whose leaf ranges are
// recorded and appended to the next level's list
__global__ void kSplit(const int32_t* __restrict__ list, int32_t count, int32_t* __restrict__ segFirst,
                       int32_t* __restrict__ segNum, const Elem* __restrict__ el, const uint32_t* __restrict__ acc,
                       const float* __restrict__ verts, const int32_t* __restrict__ tris, TreeNode* __restrict__ nodes,
                       Dipole* __restrict__ dip, int32_t* __restrict__ nextList, uint32_t* __restrict__ nextCount)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const int32_t v = list[i];
    const int32_t f = segFirst[v], s = segNum[v];
    TreeNode nd;
    if (s == 1) {
        const Elem e = el[f];
        nd.box.min[0] = e.a.x; nd.box.min[1] = e.a.y; nd.box.min[2] = e.a.z;
        nd.box.max[0] = e.a.w; nd.box.max[1] = e.b.x; nd.box.max[2] = e.b.y;
        nd.l = __float_as_int(e.b.z);
        nd.r = -1;
        dip[v] = leafDipole(verts, tris, nd.l);
    } else {
        const int32_t h = s / 2;
        nodeBox(acc, v, nd.box.min, nd.box.max);
        nd.l = v + 1;
        nd.r = v + 2 * h;
        segFirst[v + 1] = f;
        segNum[v + 1] = h;
        segFirst[v + 2 * h] = f + h;
        segNum[v + 2 * h] = s - h;
        const uint32_t slot = atomicAdd(nextCount, 2u);
        nextList[slot] = v + 1;
        nextList[slot + 1] = v + 2 * h;
    }
    nodes[v] = nd;
}

__global__ void kRelabel(Elem* __restrict__ el, int32_t n, const int32_t* __restrict__ segFirst,
                         const int32_t* __restrict__ segNum)
{
    const int32_t p = blockIdx.x * kB + threadIdx.x;
    if (p >= n) return;
    const int v = __float_as_int(el[p].b.w);
    const int32_t s = segNum[v];
    if (s < 2) return;
    const int32_t h = s / 2;
    const int child = (p - segFirst[v] < h) ? v + 1 : v + 2 * h;
    el[p].b.w = __int_as_float(child);
}

// internal dipoles of one level from their children (source/MRMesh/MRDipole.cpp:47-60)
__global__ void kInternalDipoles(const int32_t* __restrict__ list, int32_t count, const TreeNode* __restrict__ nodes,
                                 Dipole* __restrict__ dip)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const int32_t v = list[i];
    const TreeNode nd = nodes[v];
    if (nd.r < 0) return;
    const Dipole dl = dip[nd.l], dr = dip[nd.r];
    Dipole d;
    for (int k = 0; k < 3; ++k) {
        d.pos[k] = dl.pos[k] + dr.pos[k];
        d.dirArea[k] = dl.dirArea[k] + dr.dirArea[k];
    }
    d.area = dl.area + dr.area;
    d.rr = 0;
    dip[v] = d;
}

// pos /= area (as a multiplication by 1/area), rr = squared distance to the farthest box corner
// (source/MRMesh/MRDipole.cpp:11-21, 62-70)
__global__ void kFinalizeDipoles(int32_t nn, const TreeNode* __restrict__ nodes, Dipole* __restrict__ dip,
                                 int32_t* __restrict__ leafFlag)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= nn) return;
    Dipole d = dip[i];
    const TreeNode nd = nodes[i];
    if (d.area > 0) {
        const float inv = 1 / d.area;
        for (int k = 0; k < 3; ++k) d.pos[k] *= inv;
    }
    float rr = 0;
    for (int k = 0; k < 3; ++k) {
        float tmin = d.pos[k] - nd.box.min[k];
        tmin = tmin * tmin;
        float tmax = d.pos[k] - nd.box.max[k];
        tmax = tmax * tmax;
        rr += tmin > tmax ? tmin : tmax;
    }
    d.rr = rr;
    dip[i] = d;
    leafFlag[i] = nd.r < 0 ? 1 : 0;
}

__device__ __forceinline__ int32_t refOf(int32_t i, const int32_t* leafFlag, const int32_t* leafBefore)
{
    return leafFlag[i] ? ~leafBefore[i] : i - leafBefore[i];
}

// the traversal layout (mi_tree.hpp): internal nodes and leaf slots in pre-order
__global__ void kFlatten(int32_t nn, const TreeNode* __restrict__ nodes, const Dipole* __restrict__ dip,
                         const int32_t* __restrict__ leafFlag, const int32_t* __restrict__ leafBefore,
                         const float* __restrict__ verts, const int32_t* __restrict__ tris, float4* __restrict__ dnode,
                         float4* __restrict__ tri, float4* __restrict__ dint, float4* __restrict__ dleaf)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= nn) return;
    const TreeNode n = nodes[i];
    const Dipole d = dip[i];
    const float4 d0 = make_float4(d.pos[0], d.pos[1], d.pos[2], d.rr);
    const float4 d1 = make_float4(d.dirArea[0], d.dirArea[1], d.dirArea[2], d.area);
    if (n.r >= 0) {
        const int32_t j = i - leafBefore[i];
        const Box3 L = nodes[n.l].box, R = nodes[n.r].box;
        float4* r = dnode + 4 * (int64_t)j;
        r[0] = make_float4(L.min[0], L.min[1], L.min[2], L.max[0]);
        r[1] = make_float4(L.max[1], L.max[2], R.min[0], R.min[1]);
        r[2] = make_float4(R.min[2], R.max[0], R.max[1], R.max[2]);
        r[3] = make_float4(__int_as_float(refOf(n.l, leafFlag, leafBefore)),
                           __int_as_float(refOf(n.r, leafFlag, leafBefore)), 0.f, 0.f);
        dint[2 * (int64_t)j] = d0;
        dint[2 * (int64_t)j + 1] = d1;
    } else {
        const int32_t k = leafBefore[i];
        const int32_t f = n.l;
        const int32_t* t = tris + 3 * (int64_t)f;
        const float* a = verts + 3 * (int64_t)t[0];
        const float* b = verts + 3 * (int64_t)t[1];
        const float* c = verts + 3 * (int64_t)t[2];
        float4* r = tri + 3 * (int64_t)k;
        r[0] = make_float4(a[0], a[1], a[2], b[0]);
        r[1] = make_float4(b[1], b[2], c[0], c[1]);
        r[2] = make_float4(c[2], __int_as_float(f), 0.f, 0.f);
        dleaf[2 * (int64_t)k] = d0;
        dleaf[2 * (int64_t)k + 1] = d1;
    }
}

// A tree built over the valid faces of a mesh with deleted faces names the packed face index in its
// leaves; the reference's leaves hold the real FaceId (source/MRMesh/MRAABBTree.cpp:35-49).  Packing
// keeps the face order, so the tree (whose ties fall to the face index) is the same; only the ids
// written to the leaf nodes and the triangle records are mapped back.
__global__ void kRemapFaces(int32_t nn, TreeNode* __restrict__ nodes, const int32_t* __restrict__ leafFlag,
                            const int32_t* __restrict__ leafBefore, float4* __restrict__ tri,
                            const int32_t* __restrict__ faceIds)
{
    const int32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= nn || !leafFlag[i]) return;
    const int32_t f = faceIds[nodes[i].l];
    nodes[i].l = f;
    tri[3 * (int64_t)leafBefore[i] + 2].y = __int_as_float(f);
}

inline size_t al(size_t b) { return (b + 255) & ~size_t(255); }
inline int bitsFor(uint64_t x) { int b = 1; while (b < 64 && (x >> b)) ++b; return b; }
inline unsigned gridFor(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

struct Scratch {
    Elem* el[2];
    uint32_t *iota, *faceKey, *keyA, *keyB, *permA, *permB, *k1, *k2, *counters;
    uint64_t *k64, *k64A, *k64B;
    int32_t *segFirst, *segNum, *lists, *leafFlag, *leafBefore;
    uint32_t* acc;
    void* temp;
    size_t tempBytes;
};

size_t sortTempBytes(int32_t n)
{
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, n);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, n);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int32_t*)nullptr, (int32_t*)nullptr, 2 * n);
    return std::max(a, std::max(b, c));
}

size_t layout(int32_t nf, Scratch* s, char* base)
{
    const size_t n = (size_t)nf, nn = 2 * n - 1;
    size_t off = 0;
    auto take = [&](size_t bytes) { char* p = base ? base + off : nullptr; off += al(bytes); return p; };
    char* e0 = take(n * sizeof(Elem));
    char* e1 = take(n * sizeof(Elem));
    char* iota = take(n * 4);
    char* faceKey = take(n * 4);
    char* keyA = take(n * 4);
    char* keyB = take(n * 4);
    char* permA = take(n * 4);
    char* permB = take(n * 4);
    char* k1 = take(n * 4);
    char* k2 = take(n * 4);
    char* counters = take(64);
    char* k64 = take(n * 8);
    char* k64A = take(n * 8);
    char* k64B = take(n * 8);
    char* segFirst = take(nn * 4);
    char* segNum = take(nn * 4);
    char* lists = take(nn * 4);
    char* leafFlag = take(nn * 4);
    char* leafBefore = take(nn * 4);
    char* acc = take(nn * 24);
    const size_t tb = sortTempBytes(nf);
    char* temp = take(tb);
    if (s) {
        s->el[0] = reinterpret_cast<Elem*>(e0);
        s->el[1] = reinterpret_cast<Elem*>(e1);
        s->iota = reinterpret_cast<uint32_t*>(iota);
        s->faceKey = reinterpret_cast<uint32_t*>(faceKey);
        s->keyA = reinterpret_cast<uint32_t*>(keyA);
        s->keyB = reinterpret_cast<uint32_t*>(keyB);
        s->permA = reinterpret_cast<uint32_t*>(permA);
        s->permB = reinterpret_cast<uint32_t*>(permB);
        s->k1 = reinterpret_cast<uint32_t*>(k1);
        s->k2 = reinterpret_cast<uint32_t*>(k2);
        s->counters = reinterpret_cast<uint32_t*>(counters);
        s->k64 = reinterpret_cast<uint64_t*>(k64);
        s->k64A = reinterpret_cast<uint64_t*>(k64A);
        s->k64B = reinterpret_cast<uint64_t*>(k64B);
        s->segFirst = reinterpret_cast<int32_t*>(segFirst);
        s->segNum = reinterpret_cast<int32_t*>(segNum);
        s->lists = reinterpret_cast<int32_t*>(lists);
        s->leafFlag = reinterpret_cast<int32_t*>(leafFlag);
        s->leafBefore = reinterpret_cast<int32_t*>(leafBefore);
        s->acc = reinterpret_cast<uint32_t*>(acc);
        s->temp = temp;
        s->tempBytes = tb;
    }
    return off;
}

#define TG_HIP(x)                              \
    do {                                       \
        const hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_;       \
    } while (0)

} // namespace

size_t deviceTreeScratchBytes(int32_t nf)
{
    if (nf <= 0) return 0;
    return layout(nf, nullptr, nullptr);
}

hipError_t buildDeviceTree(const float* d_verts, int32_t nv, const int32_t* d_tris, int32_t nf, void* scratch,
                           size_t scratchBytes, const DeviceTreeOut& out, DeviceTreeInfo& info, hipStream_t st,
                           const int32_t* d_faceIds)
{
    info = DeviceTreeInfo{};
    if (nf <= 0) return hipSuccess;
    Scratch S;
    if (layout(nf, &S, static_cast<char*>(scratch)) > scratchBytes) return hipErrorInvalidValue;
    const int32_t n = nf, nn = 2 * nf - 1;
    const int faceBits = bitsFor((uint64_t)(n - 1));
    const int nodeBits = bitsFor((uint64_t)(nn - 1));
    uint32_t* nextCount = S.counters;
    uint32_t* err = S.counters + 16;

    kZero<<<1, 64, 0, st>>>(err);
    kLeafBoxes<<<gridFor(n), kB, 0, st>>>(d_verts, nv, d_tris, n, S.el[0], err);
    TG_HIP(hipGetLastError());
    kInit<<<gridFor(n), kB, 0, st>>>(n, S.segFirst, S.segNum, S.lists, S.iota);
    uint32_t hostErr = 0;
    TG_HIP(hipMemcpyAsync(&hostErr, err, 4, hipMemcpyDeviceToHost, st));
    TG_HIP(hipStreamSynchronize(st));
    if (hostErr) return hipErrorInvalidValue;

    std::vector<int32_t> levelOff, levelCount;
    int cur = 0;
    int32_t off = 0, count = 1;
    int64_t maxSize = n;
    while (count > 0) {
        const int32_t* list = S.lists + off;
        if (maxSize >= 2) {
            kSegInit<<<gridFor(count), kB, 0, st>>>(list, count, S.segNum, S.acc);
            kSegBox<<<gridFor(n), kB, 0, st>>>(S.el[cur], n, S.segNum, S.acc);
            kKeys<<<gridFor(n), kB, 0, st>>>(S.el[cur], n, S.segNum, S.acc, S.faceKey, S.k1, S.k2, S.k64);
            size_t tb = S.tempBytes;
            // stable LSD chain: face, k2, k1, then (node, k0)
            TG_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp, tb, S.faceKey, S.keyB, S.iota, S.permA, n, 0, faceBits, st));
            kGather<uint32_t><<<gridFor(n), kB, 0, st>>>(S.k2, S.permA, n, S.keyA);
            tb = S.tempBytes;
            TG_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp, tb, S.keyA, S.keyB, S.permA, S.permB, n, 0, 32, st));
            kGather<uint32_t><<<gridFor(n), kB, 0, st>>>(S.k1, S.permB, n, S.keyA);
            tb = S.tempBytes;
            TG_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp, tb, S.keyA, S.keyB, S.permB, S.permA, n, 0, 32, st));
            kGather<uint64_t><<<gridFor(n), kB, 0, st>>>(S.k64, S.permA, n, S.k64A);
            tb = S.tempBytes;
            TG_HIP(hipcub::DeviceRadixSort::SortPairs(S.temp, tb, S.k64A, S.k64B, S.permA, S.permB, n, 0,
                                                      32 + nodeBits, st));
            kGatherElems<<<gridFor(n), kB, 0, st>>>(S.el[cur], S.permB, n, S.el[cur ^ 1]);
            cur ^= 1;
        }
        kZero<<<1, 64, 0, st>>>(nextCount);
        kSplit<<<gridFor(count), kB, 0, st>>>(list, count, S.segFirst, S.segNum, S.el[cur], S.acc, d_verts, d_tris,
                                             out.nodes, out.dipoles, S.lists + off + count, nextCount);
        if (maxSize >= 2) kRelabel<<<gridFor(n), kB, 0, st>>>(S.el[cur], n, S.segFirst, S.segNum);
        TG_HIP(hipGetLastError());
        uint32_t next = 0;
        TG_HIP(hipMemcpyAsync(&next, nextCount, 4, hipMemcpyDeviceToHost, st));
        TG_HIP(hipStreamSynchronize(st));
        levelOff.push_back(off);
        levelCount.push_back(count);
        off += count;
        if ((int64_t)off + next > nn) return hipErrorUnknown;   // cannot happen for a binary tree of nn nodes
        count = (int32_t)next;
        maxSize = (maxSize + 1) / 2;
    }
    info.depth = (int)levelOff.size() - 1;
    for (int l = (int)levelOff.size() - 2; l >= 0; --l)
        kInternalDipoles<<<gridFor(levelCount[l]), kB, 0, st>>>(S.lists + levelOff[l], levelCount[l], out.nodes,
                                                                out.dipoles);
    kFinalizeDipoles<<<gridFor(nn), kB, 0, st>>>(nn, out.nodes, out.dipoles, S.leafFlag);
    size_t tb = S.tempBytes;
    TG_HIP(hipcub::DeviceScan::ExclusiveSum(S.temp, tb, S.leafFlag, S.leafBefore, nn, st));
    kFlatten<<<gridFor(nn), kB, 0, st>>>(nn, out.nodes, out.dipoles, S.leafFlag, S.leafBefore, d_verts, d_tris,
                                         out.dnode, out.tri, out.dint, out.dleaf);
    if (d_faceIds) kRemapFaces<<<gridFor(nn), kB, 0, st>>>(nn, out.nodes, S.leafFlag, S.leafBefore, out.tri, d_faceIds);
    TG_HIP(hipGetLastError());
    TreeNode root;
    TG_HIP(hipMemcpyAsync(&root, out.nodes, sizeof(TreeNode), hipMemcpyDeviceToHost, st));
    TG_HIP(hipStreamSynchronize(st));
    info.rootBox = root.box;
    info.rootRef = root.r < 0 ? ~0 : 0;
    return hipSuccess;
}

} // namespace mi
