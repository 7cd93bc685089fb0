// mi_walk.hpp -- device-side building blocks of the gfx950 traversal kernels (internal).
//
// Shared by the kernel translation units: record loads, the reference-order primitives (box
// distance, FP64 closest point, solid angle), the per-lane and wave-cooperative walks, the split
// form's helpers and small host-side launch helpers.  Everything has internal linkage: each
// translation unit instantiates what it uses.
#pragma once

#include "mi_kernels.hpp"

#include <atomic>
#include <cfloat>
#include <cstdint>
#include <type_traits>

// v_writelane_b32: `old` with lane `lane` replaced by the wave-uniform `value` (the compiler's
// intrinsic, reached by name: clang has no builtin for it)
__device__ int mi_writeLane(int value, int lane, int old) __asm("llvm.amdgcn.writelane");

namespace mi {
namespace {

constexpr int kBlock = 256;          // 4 waves

// pending entries of a 4-wide walk over a tree of depth <= D: each step pushes at most 3 and goes
// down two levels (one only into a leaf, where the path ends)
constexpr int quadStack(int D) { return 3 * (D / 2 + 1); }
constexpr int kXcds = 8;
constexpr float kInv4Pi = 1.0f / (4 * 3.14159265358979f);   // INV_4PI, MRDipole.cpp:125

// Mesh records are read-only while any kernel runs: reading them through the constant address
// space lets the compiler use scalar loads for wave-uniform refs even in kernels that also write
// (the split passes' atomics), instead of 64 identical vector loads.
typedef __attribute__((address_space(4))) const float cfloat;
typedef __attribute__((address_space(4))) const int32_t cint;
__device__ __forceinline__ float4 cload(const float4* p, size_t i)
{
    const cfloat* f = (const cfloat*)p + 4 * i;
    return make_float4(f[0], f[1], f[2], f[3]);
}

__device__ __forceinline__ CutEntry loadCut(const MeshDev& m, uint32_t c)
{
    const cint* e = (const cint*)(m.cut + c);
    const cfloat* f = (const cfloat*)(m.cut + c);
    CutEntry E;
    E.ref = e[0]; E.depth = e[1]; E.path = (uint32_t)e[2]; E.pad = 0;
    for (int i = 0; i < 3; ++i) { E.mn[i] = f[4 + i]; E.mx[i] = f[7 + i]; }
    return E;
}

__device__ __forceinline__ int logicalBlock(int nLogical)
{
    // blocks are dealt to XCDs round-robin; give XCD x the contiguous range
    // [x*per, (x+1)*per) of logical blocks
    const int per = (nLogical + kXcds - 1) / kXcds;
    const int b = blockIdx.x;
    return (b % kXcds) * per + b / kXcds;
}

// Grid kernels instead deal chunks of kGridChunk consecutive logical blocks (64 bricks) to the XCDs
// in turn: each L2 still sees spatially compact runs, and the costly middle of a grid (voxels deep
// inside the mesh) is spread over all eight XCDs instead of landing on two of them.  The launch
// covers whole rounds of chunks (chunkedGrid), so the mapping is a bijection onto its blocks.
constexpr int kGridChunk = 16;
__device__ __forceinline__ int gridBlock()
{
    const int b = blockIdx.x;
    const int x = b % kXcds, r = b / kXcds;
    return ((r / kGridChunk) * kXcds + x) * kGridChunk + r % kGridChunk;
}

inline int chunkedGrid(int nLogical)
{
    const int round = kXcds * kGridChunk;
    return (nLogical + round - 1) / round * round;
}

// blocks of a persistent launch: enough resident waves to fill every CU
inline int persistentBlocks()
{
    static std::atomic<int> cached[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev]) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached[dev] = cus * 8;
    }
    return cached[dev];
}

inline int paddedGrid(int nLogical)
{
    const int per = (nLogical + kXcds - 1) / kXcds;
    return per * kXcds;
}

// Box::getDistanceSq(pt) (source/MRMesh/MRBox.h:207-220): the per-dim term is the
// distance to the nearer face outside the box, zero inside; summed x, y, z.  The term is
// formed as clamp(p, mn, mx) - p (one med3 and one subtraction): below the box it is
// mn - p, above it mx - p = -(p - mx) exactly, inside p - p = 0, so its square equals the
// reference's max(mn - p, p - mx, 0)^2 bit for bit.  Requires mn <= mx; an empty slot is
// the point box at +inf (distance +inf).
__device__ __forceinline__ float boxDistSq(float mnx, float mny, float mnz, float mxx, float mxy,
                                           float mxz, float px, float py, float pz)
{
    const float tx = __builtin_amdgcn_fmed3f(px, mnx, mxx) - px;
    const float ty = __builtin_amdgcn_fmed3f(py, mny, mxy) - py;
    const float tz = __builtin_amdgcn_fmed3f(pz, mnz, mxz) - pz;
    return tx * tx + ty * ty + tz * tz;
}

__device__ __forceinline__ void xfApply(const Xf& xf, float& x, float& y, float& z)
{
    const float rx = xf.m[0] * x + xf.m[1] * y + xf.m[2] * z;
    const float ry = xf.m[3] * x + xf.m[4] * y + xf.m[5] * z;
    const float rz = xf.m[6] * x + xf.m[7] * y + xf.m[8] * z;
    x = rx + xf.m[9];
    y = ry + xf.m[10];
    z = rz + xf.m[11];
}

// transformed(box, xf): bbox of the 8 transformed corners (source/MRMesh/MRBox.h:336-344)
__device__ __forceinline__ void xfBox(const Xf& xf, float* mn, float* mx)
{
    float omn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, omx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float x = (c & 1) ? mx[0] : mn[0], y = (c & 2) ? mx[1] : mn[1], z = (c & 4) ? mx[2] : mn[2];
        xfApply(xf, x, y, z);
        omn[0] = x < omn[0] ? x : omn[0]; omx[0] = x > omx[0] ? x : omx[0];
        omn[1] = y < omn[1] ? y : omn[1]; omx[1] = y > omx[1] ? y : omx[1];
        omn[2] = z < omn[2] ? z : omn[2]; omx[2] = z > omx[2] ? z : omx[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) { mn[i] = omn[i]; mx[i] = omx[i]; }
}

struct Leaf { float a[3], b[3], c[3]; int32_t face; };

__device__ __forceinline__ Leaf loadLeaf(const float4* __restrict__ tri, int k)
{
    const float4 t0 = cload(tri, 3 * (size_t)k), t1 = cload(tri, 3 * (size_t)k + 1), t2 = cload(tri, 3 * (size_t)k + 2);
    Leaf L;
    L.a[0] = t0.x; L.a[1] = t0.y; L.a[2] = t0.z;
    L.b[0] = t0.w; L.b[1] = t1.x; L.b[2] = t1.y;
    L.c[0] = t1.z; L.c[1] = t1.w; L.c[2] = t2.x;
    L.face = __float_as_int(t2.y);
    return L;
}

__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by, double bz)
{
    return ax * bx + ay * by + az * bz;
}

// closestPointInTriangle<double> (source/MRMesh/MRClosestPointInTriangle.h:27-87): Voronoi
// regions #1 (a), #2 (b), #3 (c), #4 (ab), #5 (ac), #6 (bc), #0 (interior).
__device__ __forceinline__ void closestPointD(const double p[3], const double a[3], const double b[3],
                                              const double c[3], double pr[3], double& ba, double& bb)
{
    const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
    const double acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
    const double apx = p[0] - a[0], apy = p[1] - a[1], apz = p[2] - a[2];
    const double d1 = dot3(abx, aby, abz, apx, apy, apz);
    const double d2 = dot3(acx, acy, acz, apx, apy, apz);
    if (d1 <= 0 && d2 <= 0) { pr[0] = a[0]; pr[1] = a[1]; pr[2] = a[2]; ba = 0; bb = 0; return; }
    const double bpx = p[0] - b[0], bpy = p[1] - b[1], bpz = p[2] - b[2];
    const double d3 = dot3(abx, aby, abz, bpx, bpy, bpz);
    const double d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    if (d3 >= 0 && d4 <= d3) { pr[0] = b[0]; pr[1] = b[1]; pr[2] = b[2]; ba = 1; bb = 0; return; }
    const double cpx = p[0] - c[0], cpy = p[1] - c[1], cpz = p[2] - c[2];
    const double d5 = dot3(abx, aby, abz, cpx, cpy, cpz);
    const double d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    if (d6 >= 0 && d5 <= d6) { pr[0] = c[0]; pr[1] = c[1]; pr[2] = c[2]; ba = 0; bb = 1; return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        pr[0] = a[0] + v * abx; pr[1] = a[1] + v * aby; pr[2] = a[2] + v * abz;
        ba = v; bb = 0; return;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d6 <= 0 && d2 >= 0) {
        const double v = d2 / (d2 - d6);
        pr[0] = a[0] + v * acx; pr[1] = a[1] + v * acy; pr[2] = a[2] + v * acz;
        ba = 0; bb = v; return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0) {
        if (d4 < d3) { pr[0] = b[0]; pr[1] = b[1]; pr[2] = b[2]; ba = 1; bb = 0; return; }
        if (d5 < d6) { pr[0] = c[0]; pr[1] = c[1]; pr[2] = c[2]; ba = 0; bb = 1; return; }
        const double v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        pr[0] = b[0] + v * (c[0] - b[0]); pr[1] = b[1] + v * (c[1] - b[1]); pr[2] = b[2] + v * (c[2] - b[2]);
        ba = 1 - v; bb = v; return;
    }
    const double denom = 1 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    pr[0] = a[0] + v * abx + w * acx;
    pr[1] = a[1] + v * aby + w * acy;
    pr[2] = a[2] + v * abz + w * acz;
    ba = v; bb = w;
}

// exact squared distance of (px, py, pz) to a leaf's triangle: the FP64 closest point rounded to
// float, then the float sum of squares (the leaf test of findProjectionSubtree)
__device__ __forceinline__ float leafDistSq(const Leaf& L, float px, float py, float pz)
{
    const double pd[3] = { px, py, pz };
    const double ad[3] = { L.a[0], L.a[1], L.a[2] };
    const double bd[3] = { L.b[0], L.b[1], L.b[2] };
    const double cd[3] = { L.c[0], L.c[1], L.c[2] };
    double pr[3], ba, bb;
    closestPointD(pd, ad, bd, cd, pr, ba, bb);
    const float dx = (float)pr[0] - px, dy = (float)pr[1] - py, dz = (float)pr[2] - pz;
    return dx * dx + dy * dy + dz * dz;
}

// The leaf record through vector loads: for per-lane leaf indices.
__device__ __forceinline__ Leaf loadLeafLanes(const float4* __restrict__ tri, int k)
{
    const float4 t0 = tri[3 * (size_t)k], t1 = tri[3 * (size_t)k + 1], t2 = tri[3 * (size_t)k + 2];
    Leaf L;
    L.a[0] = t0.x; L.a[1] = t0.y; L.a[2] = t0.z;
    L.b[0] = t0.w; L.b[1] = t1.x; L.b[2] = t1.y;
    L.c[0] = t1.z; L.c[1] = t1.w; L.c[2] = t2.x;
    L.face = __float_as_int(t2.y);
    return L;
}

// triangleSolidAngle (source/MRMesh/MRDipole.cpp:81-92), det as MRMatrix3.h:226-232
__device__ __forceinline__ float solidAngle(float px, float py, float pz, const Leaf& L)
{
    const float mxx = L.a[0] - px, mxy = L.a[1] - py, mxz = L.a[2] - pz;
    const float myx = L.b[0] - px, myy = L.b[1] - py, myz = L.b[2] - pz;
    const float mzx = L.c[0] - px, mzy = L.c[1] - py, mzz = L.c[2] - pz;
    const float x = sqrtf(mxx * mxx + mxy * mxy + mxz * mxz);
    const float y = sqrtf(myx * myx + myy * myy + myz * myz);
    const float z = sqrtf(mzx * mzx + mzy * mzy + mzz * mzz);
    const float dxy = mxx * myx + mxy * myy + mxz * myz;
    const float dyz = myx * mzx + myy * mzy + myz * mzz;
    const float dzx = mzx * mxx + mzy * mxy + mzz * mxz;
    const float den = x * y * z + dxy * z + dyz * x + dzx * y;
    const float det = mxx * (myy * mzz - myz * mzy) - mxy * (myx * mzz - myz * mzx)
                    + mxz * (myx * mzy - myy * mzx);
    return 2 * atan2f(det, den);
}

struct BestHit { float distSq; int32_t face; float px, py, pz, a, b; };

// findProjectionSubtree (source/MRMesh/MRMeshProject.cpp:13-103) for one lane.
// FULL: keep the winning face/point/barycentrics (projections) or only distSq (SDF).
// XF: apply a non-rigid tree transform to boxes and triangles.
template <bool FULL, bool XF, int D>
__device__ __forceinline__ void distanceDfs(const MeshDev& m, float px, float py, float pz,
                                            float up, float lo, const Xf& treeXf, int2* stk,
                                            BestHit& res)
{
    res.distSq = up;
    res.face = -1;
    if (m.numFaces <= 0) return;
    float rmn[3] = { m.rootMin[0], m.rootMin[1], m.rootMin[2] };
    float rmx[3] = { m.rootMax[0], m.rootMax[1], m.rootMax[2] };
    if (XF) xfBox(treeXf, rmn, rmx);
    int32_t cur = m.rootRef;
    float curd = boxDistSq(rmn[0], rmn[1], rmn[2], rmx[0], rmx[1], rmx[2], px, py, pz);
    bool have = curd < res.distSq;
    int sp = 0;
    const double pd[3] = { px, py, pz };
    for (;;) {
        if (!have) {
            // pop; entries whose box is no closer than the best so far are dropped
            if (sp == 0) break;
            --sp;
            const int2 e = stk[sp * kBlock];
            cur = e.x;
            curd = __int_as_float(e.y);
            if (curd >= res.distSq) continue;
        }
        have = false;
        if (cur < 0) {
            Leaf L = loadLeaf(m.tri, ~cur);
            if (XF) {
                xfApply(treeXf, L.a[0], L.a[1], L.a[2]);
                xfApply(treeXf, L.b[0], L.b[1], L.b[2]);
                xfApply(treeXf, L.c[0], L.c[1], L.c[2]);
            }
            const double ad[3] = { L.a[0], L.a[1], L.a[2] };
            const double bd[3] = { L.b[0], L.b[1], L.b[2] };
            const double cd[3] = { L.c[0], L.c[1], L.c[2] };
            double pr[3], ba, bb;
            closestPointD(pd, ad, bd, cd, pr, ba, bb);
            const float qx = (float)pr[0], qy = (float)pr[1], qz = (float)pr[2];
            const float dx = qx - px, dy = qy - py, dz = qz - pz;
            const float d = dx * dx + dy * dy + dz * dz;
            if (d < res.distSq) {
                res.distSq = d;
                if (FULL) {
                    res.face = L.face;
                    res.px = qx; res.py = qy; res.pz = qz;
                    res.a = (float)ba; res.b = (float)bb;
                }
                if (d <= lo) break;
            }
            continue;
        }
        const float4 n0 = cload(m.dnode, 4 * cur), n1 = cload(m.dnode, 4 * cur + 1);
        const float4 n2 = cload(m.dnode, 4 * cur + 2), n3 = cload(m.dnode, 4 * cur + 3);
        float lmn[3] = { n0.x, n0.y, n0.z }, lmx[3] = { n0.w, n1.x, n1.y };
        float rmn2[3] = { n1.z, n1.w, n2.x }, rmx2[3] = { n2.y, n2.z, n2.w };
        if (XF) { xfBox(treeXf, lmn, lmx); xfBox(treeXf, rmn2, rmx2); }
        const float d1 = boxDistSq(lmn[0], lmn[1], lmn[2], lmx[0], lmx[1], lmx[2], px, py, pz);
        const float d2 = boxDistSq(rmn2[0], rmn2[1], rmn2[2], rmx2[0], rmx2[1], rmx2[2], px, py, pz);
        const int32_t lref = __float_as_int(n3.x), rref = __float_as_int(n3.y);
        // nearer child first; equal distances explore the right child first
        const bool leftNear = d1 < d2;
        const int32_t nearRef = leftNear ? lref : rref, farRef = leftNear ? rref : lref;
        const float nearD = leftNear ? d1 : d2, farD = leftNear ? d2 : d1;
        if (farD < res.distSq) {
            stk[sp * kBlock] = make_int2(farRef, __float_as_int(farD));
            ++sp;
        }
        if (nearD < res.distSq) { cur = nearRef; curd = nearD; have = true; }
    }
}

// calcFastWindingNumber (source/MRMesh/MRDipole.cpp:94-128) for one lane; the stack
// holds node refs in the .x component of the shared LDS column.
__device__ __forceinline__ float windingNumber(const MeshDev& m, float qx, float qy, float qz,
                                               float beta, int32_t skipFace, int2* stk)
{
    if (m.numFaces <= 0) return 0;
    const float betaSq = beta * beta;
    float res = 0;
    int sp = 0;
    int32_t cur = m.rootRef;
    for (;;) {
        float4 d0, d1;
        if (cur >= 0) { d0 = cload(m.dint, 2 * cur); d1 = cload(m.dint, 2 * cur + 1); }
        else { d0 = cload(m.dleaf, 2 * (~cur)); d1 = cload(m.dleaf, 2 * (~cur) + 1); }
        // Dipole::addIfGoodApprox (source/MRMesh/MRDipole.h:18-27)
        const float dpx = d0.x - qx, dpy = d0.y - qy, dpz = d0.z - qz;
        const float dd = dpx * dpx + dpy * dpy + dpz * dpz;
        if (!(dd <= betaSq * d0.w)) {
            const float dist = sqrtf(dd);
            if (dist > 0) res += (dpx * d1.x + dpy * d1.y + dpz * d1.z) / (dist * dd);
        } else if (cur >= 0) {
            const float4 n3 = cload(m.dnode, 4 * cur + 3);
            stk[sp * kBlock].x = __float_as_int(n3.y);   // right child later
            ++sp;
            cur = __float_as_int(n3.x);                   // left child next
            continue;
        } else {
            const Leaf L = loadLeaf(m.tri, ~cur);
            if (L.face != skipFace) res += solidAngle(qx, qy, qz, L);
        }
        if (sp == 0) break;
        --sp;
        cur = stk[sp * kBlock].x;
    }
    return kInv4Pi * res;
}

// voxel of lane `lane` in brick `brick` (4x4x4 per wave, bricks x-fastest)
__device__ __forceinline__ bool brickVoxel(int brick, int lane, int nbx, int nby, int nbz, int dx,
                                           int dy, int zBegin, int zEnd, int& x, int& y, int& z)
{
    const int bx = brick % nbx;
    const int t = brick / nbx;
    const int by = t % nby;
    const int bz = t / nby;
    x = bx * 4 + (lane & 3);
    y = by * 4 + ((lane >> 2) & 3);
    z = zBegin + bz * 4 + (lane >> 4);
    return bz < nbz && x < dx && y < dy && z < zEnd;
}

// ---------------------------------------------------------------- wave-cooperative traversals
//
// One wave walks ONE depth-first traversal for its 64 coherent queries: the node ref is
// wave-uniform (readfirstlane), so every node record is fetched once per wave at a uniform
// address instead of up to 64 divergent ones, and the stack is one short wave-private array of
// uniform entries in LDS (a few hundred bytes per wave instead of D*8 bytes per lane).

constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ uint64_t waveBallot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float laneValue(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// adds a wave-uniform count to the mesh's measurement counter `slot` (lane 0; nothing without counters)
__device__ __forceinline__ void statAdd(const MeshDev& m, int slot, uint32_t v)
{
    if (m.stats && v && (threadIdx.x & 63) == 0) atomicAdd(&m.stats[slot], (unsigned long long)v);
}

// Upper bound of one query's squared distance: a greedy descent (nearer child box at every level,
// per lane) to one leaf, then that leaf's exact distance (the leaf test of the walks).  It is a real
// candidate, so seeding a walk with it leaves the result unchanged while every lane prunes from the
// first node on -- the walks of a wave whose lanes are far apart no longer follow one lane's order
// with the other lanes' limits still open.
// Per-lane refs: plain (vector) loads, kept apart from the scalar-load pointers of the walks so the
// compiler does not move the walks' record addresses into vector registers as well.
__device__ __forceinline__ float greedyLeafDistSq(const MeshDev& m, float px, float py, float pz)
{
    int cur = m.rootRef;
    while (cur >= 0) {
        const float4 n0 = m.dnode[4 * cur], n1 = m.dnode[4 * cur + 1];
        const float4 n2 = m.dnode[4 * cur + 2], n3 = m.dnode[4 * cur + 3];
        const float d1 = boxDistSq(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, px, py, pz);
        const float d2 = boxDistSq(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, px, py, pz);
        cur = d1 <= d2 ? __float_as_int(n3.x) : __float_as_int(n3.y);
    }
    return leafDistSq(loadLeafLanes(m.tri, ~cur), px, py, pz);
}

// Minimum squared distance of each lane's point to the mesh below `up` (every leaf whose box is
// closer than the lane's current best is tested, as in findProjectionSubtree).  The result is a
// minimum over candidates, so it does not depend on the visit order; lanes pass up < 0 to opt out.
// Slots of a 4-wide record q[0..7] (built by kBuildQuads): slot i's box and ref.
struct Quad {
    float mn[4][3], mx[4][3];
    int32_t ref[4];
    bool leafA, leafB;      // left / right child is a leaf (it fills slot 0 / 2, slot 1 / 3 empty)
};

__device__ __forceinline__ Quad loadQuad(const MeshDev& m, int j)
{
    const size_t b = 8 * (size_t)j;
    const float4 q0 = cload(m.dquad, b), q1 = cload(m.dquad, b + 1), q2 = cload(m.dquad, b + 2);
    const float4 q3 = cload(m.dquad, b + 3), q4 = cload(m.dquad, b + 4), q5 = cload(m.dquad, b + 5);
    const float4 q6 = cload(m.dquad, b + 6), q7 = cload(m.dquad, b + 7);
    Quad Q;
    Q.mn[0][0] = q0.x; Q.mn[0][1] = q0.y; Q.mn[0][2] = q0.z; Q.mx[0][0] = q0.w; Q.mx[0][1] = q1.x; Q.mx[0][2] = q1.y;
    Q.mn[1][0] = q1.z; Q.mn[1][1] = q1.w; Q.mn[1][2] = q2.x; Q.mx[1][0] = q2.y; Q.mx[1][1] = q2.z; Q.mx[1][2] = q2.w;
    Q.mn[2][0] = q3.x; Q.mn[2][1] = q3.y; Q.mn[2][2] = q3.z; Q.mx[2][0] = q3.w; Q.mx[2][1] = q4.x; Q.mx[2][2] = q4.y;
    Q.mn[3][0] = q4.z; Q.mn[3][1] = q4.w; Q.mn[3][2] = q5.x; Q.mx[3][0] = q5.y; Q.mx[3][1] = q5.z; Q.mx[3][2] = q5.w;
    Q.ref[0] = __float_as_int(q6.x); Q.ref[1] = __float_as_int(q6.y);
    Q.ref[2] = __float_as_int(q6.z); Q.ref[3] = __float_as_int(q6.w);
    Q.leafA = __float_as_int(q7.x) != 0;
    Q.leafB = __float_as_int(q7.y) != 0;
    return Q;
}

// A lower-cost box distance for pruning only: the same clamped terms, summed with fused
// multiply-adds.  Every term is non-negative, so it and boxDistSq are both within 3 units in the
// last place of the exact sum, i.e. within 2^-21 relative of each other: a walk that prunes with
// `slotDistFast < best * (1 + 2^-19)` (boundAbove) tests everything the exact test keeps.
__device__ __forceinline__ float slotDistFast(const Quad& Q, int i, float px, float py, float pz)
{
    const float tx = __builtin_amdgcn_fmed3f(px, Q.mn[i][0], Q.mx[i][0]) - px;
    const float ty = __builtin_amdgcn_fmed3f(py, Q.mn[i][1], Q.mx[i][1]) - py;
    const float tz = __builtin_amdgcn_fmed3f(pz, Q.mn[i][2], Q.mx[i][2]) - pz;
    return __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz));
}

// best * (1 + 2^-19): the threshold of slotDistFast tests (+inf above FLT_MAX / (1 + 2^-19); a
// negative best, the opted-out lanes', stays negative)
__device__ __forceinline__ float boundAbove(float best) { return best * 1.0000019073486328f; }

// a value no larger than the exact box distance behind a slotDistFast result `d` (16 units in the
// last place below it), for the stack's lower bounds
__device__ __forceinline__ float belowFast(float d)
{
    const uint32_t b = __float_as_uint(d);
    return __uint_as_float(b > 16u ? b - 16u : 0u);
}

__device__ __forceinline__ float slotDistSq(const Quad& Q, int i, float px, float py, float pz)
{
    return boxDistSq(Q.mn[i][0], Q.mn[i][1], Q.mn[i][2], Q.mx[i][0], Q.mx[i][1], Q.mx[i][2], px, py, pz);
}

// lanes with a < b (ordered compare) as a wave mask, straight from the compare
__device__ __forceinline__ uint64_t ltMask(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, 4); }

// A lower bound of the squared distance from (px, py, pz) to the triangle of leaf slot `slot`, at
// most the leaf test's own result (the FP64 closest point, rounded).  The slot's record holds a unit
// normal n with the range [lo, hi] of n.x over the triangle, and for each edge an in-plane unit
// normal m_e with the maximum of m_e.x over the triangle: the triangle lies in the slab and in
// each half-space, so its distance is at least sqrt(a^2 + b^2), a = distance to the slab, b = the
// largest distance to an edge half-space (n and m_e are orthogonal).  Rounding is covered by
// eps = 2^-20 (|p|_1 + r), r = max |v|_1 over the triangle's vertices (16 units of 2^-24 on every
// magnitude involved, here and in the leaf test's rounding to float), and by the factor 1 - 2^-13
// (the float normals' length and orthogonality, the sums of squares).  A degenerate or non-finite
// triangle has an all-zero record and the bound 0.
// The record is read with scalar loads: `slot` is wave-uniform.
__device__ __forceinline__ float leafLowerBound(const MeshDev& m, int slot, float px, float py, float pz)
{
    const float4 r0 = cload(m.prism, 5 * (size_t)slot), r1 = cload(m.prism, 5 * (size_t)slot + 1);
    const float4 r2 = cload(m.prism, 5 * (size_t)slot + 2), r3 = cload(m.prism, 5 * (size_t)slot + 3);
    const float4 r4 = cload(m.prism, 5 * (size_t)slot + 4);
    const float np = __builtin_fmaf(r0.z, pz, __builtin_fmaf(r0.y, py, r0.x * px));
    const float a = fmaxf(np - r1.x, r0.w - np);
    const float e0 = __builtin_fmaf(r1.w, pz, __builtin_fmaf(r1.z, py, __builtin_fmaf(r1.y, px, -r2.x)));
    const float e1 = __builtin_fmaf(r2.w, pz, __builtin_fmaf(r2.z, py, __builtin_fmaf(r2.y, px, -r3.x)));
    const float e2 = __builtin_fmaf(r3.w, pz, __builtin_fmaf(r3.z, py, __builtin_fmaf(r3.y, px, -r4.x)));
    const float b = fmaxf(e0, fmaxf(e1, e2));
    const float eps = (fabsf(px) + fabsf(py) + fabsf(pz) + r4.y) * 0x1p-20f;
    const float as = fmaxf(a - eps, 0.f), bs = fmaxf(b - eps, 0.f);
    return __builtin_fmaf(as, as, bs * bs) * (1.f - 0x1p-13f);
}

// Stack of a wave-cooperative distance walk.  The node ref of entry e is held by lane e of one
// VGPR (writelane / readlane move wave-uniform values without memory), and each lane's box
// distance of the entry sits in LDS at d[e][lane]: a pop reads one value per lane instead of
// reloading the box and recomputing its distance.  Only the distance's top 16 bits are kept
// (5 KB per wave at depth 24, so LDS does not limit occupancy): for a non-negative float that is
// a lower bound within 2^-7 relative, so a pop can wake a lane whose box lies just past its best
// -- that lane then finds nothing closer (slot and leaf tests are exact) -- and never sleeps a
// lane whose box is closer.
//
// Leaf tests are queued rather than run at the visit: a leaf visit appends (leaf, lane) for its
// active lanes, and a full queue is run one test per lane (flushLeaves), so the FP64 closest point
// runs on all 64 lanes instead of the few whose box was close.  Results meet in best[] by LDS
// atomic minimum.  Pruning between flushes uses a best that may lag behind, which only adds
// visits: every leaf closer than the final best is still tested, so the minimum is unchanged.
template <int D>
struct DistStack {
    uint16_t d[quadStack(D)][64];
    int32_t leaf[64];       // queued leaf tests: leaf slot
    uint8_t owner[64];      //   and the lane that asked
    uint32_t best[64];      // each lane's best squared distance (float bits; >= 0 for searching lanes)
};
__device__ __forceinline__ uint16_t stackDist(float d) { return (uint16_t)(__float_as_uint(d) >> 16); }
__device__ __forceinline__ float stackDist(uint16_t s) { return __uint_as_float((uint32_t)s << 16); }
static_assert(quadStack(32) <= 64, "stack entries are held one per lane");

// Runs the `qn` queued leaf tests, one per lane: lane j takes entry j, fetches its leaf (a vector
// load: entries name different leaves) and its owner's point (cross-lane read), and lowers the
// owner's best.  Returns this lane's best afterwards.
template <int D>
__device__ __forceinline__ float flushLeaves(const MeshDev& m, float px, float py, float pz, DistStack<D>* stk, int qn)
{
    const int lane = threadIdx.x & 63;
    const int owner = stk->owner[lane];
    const float ox = __shfl(px, owner), oy = __shfl(py, owner), oz = __shfl(pz, owner);
    if (lane < qn) {
        const float d = leafDistSq(loadLeafLanes(m.tri, stk->leaf[lane]), ox, oy, oz);
        atomicMin(&stk->best[owner], __float_as_uint(d));
    }
    return __uint_as_float(stk->best[lane]);
}

// Squared distance to the nearest leaf (calcDistanceSq / findProjectionSubtree distance part,
// source/MRCuda/MRCudaFastWindingNumber.cu:65-133) for each lane, on a wave-shared walk over the
// 4-wide records: the minimum does not depend on the visit order, so it equals the per-lane
// result bit for bit.  A step tests the four grandchildren of a node; the nearest active one (for
// the lowest active lane) continues, the other active ones are pushed farthest first.
// The walk of coopDistance from node `cur` (box distance `curd`), for lanes whose best so far is
// `best`; after `budget` visited steps (0 = no limit) it stops and sets `stopped` (wave-uniform),
// returning the best found so far -- a real candidate, so a later walk can finish from it.
// Pruning compares slotDistFast distances with boundAbove(best), a superset of the exact tests:
// it may add visits, never drop one.  Only the leaf tests decide the result.
template <int D, bool FILTER = true>
__device__ __forceinline__ float coopDistanceWalk(const MeshDev& m, float px, float py, float pz, float best,
                                                  int cur, float curd, DistStack<D>* stk, int budget, bool& stopped)
{
    const int lane = threadIdx.x & 63;
    stopped = false;
    int sp = 0, steps = 0, qn = 0;
    uint32_t nQuad = 0, nLeaf = 0, nTest = 0;     // measurement counts
    int refs = 0;      // lane e: node ref of stack entry e
    bool have = true;
    stk->best[lane] = __float_as_uint(best);
    float bestHi = boundAbove(best);     // every pruning test compares against this (slotDistFast)
    for (;;) {
        if (!have) {
            if (sp == 0) break;
            --sp;
            cur = __builtin_amdgcn_readlane(refs, sp);
            curd = stackDist(stk->d[sp][lane]);
        }
        have = false;
        uint64_t act = ltMask(curd, bestHi);
        if (!act) continue;
        if (budget && ++steps > budget) {
            stopped = true;
            break;
        }
        if (cur < 0) {
            // lanes whose triangle may still lie closer than their best (the leaf test decides)
            const uint64_t may =
                FILTER ? ~(uint64_t)__builtin_amdgcn_fcmpf(leafLowerBound(m, ~cur, px, py, pz), bestHi, 2) : ~0ull;
            nLeaf += FILTER ? 1u : 0u;
            act &= may;
            if (qn + __popcll(act) > 64) {
                nTest += qn;
                best = flushLeaves<D>(m, px, py, pz, stk, qn);
                bestHi = boundAbove(best);
                qn = 0;
                act = ltMask(curd, bestHi) & may;
            }
            if ((act >> lane) & 1ull) {
                const int r = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                stk->leaf[qn + r] = ~cur;
                stk->owner[qn + r] = (uint8_t)lane;
            }
            qn += __popcll(act);
            continue;
        }
        const Quad Q = loadQuad(m, cur);
        ++nQuad;
        float d[4];
        uint64_t b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            d[i] = slotDistFast(Q, i, px, py, pz);
            b[i] = ltMask(d[i], bestHi);
        }
        if (Q.leafA) b[1] = 0;     // empty slots
        if (Q.leafB) b[3] = 0;
        const uint64_t any = b[0] | b[1] | b[2] | b[3];
        if (!any) continue;
        // order keys: the lowest active lane's distance bits (non-negative floats order like their
        // bits, all below 2^31) with the slot index in the two low bits; 0 for inactive slots
        const int lead = __builtin_ctzll(any);
        uint32_t k[4];
        int cnt = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            k[i] = b[i] ? ((__float_as_uint(laneValue(d[i], lead)) & ~3u) | (uint32_t)i) : 0u;
            cnt += b[i] ? 1 : 0;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!b[i]) continue;
            int farther = 0;     // active slots farther than slot i (an inactive key 0 never counts)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j != i) farther += (int)((k[i] - k[j]) >> 31);
            if (farther == cnt - 1) {
                cur = Q.ref[i];
                curd = d[i];
            } else {
                const int pos = sp + farther;
                refs = mi_writeLane(Q.ref[i], pos, refs);
                stk->d[pos][lane] = stackDist(belowFast(d[i]));
            }
        }
        sp += cnt - 1;
        have = true;
    }
    if (qn) best = flushLeaves<D>(m, px, py, pz, stk, qn);
    if (m.stats) {
        statAdd(m, kStatWalks, 1);
        statAdd(m, kStatQuadSteps, nQuad);
        statAdd(m, kStatPrismReads, nLeaf);
        statAdd(m, kStatLeafTests, nTest + (uint32_t)qn);
    }
    return best;
}

// Squared distance to the nearest leaf (calcDistanceSq / findProjectionSubtree distance part,
// source/MRCuda/MRCudaFastWindingNumber.cu:65-133) for each lane, on a wave-shared walk over the
// 4-wide records: the minimum does not depend on the visit order, so it equals the per-lane
// result bit for bit.  A step tests the four grandchildren of a node; the nearest active one (for
// the lowest active lane) continues, the other active ones are pushed farthest first.
template <int D, bool FILTER = true>
__device__ __forceinline__ float coopDistance(const MeshDev& m, float px, float py, float pz, float up,
                                              DistStack<D>* stk, int budget = 0, bool* stopped = nullptr)
{
    float best = up;
    if (stopped) *stopped = false;
    if (m.numFaces <= 0) return best;
    if (up > 0) {
        const float g = greedyLeafDistSq(m, px, py, pz);
        if (g < best) best = g;
    }
    statAdd(m, kStatSeeds, 1);
    statAdd(m, kStatSeedLevels, (uint32_t)m.depth);
    const float rootd = boxDistSq(m.rootMin[0], m.rootMin[1], m.rootMin[2], m.rootMax[0], m.rootMax[1],
                                  m.rootMax[2], px, py, pz);
    bool st = false;
    best = coopDistanceWalk<D, FILTER>(m, px, py, pz, best, m.rootRef, rootd, stk, budget, st);
    if (stopped) *stopped = st;
    return best;
}

// Winding number of each lane's point: the traversal order of calcFastWindingNumber is left-first
// for every query, so a wave walks one left-first traversal and each entry carries the mask of
// lanes that still descend there.  Every lane visits exactly its own node sequence in its own
// order, so the sums are bit-identical to the per-lane traversal.  valid = false opts a lane out.
__device__ __forceinline__ float coopWinding(const MeshDev& m, float qx, float qy, float qz, bool valid,
                                             float beta, int32_t skipFace, uint4* stk)
{
    if (m.numFaces <= 0) return 0;
    const float betaSq = beta * beta;
    const int lane = threadIdx.x & 63;
    float res = 0;
    uint64_t curMask = waveBallot(valid);
    int cur = m.rootRef;
    int sp = 0;
    uint32_t nVisit = 0, nDesc = 0, nLeaf = 0;     // measurement counts
    while (curMask) {
        const bool mine = (curMask >> lane) & 1;
        cur = uniform(cur);
        ++nVisit;
        float4 d0, d1;
        if (cur >= 0) { d0 = cload(m.dint, 2 * cur); d1 = cload(m.dint, 2 * cur + 1); }
        else { d0 = cload(m.dleaf, 2 * (~cur)); d1 = cload(m.dleaf, 2 * (~cur) + 1); }
        const float dpx = d0.x - qx, dpy = d0.y - qy, dpz = d0.z - qz;
        const float dd = dpx * dpx + dpy * dpy + dpz * dpz;
        const bool approx = !(dd <= betaSq * d0.w);
        if (mine && approx) {
            const float dist = sqrtf(dd);
            if (dist > 0) res += (dpx * d1.x + dpy * d1.y + dpz * d1.z) / (dist * dd);
        }
        const uint64_t desc = waveBallot(mine && !approx);
        if (desc) {
            if (cur >= 0) {
                const float4 n3 = cload(m.dnode, 4 * cur + 3);
                stk[sp] = make_uint4(__float_as_uint(n3.y), (uint32_t)desc, (uint32_t)(desc >> 32), 0u);
                ++sp;
                ++nDesc;
                cur = __float_as_int(n3.x);
                curMask = desc;
                continue;
            }
            ++nLeaf;
            const Leaf L = loadLeaf(m.tri, ~cur);
            if (mine && !approx && L.face != skipFace) res += solidAngle(qx, qy, qz, L);
        }
        if (sp == 0) break;
        --sp;
        const uint4 e = stk[sp];
        cur = (int)e.x;
        curMask = (uint64_t)e.y | ((uint64_t)e.z << 32);
    }
    if (m.stats) {
        statAdd(m, kStatWalks, 1);
        statAdd(m, kStatFwnVisits, nVisit);
        statAdd(m, kStatFwnDescents, nDesc);
        statAdd(m, kStatFwnLeaves, nLeaf);
    }
    return kInv4Pi * res;
}

// Closest point for each lane (findProjectionSubtree, source/MRMesh/MRMeshProject.cpp:13-103) on a
// wave-shared walk over the 4-wide records.  The reference's answer is the leaf that minimises
// (distSq, k) where k is the leaf's position in that lane's own depth-first order (nearer child
// first, ties to the right): candidates are accepted only on a strict '<' and, while the best is
// still above the minimum, no ancestor of the first minimal leaf can be pruned.  k is the lane's
// path of near(0)/far(1) choices, MSB first (bit 31 - depth for a child one level below `depth`),
// so comparing keys compares positions in the lane's order.  Pruning and acceptance both compare
// (distSq, key) lexicographically, which makes the record bit-identical to the reference for any
// visit order.  A slot passes only if its parent would have passed, so the 4-wide step never
// visits what the binary walk would have pruned.
// Valid when loDistLimitSq <= 0 (the reference's early exit then only fires at distSq == 0,
// which is the minimum) and for trees of depth <= 32.
// Keys are not stored per lane: a popped entry is related to the slot popped or continued just
// before it in the same step -- its sibling (flip the entry's bit, clear below), or a child of
// the other child (flip the bit one level up, then the entry's own bit from a 64-lane mask).
// The walk of coopProject from node `cur` at depth `curDepth` (box distance `curd`, per-lane key
// prefix `curKey`), improving each lane's (res.distSq, bestKey); after `budget` visited steps (0 =
// no limit) it stops and sets `stopped` (wave-uniform).
// Stack and leaf-test queue of a projection walk (see DistStack).  Queued tests carry the owner's
// key of the leaf; results meet in best[] as (distSq bits << 32 | key) by 64-bit LDS atomic minimum
// -- the order that decides the record -- and win[] keeps the leaf behind each lane's best.
template <int D>
struct ProjStack {
    uint16_t d[quadStack(D)][64];     // each lane's box distance of entry e (top 16 bits: a lower bound)
    int32_t leaf[64];
    uint32_t key[64];
    uint8_t owner[64];
    unsigned long long best[64];
    int32_t win[64];                  // leaf of the lane's best found in this walk, -1 before
};

// The order of candidates under loDistLimitSq (findProjectionSubtree, source/MRMesh/MRMeshProject.cpp:78-85):
// the reference accepts a leaf on a strict '<' and stops at the first accepted one at or below lo.  With
// loE = min(lo, the largest float below the search limit), every accepted leaf at or below lo is the
// first such leaf of the lane's depth-first order, so the result is the leaf minimising
// (max(d, loE), key): all candidates <= loE tie and the lowest key (the first visited) wins, and
// without such a candidate the order is (d, key) as before.  loE <= 0 leaves d unchanged.  A NaN d
// stays NaN (never accepted, as in the reference).
__device__ __forceinline__ float effectiveLo(float lo, float lim)
{
    if (!(lo > 0) || !(lim > 0)) return 0.f;
    const float below = __uint_as_float(__float_as_uint(lim) - 1u);     // largest float < lim (lim > 0)
    return lo < below ? lo : below;
}
__device__ __forceinline__ float loOrder(float d, float loE) { return d < loE ? loE : d; }

// Runs the `qn` queued projection leaf tests, one per lane, and refreshes this lane's (bestD, bestKey);
// candidates compare by (loOrder(d, owner's loE), key), see effectiveLo.
template <int D>
__device__ __forceinline__ void flushProjLeaves(const MeshDev& m, float px, float py, float pz, float loE,
                                                ProjStack<D>* stk, int qn, float& bestD, uint32_t& bestKey)
{
    const int lane = threadIdx.x & 63;
    const int owner = stk->owner[lane];
    const float ox = __shfl(px, owner), oy = __shfl(py, owner), oz = __shfl(pz, owner);
    const float oLo = __shfl(loE, owner);
    unsigned long long packed = ~0ull;
    int k = 0;
    if (lane < qn) {
        k = stk->leaf[lane];
        const float d = loOrder(leafDistSq(loadLeafLanes(m.tri, k), ox, oy, oz), oLo);
        packed = ((unsigned long long)__float_as_uint(d) << 32) | stk->key[lane];
        atomicMin(&stk->best[owner], packed);
    }
    // (distSq, key) names one leaf of one lane's order, so exactly the winning entry matches
    if (lane < qn && stk->best[owner] == packed) stk->win[owner] = k;
    const unsigned long long b = stk->best[lane];
    bestD = __uint_as_float((uint32_t)(b >> 32));
    bestKey = (uint32_t)b;
}

// lanes with a <= b (ordered compare)
__device__ __forceinline__ uint64_t leMask(float a, float b) { return __builtin_amdgcn_fcmpf(a, b, 5); }

// the `<=` pruning limit of a projection walk: its best, with +inf lowered to FLT_MAX so that an
// empty quad slot (+inf box) never passes; every finite box distance still does
__device__ __forceinline__ float pruneLimit(float best) { return best == INFINITY ? FLT_MAX : best; }

// record = false leaves res's point, face and barycentrics unset (res.face >= 0 still marks an
// improvement): for callers that keep only (distSq, key).
// loE (per lane, see effectiveLo): res.distSq and the merged bests are in loOrder space; with record the
// winning leaf's record (and its true squared distance) is recomputed into res.
template <int D>
__device__ __forceinline__ void coopProjectWalk(const MeshDev& m, float px, float py, float pz, float loE, int cur,
                                                int curDepth, float curd, uint32_t curKey, ProjStack<D>* stk,
                                                BestHit& res, uint32_t& bestKey, int budget, bool& stopped,
                                                bool record = true)
{
    stopped = false;
    const int lane = threadIdx.x & 63;
    int sp = 0, steps = 0, qn = 0;
    uint32_t nStep = 0, nTest = 0;     // measurement counts
    stk->best[lane] = ((unsigned long long)__float_as_uint(res.distSq) << 32) | bestKey;
    stk->win[lane] = -1;
    float bestT = pruneLimit(res.distSq);
    // lane e of each: stack entry e's node ref, depth | 0x100 (relation: child of the other child),
    // and for that relation the 64-lane mask of the entry's own key bit
    int refs = 0, infos = 0, masksLo = 0, masksHi = 0;
    bool have = true;
    for (;;) {
        if (!have) {
            if (sp == 0) break;
            --sp;
            cur = __builtin_amdgcn_readlane(refs, sp);
            curd = stackDist(stk->d[sp][lane]);
            const int info = __builtin_amdgcn_readlane(infos, sp);
            curDepth = info & 0xff;                        // >= 1
            const uint32_t bit = 1u << (32 - curDepth);
            if (info & 0x100) {
                const uint32_t pbit = bit << 1;
                curKey = (curKey & ~(pbit - 1u)) ^ pbit;
                const uint64_t mask = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(masksLo, sp) |
                                      ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(masksHi, sp) << 32);
                if ((mask >> lane) & 1ull) curKey |= bit;
            } else {
                curKey = (curKey & ~(bit - 1u)) ^ bit;
            }
        }
        have = false;
        // popped distances are lower bounds: the test may keep a lane whose box lies just past its
        // best (its slot and leaf tests are exact), never drops one that could still win
        // visits keep every (distance, key) below (best, bestKey): distance <= best is a superset
        // (distance ties with a larger key only add visits; the leaf tests order by key exactly)
        uint64_t act = leMask(curd, bestT);
        if (!act) continue;
        if (budget && ++steps > budget) {
            stopped = true;
            break;
        }
        if (cur < 0) {
            if (qn + __popcll(act) > 64) {
                nTest += qn;
                flushProjLeaves<D>(m, px, py, pz, loE, stk, qn, res.distSq, bestKey);
                bestT = pruneLimit(res.distSq);
                qn = 0;
                act = leMask(curd, bestT);
            }
            if ((act >> lane) & 1ull) {
                const int r = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                stk->leaf[qn + r] = ~cur;
                stk->key[qn + r] = curKey;
                stk->owner[qn + r] = (uint8_t)lane;
            }
            qn += __popcll(act);
            continue;
        }
        const Quad Q = loadQuad(m, cur);
        ++nStep;
        // the two children's boxes (the node's binary record) decide this lane's first choice, the
        // grandchildren's its second
        const float4 n0 = cload(m.dnode, 4 * (size_t)cur), n1 = cload(m.dnode, 4 * (size_t)cur + 1);
        const float4 n2 = cload(m.dnode, 4 * (size_t)cur + 2);
        float d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = slotDistSq(Q, i, px, py, pz);
        const float dA = boxDistSq(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, px, py, pz);
        const float dB = boxDistSq(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, px, py, pz);
        const bool nearA = dA < dB;                           // ties: right child first
        const uint32_t bit1 = 1u << (31 - curDepth), bit2 = bit1 >> 1;
        const uint32_t kA = curKey | (nearA ? 0u : bit1), kB = curKey | (nearA ? bit1 : 0u);
        const bool nearA0 = d[0] < d[1], nearB0 = d[2] < d[3];
        uint32_t key[4];
        key[0] = Q.leafA ? kA : (kA | (nearA0 ? 0u : bit2));
        key[1] = kA | (nearA0 ? bit2 : 0u);
        key[2] = Q.leafB ? kB : (kB | (nearB0 ? 0u : bit2));
        key[3] = kB | (nearB0 ? bit2 : 0u);
        // a slot's (distance, key) is never below its parent's, so a slot that passes implies that
        // the binary walk would have reached it
        uint64_t b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = leMask(d[i], bestT);
        const uint64_t any = b[0] | b[1] | b[2] | b[3];
        if (!any) continue;
        // order keys as in coopDistanceWalk: lowest active lane's distance bits, slot index below
        const int lead = __builtin_ctzll(any);
        uint32_t k[4];
        int cnt = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            k[i] = b[i] ? ((__float_as_uint(laneValue(d[i], lead)) & ~3u) | (uint32_t)i) : 0u;
            cnt += b[i] ? 1 : 0;
        }
        int rank[4];
        uint32_t sideAt = 0;      // side (0: under the left child, 1: right) of the slot at each rank
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int farther = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j != i) farther += (int)((k[i] - k[j]) >> 31);
            rank[i] = cnt - 1 - farther;
            if (b[i]) sideAt |= (uint32_t)(i >> 1) << rank[i];
        }
        const int baseDepth = curDepth;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!b[i]) continue;
            const int side = i >> 1;
            const bool leafSide = side ? Q.leafB : Q.leafA;
            const int depthI = baseDepth + (leafSide ? 1 : 2);
            if (rank[i] == 0) {
                cur = Q.ref[i];
                curd = d[i];
                curKey = key[i];
                curDepth = depthI;
                continue;
            }
            // relation to the slot handled just before this one
            const int prevSide = (int)((sideAt >> (rank[i] - 1)) & 1u);
            int info = depthI;
            const int pos = sp + cnt - 1 - rank[i];
            if (prevSide != side && !leafSide) {
                info |= 0x100;
                const uint64_t mask = __builtin_amdgcn_uicmp(key[i] & bit2, 0u, 33);
                masksLo = mi_writeLane((int)(uint32_t)mask, pos, masksLo);
                masksHi = mi_writeLane((int)(uint32_t)(mask >> 32), pos, masksHi);
            }
            refs = mi_writeLane(Q.ref[i], pos, refs);
            infos = mi_writeLane(info, pos, infos);
            stk->d[pos][lane] = stackDist(d[i]);
        }
        sp += cnt - 1;
        have = true;
    }
    if (qn) flushProjLeaves<D>(m, px, py, pz, loE, stk, qn, res.distSq, bestKey);
    if (m.stats) {
        statAdd(m, kStatWalks, 1);
        statAdd(m, kStatKeyedSteps, nStep);
        statAdd(m, kStatKeyedTests, nTest + (uint32_t)qn);
    }
    // the record of the winning leaf (its distance is the one just merged, recomputed identically)
    const int wl = stk->win[lane];
    if (wl >= 0) {
        if (record) {
            const Leaf L = loadLeafLanes(m.tri, wl);
            const double pd[3] = { px, py, pz };
            const double ad[3] = { L.a[0], L.a[1], L.a[2] };
            const double bd[3] = { L.b[0], L.b[1], L.b[2] };
            const double cd[3] = { L.c[0], L.c[1], L.c[2] };
            double pr[3], ba, bb;
            closestPointD(pd, ad, bd, cd, pr, ba, bb);
            res.face = L.face;
            res.px = (float)pr[0]; res.py = (float)pr[1]; res.pz = (float)pr[2];
            res.a = (float)ba; res.b = (float)bb;
            // the leaf test's value; a stopped walk keeps the loOrder value its hand-over merges with
            if (!stopped) {
                const float dx = res.px - px, dy = res.py - py, dz = res.pz - pz;
                res.distSq = dx * dx + dy * dy + dz * dz;
            }
        } else {
            res.face = 0;
        }
    }
}

// lo: the caller's loDistLimitSq (<= 0: none); the seed, pruning and hand-over bests are in loOrder
// space (see effectiveLo), the record's distSq is the winner's true squared distance.
template <int D>
__device__ __forceinline__ void coopProject(const MeshDev& m, float px, float py, float pz, float up, bool valid,
                                            ProjStack<D>* stk, BestHit& res, uint32_t* bestKeyOut = nullptr,
                                            int budget = 0, bool* stopped = nullptr, float lo = 0.f)
{
    res.distSq = valid ? up : -1.f;
    res.face = -1;
    res.px = res.py = res.pz = 0;
    res.a = res.b = 0;
    uint32_t bestKey = 0;
    if (stopped) *stopped = false;
    if (bestKeyOut) *bestKeyOut = 0;
    if (m.numFaces <= 0) return;
    // search radius just above the greedy leaf's distance: that leaf (and every leaf at the minimum)
    // is still accepted, so the record is unchanged; an empty search keeps the caller's limit
    const float loE = effectiveLo(lo, up);
    statAdd(m, kStatSeeds, 1);
    statAdd(m, kStatSeedLevels, (uint32_t)m.depth);
    if (valid && up > 0) {
        const float g = loOrder(greedyLeafDistSq(m, px, py, pz), loE);
        if (g < up) {
            // any radius above g keeps that leaf; at least FLT_MIN so no comparison meets a denormal
            const float above = g < FLT_MIN ? FLT_MIN : __uint_as_float(__float_as_uint(g) + 1u);
            res.distSq = above < up ? above : up;
        }
    }
    const float rootd = boxDistSq(m.rootMin[0], m.rootMin[1], m.rootMin[2], m.rootMax[0], m.rootMax[1],
                                  m.rootMax[2], px, py, pz);
    bool st = false;
    coopProjectWalk<D>(m, px, py, pz, loE, m.rootRef, 0, rootd, 0u, stk, res, bestKey, budget, st);
    if (stopped) *stopped = st;
    if (bestKeyOut) *bestKeyOut = bestKey;
}

// The cut's subtrees ordered by box distance to the first valid lane's query (ties by index), for
// handed-over wave `slot`: lane c ranks subtree c, the lane of rank r writes order[r].  The lead is
// the first valid lane with finite coordinates, and a distance that is not a number counts as +inf,
// so the ranks are always a permutation of [0, nCut) (a query point may be NaN or infinite: its
// distances are then NaN, which would otherwise give every subtree rank 0).
__device__ __forceinline__ void rankCut(const MeshDev& m, bool valid, float px, float py, float pz, uint8_t* order)
{
    const int lane = threadIdx.x & 63;
    const bool finite = valid && __builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz);
    const uint64_t fm = waveBallot(finite), vm = waveBallot(valid);
    const int lead = fm ? __builtin_ctzll(fm) : vm ? __builtin_ctzll(vm) : 0;
    const float lx = laneValue(px, lead), ly = laneValue(py, lead), lz = laneValue(pz, lead);
    float d = INFINITY;
    if (lane < m.nCut) {
        const CutEntry& E = m.cut[lane];
        d = boxDistSq(E.mn[0], E.mn[1], E.mn[2], E.mx[0], E.mx[1], E.mx[2], lx, ly, lz);
        if (!(d <= INFINITY)) d = INFINITY;
    }
    int rank = 0;
    for (int j = 0; j < m.nCut; ++j) {
        const float dj = laneValue(d, j);
        rank += (dj < d || (dj == d && j < lane)) ? 1 : 0;
    }
    if (lane < m.nCut) order[rank] = (uint8_t)lane;
}

// appends `value` of every lane with p set to list[] (one atomic per wave)
__device__ __forceinline__ void appendWave(bool p, uint32_t value, uint32_t* list, uint32_t* count)
{
    const uint64_t mask = waveBallot(p);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(mask);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    if (p) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// the query of lane `lane` of projection wave `wv` (as in kProjectionsCoop)
__device__ __forceinline__ bool projQuery(const float* __restrict__ pts, const uint32_t* __restrict__ perm, int64_t n,
                                          const Xf& ptXf, int hasPtXf, float up, const float* __restrict__ bound,
                                          uint32_t wv, int lane, int64_t& q, float& px, float& py, float& pz, float& lim)
{
    const int64_t i = (int64_t)wv * 64 + lane;
    const bool valid = i < n;
    q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
    px = py = pz = 0;
    lim = up;
    if (valid) {
        px = pts[3 * q]; py = pts[3 * q + 1]; pz = pts[3 * q + 2];
        if (hasPtXf) xfApply(ptXf, px, py, pz);
        if (bound) { const float b = bound[q]; lim = b < up ? b : up; }
    }
    return valid;
}

// shared tail of the signed-distance kernels: NaN band, winding-number sign, store
// (FastWindingNumber::calcWithDistances, source/MRMesh/MRFastWindingNumber.cpp:69-76)
__device__ __forceinline__ void sdfStore(const MeshDev& m, float px, float py, float pz, bool valid, float resSq,
                                         const DistOpts& o, uint4* stk, float* out, size_t idx)
{
    const bool nan = o.nullOutsideMinMax && (resSq < o.minDistSq || resSq >= o.maxDistSq);
    const float wn = coopWinding(m, px, py, pz, valid && !nan, o.beta, -1, stk);
    if (!valid) return;
    if (nan) {
        out[idx] = __int_as_float(0x7fc00000);
        return;
    }
    const float sign = wn > o.threshold ? -1.f : +1.f;
    out[idx] = sign * sqrtf(resSq);
}

template <typename Fn>
hipError_t dispatchDepth(int D, Fn&& fn)
{
    switch (D) {
    case 8: return fn(std::integral_constant<int, 8>{});
    case 16: return fn(std::integral_constant<int, 16>{});
    case 24: return fn(std::integral_constant<int, 24>{});
    case 32: return fn(std::integral_constant<int, 32>{});
    default: return hipErrorInvalidValue;
    }
}

} // namespace
} // namespace mi
