// mi_kernels_keyed.hip -- gfx950 kernels of the keyed closest-point walks: findProjections and the
// signed distance with a band (minDistSq > 0).
//
// "Keyed": the result is the leaf minimising (order value, key), where the key is the leaf's position
// in the query's own depth-first order (near child first, ties to the right) and the order value is
// the candidate's squared distance raised to loE (effectiveLo): the reference's walk
// (findProjectionSubtree, source/MRMesh/MRMeshProject.cpp:13-103) accepts a leaf on a strict '<' and
// stops at the first accepted leaf at or below loDistLimitSq, so its record is
//   * the first leaf of its order at or below loDistLimitSq, if one is below the search limit;
//   * else the first leaf of its order at the minimum distance.
// Both are the minimum of (max(d, loE), key), which does not depend on the visit order: the shared
// wave walks, the split form's partial walks and their atomic merges all reproduce it exactly.
#include "mi_walk.hpp"

namespace mi {
namespace {

// ---------------------------------------------------------------- per-lane form (reference form)

template <bool XF, int D>
__global__ __launch_bounds__(kBlock) void kProjections(MeshDev m, const float* __restrict__ pts,
                                                       const uint32_t* __restrict__ perm, int64_t n,
                                                       int nLogical, Xf ptXf, int hasPtXf, Xf treeXf,
                                                       float up, float lo, const float* __restrict__ bound,
                                                       ProjRec* __restrict__ out)
{
    __shared__ int2 smem[D * kBlock];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t q = perm ? (int64_t)perm[i] : i;
    float px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    if (hasPtXf) xfApply(ptXf, px, py, pz);
    float lim = up;
    if (bound) { const float b = bound[q]; lim = b < up ? b : up; }
    BestHit h;
    h.px = h.py = h.pz = 0; h.a = h.b = 0;
    distanceDfs<true, XF, D>(m, px, py, pz, lim, lo, treeXf, smem + threadIdx.x, h);
    float4* o = reinterpret_cast<float4*>(out + q);
    o[0] = make_float4(__int_as_float(h.face), h.px, h.py, h.pz);
    o[1] = make_float4(__int_as_float(-1), h.a, h.b, h.distSq);
}

// ---------------------------------------------------------------- query sources of the split passes

// Free points (Morton-sorted through perm): unit = 64-query wave of the first launch.
struct PointSource {
    const float* pts;
    const uint32_t* perm;
    int64_t n;
    Xf ptXf;
    int hasPtXf;
    float up, lo;
    const float* bound;
    __device__ __forceinline__ bool get(uint32_t unit, int lane, int64_t& q, float& px, float& py, float& pz,
                                        float& lim) const
    {
        return projQuery(pts, perm, n, ptXf, hasPtXf, up, bound, unit, lane, q, px, py, pz, lim);
    }
};

// Grid voxels of the z-slab [zBegin, zEnd): unit = 4x4x4 brick; q = slab-relative output index.
struct BrickSource {
    int dx, dy, zBegin, zEnd, nbx, nby, nbz;
    Xf xf;
    float up, lo;
    __device__ __forceinline__ bool get(uint32_t brick, int lane, int64_t& q, float& px, float& py, float& pz,
                                        float& lim) const
    {
        int x, y, z;
        const bool valid = brickVoxel((int)brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z);
        px = (float)x; py = (float)y; pz = (float)z;
        xfApply(xf, px, py, pz);
        q = valid ? ((int64_t)(z - zBegin) * dy + y) * dx + x : 0;
        lim = up;
        return valid;
    }
};

// hands the wave's queries to the split pass: unit id, every lane's (order value, key) and the cut order
__device__ __forceinline__ void handOver(const MeshDev& m, const SplitScratch& sc, uint32_t unit, bool valid, float px,
                                         float py, float pz, float bestE, uint32_t key)
{
    const int lane = threadIdx.x & 63;
    uint32_t slot = 0;
    if (lane == 0) slot = atomicAdd(&sc.counters[0], 1u);
    slot = (uint32_t)uniform((int)slot);
    if (lane == 0) sc.list[slot] = unit;
    reinterpret_cast<unsigned long long*>(sc.best)[(size_t)slot * 64 + lane] =
        ((unsigned long long)__float_as_uint(bestE) << 32) | key;
    rankCut(m, valid, px, py, pz, sc.order + (size_t)slot * 64);
}

// The leaf named by a key: walk down from the root taking, at each level, the child this lane saw as
// near (key bit 0) or far (1); returns the leaf slot, or -1 if the key does not end at a leaf.
__device__ __forceinline__ int keyLeaf(const MeshDev& m, uint32_t key, float px, float py, float pz)
{
    int node = m.rootRef;
    for (int l = 0; node >= 0 && l < 32; ++l) {
        const float4 n0 = cload(m.dnode, 4 * (size_t)node), n1 = cload(m.dnode, 4 * (size_t)node + 1);
        const float4 n2 = cload(m.dnode, 4 * (size_t)node + 2), n3 = cload(m.dnode, 4 * (size_t)node + 3);
        const float d1 = boxDistSq(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, px, py, pz);
        const float d2 = boxDistSq(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, px, py, pz);
        const bool leftNear = d1 < d2;
        const bool far = (key >> (31 - l)) & 1u;
        node = (far == leftNear) ? __float_as_int(n3.y) : __float_as_int(n3.x);
    }
    return node < 0 ? ~node : -1;
}

// ---------------------------------------------------------------- projections

// bound (optional, indexed like the output): a per-query cap on upDistLimitSq.  With rerun lists
// given, a query whose cap was below `up` and that found nothing below the cap is not written but
// appended to the list, to be searched again without the cap (any cap above the true distance
// gives the reference's record, so the combination is exact).  lo: loDistLimitSq.  CAPPED: rerun lists
// given (the capped form); without them a bound is a plain per-query limit (the bounded form).
template <int D, bool CAPPED>
__global__ __launch_bounds__(kBlock) void kProjectionsCoop(MeshDev m, const float* __restrict__ pts,
                                                           const uint32_t* __restrict__ perm, int64_t n,
                                                           int nLogical, Xf ptXf, int hasPtXf, float up, float lo,
                                                           const float* __restrict__ bound,
                                                           uint32_t* __restrict__ rerun,
                                                           uint32_t* __restrict__ rerunCount,
                                                           SplitScratch sc, ProjRec* __restrict__ out)
{
    __shared__ ProjStack<D> stkMem[kWaves];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    const bool valid = i < n;
    if (!waveBallot(valid)) return;
    const int64_t q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
    float px = 0, py = 0, pz = 0;
    float lim = up;
    if (valid) {
        px = pts[3 * q]; py = pts[3 * q + 1]; pz = pts[3 * q + 2];
        if (hasPtXf) xfApply(ptXf, px, py, pz);
        if (bound) { const float b = bound[q]; lim = b < up ? b : up; }
    }
    const int w = threadIdx.x >> 6;
    BestHit h;
    uint32_t key = 0;
    bool stopped = false;
    coopProject<D>(m, px, py, pz, lim, valid, &stkMem[w], h, &key, sc.list ? sc.budget : 0, &stopped, lo);
    // capped: a lane whose cap (below up) has found nothing
    const bool uncapped = CAPPED && valid && h.face < 0 && lim < up;
    uint64_t failed = 0;
    if (CAPPED) {
        failed = waveBallot(uncapped && !stopped);      // searched to the end below a cap too small
    }
    if (stopped || (CAPPED && sc.list && failed)) {
        // capped: a lane with nothing below its cap continues uncapped, from a radius just above the
        // greedy leaf as coopProject starts, so every handed-over record is exact and nothing is searched
        // a second time after the split pass
        float seedE = h.distSq;
        uint32_t seedKey = key;
        if (uncapped) {
            const float loE = effectiveLo(lo, up);
            const float g = loOrder(greedyLeafDistSq(m, px, py, pz), loE);
            const float above = g < FLT_MIN ? FLT_MIN : __uint_as_float(__float_as_uint(g) + 1u);
            seedE = g < up && above < up ? above : up;
            seedKey = 0;
        }
        handOver(m, sc, (uint32_t)(lb * kWaves + w), valid, px, py, pz, seedE, seedKey);
        return;
    }
    // without a split scratch (a tree too small to cut) the failed lanes go to the rerun list
    const bool again = uncapped;
    if (CAPPED) appendWave(again, (uint32_t)q, rerun, rerunCount);
    if (!valid || again) return;
    float4* o = reinterpret_cast<float4*>(out + q);
    o[0] = make_float4(__int_as_float(h.face), h.px, h.py, h.pz);
    o[1] = make_float4(__int_as_float(-1), h.a, h.b, h.distSq);
}

// Second launch of the split form: persistent waves take (handed-over unit, cut subtree) pairs; each
// lane's key prefix for the subtree is its near/far choices along the subtree's path, the walk starts
// from the lane's (order value, key) so far, and improvements are merged with an atomic minimum on
// value bits << 32 | key -- the order that decides the record.
template <int D, class Src>
__global__ __launch_bounds__(kBlock) void kKeyedSplit(MeshDev m, Src src, SplitScratch sc)
{
    __shared__ ProjStack<D> stkMem[kWaves];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    unsigned long long* best = reinterpret_cast<unsigned long long*>(sc.best);
    const uint32_t nHard = sc.counters[0] < sc.cap ? sc.counters[0] : sc.cap;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&sc.counters[1], 1u);
        t = (uint32_t)uniform((int)t);
        // a handed-over unit's subtrees in consecutive tickets, nearest first
        const uint32_t hIdx = t / (uint32_t)m.nCut, r = t % (uint32_t)m.nCut;
        if (hIdx >= nHard) break;
        const uint32_t c = sc.order[(size_t)hIdx * 64 + r];
        if (c >= (uint32_t)m.nCut) continue;     // rankCut writes a permutation of [0, nCut)
        int64_t q;
        float px, py, pz, lim;
        const bool valid = src.get(sc.list[hIdx], lane, q, px, py, pz, lim);
        const float loE = effectiveLo(src.lo, lim);
        const unsigned long long seed = best[(size_t)hIdx * 64 + lane];
        BestHit res;
        res.distSq = valid ? __uint_as_float((uint32_t)(seed >> 32)) : -1.f;
        res.face = -1;
        res.px = res.py = res.pz = res.a = res.b = 0;
        uint32_t bestKey = valid ? (uint32_t)seed : 0u;
        const CutEntry E = loadCut(m, c);
        const float cd = boxDistSq(E.mn[0], E.mn[1], E.mn[2], E.mx[0], E.mx[1], E.mx[2], px, py, pz);
        if (!waveBallot(cd <= res.distSq)) continue;     // no lane can improve inside this subtree
        // key prefix: this lane's near(0)/far(1) choice at every level of the path to the subtree
        uint32_t key = 0;
        int node = m.rootRef;
        for (int l = 0; l < E.depth; ++l) {
            const float4 n0 = cload(m.dnode, 4 * (size_t)node), n1 = cload(m.dnode, 4 * (size_t)node + 1);
            const float4 n2 = cload(m.dnode, 4 * (size_t)node + 2), n3 = cload(m.dnode, 4 * (size_t)node + 3);
            const float d1 = boxDistSq(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, px, py, pz);
            const float d2 = boxDistSq(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, px, py, pz);
            const bool leftNear = d1 < d2;
            const bool right = (E.path >> l) & 1u;
            if (right == leftNear) key |= 1u << (31 - l);
            node = right ? __float_as_int(n3.y) : __float_as_int(n3.x);
        }
        bool st;
        coopProjectWalk<D>(m, px, py, pz, loE, E.ref, E.depth, cd, key, &stkMem[w], res, bestKey, 0, st, false);
        if (valid && res.face >= 0) {
            const unsigned long long packed = ((unsigned long long)__float_as_uint(res.distSq) << 32) | bestKey;
            if (packed < seed) atomicMin(&best[(size_t)hIdx * 64 + lane], packed);
        }
    }
}

// Third launch of the split projections: the handed-over waves' records, from the leaf the winning
// key names (its closest point and true squared distance recomputed exactly as the leaf test).  caps
// (the capped form): failCount counts the queries whose cap (below up) was below their result.
template <int D>
__global__ __launch_bounds__(kBlock) void kProjSplitFinish(MeshDev m, PointSource src, SplitScratch sc,
                                                           const float* __restrict__ caps,
                                                           uint32_t* __restrict__ failCount, ProjRec* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long* best = reinterpret_cast<const unsigned long long*>(sc.best);
    const uint32_t nHard = sc.counters[0] < sc.cap ? sc.counters[0] : sc.cap;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&sc.counters[2], 1u);
        t = (uint32_t)uniform((int)t);
        if (t >= nHard) break;
        int64_t q;
        float px, py, pz, lim;
        const bool valid = src.get(sc.list[t], lane, q, px, py, pz, lim);
        const unsigned long long b = valid ? best[(size_t)t * 64 + lane] : 0ull;
        const float e = __uint_as_float((uint32_t)(b >> 32));
        if (caps) {
            const float c = valid ? caps[q] : src.up;
            const uint64_t f = waveBallot(valid && c < src.up && !(e < c));
            if (f && lane == 0) atomicAdd(failCount, (uint32_t)__popcll(f));
        }
        if (!valid) continue;
        float4* o = reinterpret_cast<float4*>(out + q);
        // a winner has d < lim, i.e. order value < lim (loE < lim); otherwise the reference's empty result
        const int leaf = e < lim ? keyLeaf(m, (uint32_t)b, px, py, pz) : -1;
        if (leaf < 0) {
            o[0] = make_float4(__int_as_float(-1), 0.f, 0.f, 0.f);
            o[1] = make_float4(__int_as_float(-1), 0.f, 0.f, lim);
            continue;
        }
        const Leaf L = loadLeafLanes(m.tri, leaf);
        const double pd[3] = { px, py, pz };
        const double ad[3] = { L.a[0], L.a[1], L.a[2] };
        const double bd[3] = { L.b[0], L.b[1], L.b[2] };
        const double cd[3] = { L.c[0], L.c[1], L.c[2] };
        double pr[3], ba, bb;
        closestPointD(pd, ad, bd, cd, pr, ba, bb);
        const float qx = (float)pr[0], qy = (float)pr[1], qz = (float)pr[2];
        const float dx = qx - px, dy = qy - py, dz = qz - pz;
        o[0] = make_float4(__int_as_float(L.face), qx, qy, qz);
        o[1] = make_float4(__int_as_float(-1), (float)ba, (float)bb, dx * dx + dy * dy + dz * dz);
    }
}

// ---------------------------------------------------------------- signed distance with a band

// calcWithDistances (source/MRMesh/MRFastWindingNumber.cpp:69-76) with minDistSq > 0: the distance is
// findProjection(p, maxDistSq, minDistSq).distSq -- the keyed walk's winner -- then the NaN band and
// the winding-number sign as for the plain grid.  Same brick mapping as kSdfGridCoop.
template <int D>
__global__ __launch_bounds__(kBlock) void kSdfBandGrid(MeshDev m, BrickSource src, int nLogical, DistOpts o,
                                                       SplitScratch sc, float* __restrict__ out)
{
    __shared__ ProjStack<D> stkMem[kWaves];
    const int lb = gridBlock();
    if (lb >= nLogical) return;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t brick = (uint32_t)(lb * kWaves + w);
    int64_t idx;
    float px, py, pz, lim;
    const bool valid = src.get(brick, lane, idx, px, py, pz, lim);
    if (!waveBallot(valid)) return;
    BestHit h;
    uint32_t key = 0;
    bool stopped = false;
    coopProject<D>(m, px, py, pz, lim, valid, &stkMem[w], h, &key, sc.list ? sc.budget : 0, &stopped, src.lo);
    if (stopped) {
        handOver(m, sc, brick, valid, px, py, pz, h.distSq, key);
        return;
    }
    // the winner's true squared distance, or the limit itself when nothing is below it
    sdfStore(m, px, py, pz, valid, h.distSq, o, reinterpret_cast<uint4*>(&stkMem[w]), out, (size_t)idx);
}

// Third launch of the split band grid: the handed-over bricks' distances (the winning leaf's test
// recomputed), then NaN band, sign and store.
template <int D>
__global__ __launch_bounds__(kBlock) void kSdfBandFinish(MeshDev m, BrickSource src, DistOpts o, SplitScratch sc,
                                                         float* __restrict__ out)
{
    __shared__ uint4 stkMem[kWaves][D];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const unsigned long long* best = reinterpret_cast<const unsigned long long*>(sc.best);
    const uint32_t nHard = sc.counters[0] < sc.cap ? sc.counters[0] : sc.cap;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&sc.counters[2], 1u);
        t = (uint32_t)uniform((int)t);
        if (t >= nHard) break;
        int64_t idx;
        float px, py, pz, lim;
        const bool valid = src.get(sc.list[t], lane, idx, px, py, pz, lim);
        const unsigned long long b = best[(size_t)t * 64 + lane];
        const float e = __uint_as_float((uint32_t)(b >> 32));
        float resSq = lim;
        if (valid && e < lim) {
            const int leaf = keyLeaf(m, (uint32_t)b, px, py, pz);
            if (leaf >= 0) resSq = leafDistSq(loadLeafLanes(m.tri, leaf), px, py, pz);
        }
        sdfStore(m, px, py, pz, valid, resSq, o, stkMem[w], out, (size_t)idx);
    }
}

} // namespace

hipError_t launchProjections(const MeshDev& m, const float* pts, const uint32_t* perm, int64_t n,
                             const Xf* ptXf, const Xf* treeXf, float up, float lo, ProjRec* out,
                             int stackD, hipStream_t s, const float* bound, uint32_t* rerun,
                             uint32_t* rerunCount, const SplitScratch* split)
{
    if (n <= 0) return hipSuccess;
    const int nLogical = (int)((n + kBlock - 1) / kBlock);
    Xf px{}, tx{};
    if (ptXf) px = *ptXf;
    if (treeXf) tx = *treeXf;
    // the keyed walk reproduces the reference's record for every loDistLimitSq; a non-rigid tree
    // transform (applied per box) keeps the per-lane form
    const bool coopOk = !treeXf;
    const bool coop = coopOk && (m.traversal == (int)Traversal::WaveCoop || rerun);
    if (rerun && (!coop || !rerunCount || !bound)) return hipErrorInvalidValue;
    SplitScratch sc;
    // the capped form takes the split scratch even with no step budget (its failed waves use it)
    if (coop && split && split->list && (split->budget > 0 || rerun) && m.nCut > 0) {
        if ((int64_t)split->cap < projectionWaves(n)) return hipErrorInvalidValue;
        sc = *split;
        hipError_t e = fillU32(sc.counters, 3, 0u, s);
        if (e != hipSuccess) return e;
    }
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (coop) {
            if (rerun)
                hipLaunchKernelGGL((kProjectionsCoop<D, true>), dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m,
                                   pts, perm, n, nLogical, px, ptXf ? 1 : 0, up, lo, bound, rerun, rerunCount, sc, out);
            else
                hipLaunchKernelGGL((kProjectionsCoop<D, false>), dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m,
                                   pts, perm, n, nLogical, px, ptXf ? 1 : 0, up, lo, bound, rerun, rerunCount, sc, out);
            if (sc.list) {
                // the capped form hands over uncapped searches (the bound no longer applies); the bounded
                // form keeps its per-query limits
                const PointSource src{ pts, perm, n, px, ptXf ? 1 : 0, up, lo, rerun ? nullptr : bound };
                const int blocks = persistentBlocks();
                hipLaunchKernelGGL((kKeyedSplit<D, PointSource>), dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 1),
                                   src, sc);
                hipLaunchKernelGGL(kProjSplitFinish<D>, dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 2), src, sc,
                                   rerun ? bound : nullptr, rerunCount, out);
            }
        } else if (treeXf)
            hipLaunchKernelGGL((kProjections<true, D>), dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s,
                               m, pts, perm, n, nLogical, px, ptXf ? 1 : 0, tx, up, lo, bound, out);
        else
            hipLaunchKernelGGL((kProjections<false, D>), dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s,
                               m, pts, perm, n, nLogical, px, ptXf ? 1 : 0, tx, up, lo, bound, out);
        return hipGetLastError();
    });
}

hipError_t launchSdfGridKeyed(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                              int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s,
                              const SplitScratch* split)
{
    const int nbx = (dims[0] + 3) / 4, nby = (dims[1] + 3) / 4, nbz = (zEnd - zBegin + 3) / 4;
    const int64_t bricks = (int64_t)nbx * nby * nbz;
    if (bricks <= 0) return hipSuccess;
    const int nLogical = (int)((bricks + 3) / 4);
    SplitScratch sc;
    if (split && split->list && split->budget > 0 && m.nCut > 0) {
        if ((int64_t)split->cap < sdfGridWaves(dims, zBegin, zEnd)) return hipErrorInvalidValue;
        sc = *split;
        hipError_t e = fillU32(sc.counters, 3, 0u, s);
        if (e != hipSuccess) return e;
    }
    const BrickSource src{ dims[0], dims[1], zBegin, zEnd, nbx, nby, nbz, xf, o.maxDistSq, o.minDistSq };
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        hipLaunchKernelGGL(kSdfBandGrid<D>, dim3(chunkedGrid(nLogical)), dim3(kBlock), 0, s, m, src, nLogical, o, sc,
                           out);
        if (sc.list) {
            const int blocks = persistentBlocks();
            hipLaunchKernelGGL((kKeyedSplit<D, BrickSource>), dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 1), src,
                               sc);
            hipLaunchKernelGGL(kSdfBandFinish<D>, dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 2), src, o, sc, out);
        }
        return hipGetLastError();
    });
}

} // namespace mi
