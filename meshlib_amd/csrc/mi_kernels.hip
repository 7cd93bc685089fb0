// mi_kernels.hip -- gfx950 kernels of the distance-field path.
//
// Compiled with -ffp-contract=off: the box distance, the FP64 closest point, the
// dipole test and the solid angle are evaluated with the reference CPU build's
// rounding and operation order, so distances and projections are bit-identical to
// the CPU path (source/MRMesh/MRMeshProject.cpp:13-103) and winding numbers differ
// only by atan2f ulps.
//
// Execution model (one lane = one query, 64-lane waves):
//  * queries of a wave are spatially coherent: grid voxels are mapped as 4x4x4
//    bricks per wave; free points are Morton-sorted first, so the lanes of a wave
//    walk mostly the same nodes and their loads hit the same cache lines;
//  * per-lane form (kSdfGrid, kProjections, ...): each lane keeps the reference's
//    depth-first order exactly (nearer child first, strict '<' on distances), with
//    the pending far children on a per-lane stack in LDS laid out [depth][lane];
//  * wave-cooperative form (the default; coopDistanceWalk, coopProjectWalk,
//    coopWinding): one walk per wave over 4-wide records with wave-uniform node refs
//    (scalar loads), a stack whose uniform part is held in lanes of a few VGPRs and
//    whose per-lane part (16-bit lower bounds of box distances) sits in LDS, and leaf
//    tests queued and run one per lane; results equal the per-lane form's;
//  * split form: waves past a step budget hand their queries to a second launch that
//    walks the tree's top subtrees for them in parallel (results merged by atomic
//    minimum) and a third that stores them;
//  * blocks are remapped so that each XCD (blocks are dealt round-robin over the 8
//    XCDs) gets one contiguous range of bricks / sorted queries, keeping each L2's
//    working set spatially local.
#include "mi_walk.hpp"

#include <hipcub/hipcub.hpp>

namespace mi {
namespace {

template <int D>
__global__ __launch_bounds__(kBlock) void kSdfGrid(MeshDev m, int dx, int dy, int zBegin, int zEnd,
                                                   int nbx, int nby, int nbz, int nLogical, Xf xf,
                                                   DistOpts o, float* __restrict__ out)
{
    __shared__ int2 smem[D * kBlock];
    const int lb = gridBlock();
    if (lb >= nLogical) return;
    const int lane = threadIdx.x & 63;
    const int brick = lb * (kBlock / 64) + (threadIdx.x >> 6);
    int x, y, z;
    if (!brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z)) return;
    float px = (float)x, py = (float)y, pz = (float)z;
    xfApply(xf, px, py, pz);
    int2* stk = smem + threadIdx.x;
    BestHit h;
    distanceDfs<false, false, D>(m, px, py, pz, o.maxDistSq, o.minDistSq, xf, stk, h);
    const size_t idx = ((size_t)(z - zBegin) * dy + y) * dx + x;
    const float resSq = h.distSq;
    // FastWindingNumber::calcWithDistances (source/MRMesh/MRFastWindingNumber.cpp:69-76)
    if (o.nullOutsideMinMax && (resSq < o.minDistSq || resSq >= o.maxDistSq)) {
        out[idx] = __int_as_float(0x7fc00000);
        return;
    }
    const float wn = windingNumber(m, px, py, pz, o.beta, -1, stk);
    const float sign = wn > o.threshold ? -1.f : +1.f;
    out[idx] = sign * sqrtf(resSq);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kFwnGrid(MeshDev m, int dx, int dy, int zBegin, int zEnd,
                                                   int nbx, int nby, int nbz, int nLogical, Xf xf,
                                                   float beta, float* __restrict__ out)
{
    __shared__ int2 smem[D * kBlock];
    const int lb = gridBlock();
    if (lb >= nLogical) return;
    const int lane = threadIdx.x & 63;
    const int brick = lb * (kBlock / 64) + (threadIdx.x >> 6);
    int x, y, z;
    if (!brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z)) return;
    float px = (float)x, py = (float)y, pz = (float)z;
    xfApply(xf, px, py, pz);
    const size_t idx = ((size_t)(z - zBegin) * dy + y) * dx + x;
    out[idx] = windingNumber(m, px, py, pz, beta, -1, smem + threadIdx.x);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kFwnPoints(MeshDev m, const float* __restrict__ pts,
                                                     const uint32_t* __restrict__ perm, int64_t n,
                                                     int nLogical, float beta, int32_t skipFace,
                                                     float* __restrict__ out)
{
    __shared__ int2 smem[D * kBlock];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t q = perm ? (int64_t)perm[i] : i;
    const float px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    out[q] = windingNumber(m, px, py, pz, beta, skipFace, smem + threadIdx.x);
}

// FastWindingNumber::calcSelfIntersections (source/MRMesh/MRFastWindingNumber.cpp:40-52): one
// lane per leaf slot (tree order = spatial order), query = triCenter, own face skipped
template <int D>
__global__ __launch_bounds__(kBlock) void kSelfIntersections(MeshDev m, int nLogical, float beta,
                                                             uint8_t* __restrict__ out)
{
    __shared__ int2 smem[D * kBlock];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int k = lb * kBlock + threadIdx.x;
    if (k >= m.numFaces) return;
    const Leaf L = loadLeaf(m.tri, k);
    const float third = 1 / 3.0f;
    const float qx = third * ((L.a[0] + L.b[0]) + L.c[0]);
    const float qy = third * ((L.a[1] + L.b[1]) + L.c[1]);
    const float qz = third * ((L.a[2] + L.b[2]) + L.c[2]);
    const float wn = windingNumber(m, qx, qy, qz, beta, L.face, smem + threadIdx.x);
    out[L.face] = (wn < 0 || wn > 1) ? 1 : 0;
}

// At most 72 VGPRs: seven waves per SIMD (the LDS of seven blocks fits a CU); the walks are bound by
// the latency of their dependent scalar loads, and the seventh wave hides more of it than the
// registers it costs (signed distance 105 -> 98 ms at k=8 256^3).
// CAPPED: the capped form (bound and rerun lists given).
template <int D, bool FILTER, bool CAPPED>
__global__ __launch_bounds__(kBlock, 7) void kSdfGridCoop(MeshDev m, int dx, int dy, int zBegin, int zEnd,
                                                       int nbx, int nby, int nbz, int nLogical, Xf xf,
                                                       DistOpts o, const float* __restrict__ bound,
                                                       uint32_t* __restrict__ rerun,
                                                       uint32_t* __restrict__ rerunCount,
                                                       SplitScratch sc, float* __restrict__ out)
{
    __shared__ DistStack<D> stkMem[kWaves];
    const int lb = gridBlock();
    if (lb >= nLogical) return;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int brick = lb * kWaves + w;
    int x, y, z;
    const bool valid = brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z);
    if (!waveBallot(valid)) return;
    float px = (float)x, py = (float)y, pz = (float)z;
    xfApply(xf, px, py, pz);
    const size_t idx = valid ? ((size_t)(z - zBegin) * dy + y) * dx + x : 0;
    float lim = o.maxDistSq;
    if (CAPPED && valid) { const float b = bound[idx]; lim = b < lim ? b : lim; }
    bool stopped = false;
    const float resSq = coopDistance<D, FILTER>(m, px, py, pz, valid ? lim : -1.f, &stkMem[w],
                                                sc.list ? sc.budget : 0, &stopped);
    // capped: a lane whose cap (below the plain limit) has found nothing
    const bool uncapped = CAPPED && valid && !(resSq < lim) && lim < o.maxDistSq;
    uint64_t failed = 0;
    if (CAPPED) {
        failed = waveBallot(uncapped && !stopped);      // searched to the end below a cap too small
    }
    if (stopped || (CAPPED && sc.list && failed)) {
        // hand the brick to the split pass with every lane's best so far.  Capped: a lane with nothing
        // below its cap continues uncapped from the greedy leaf (a real candidate), so every handed-over
        // result is exact and nothing is searched a second time after the split pass.
        float seed = resSq;
        if (uncapped) {
            const float g = greedyLeafDistSq(m, px, py, pz);
            seed = g < o.maxDistSq ? g : o.maxDistSq;
        }
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(&sc.counters[0], 1u);
        slot = (uint32_t)uniform((int)slot);
        if (lane == 0) sc.list[slot] = (uint32_t)brick;
        sc.best[(size_t)slot * 64 + lane] = __float_as_uint(valid ? seed : -1.f);
        rankCut(m, valid, px, py, pz, sc.order + (size_t)slot * 64);
        return;
    }
    // without a split scratch (a tree too small to cut) the failed lanes go to the rerun list
    const bool again = CAPPED && uncapped;
    if (CAPPED) appendWave(again, (uint32_t)idx, rerun, rerunCount);
    sdfStore(m, px, py, pz, valid && !again, resSq, o, reinterpret_cast<uint4*>(&stkMem[w]), out, idx);
}

// Second launch of the split form: persistent waves take (handed-over brick, cut subtree) pairs and
// walk that subtree for the brick's voxels from their best so far; improvements are merged by an
// atomic minimum on the float bits (non-negative floats order like their bit patterns).
template <int D, bool FILTER>
__global__ __launch_bounds__(kBlock) void kSdfSplit(MeshDev m, int dx, int dy, int zBegin, int zEnd, int nbx,
                                                      int nby, int nbz, Xf xf, SplitScratch sc)
{
    __shared__ DistStack<D> stkMem[kWaves];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t nHard = sc.counters[0] < sc.cap ? sc.counters[0] : sc.cap;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&sc.counters[1], 1u);
        t = (uint32_t)uniform((int)t);
        // a handed-over wave's subtrees in consecutive tickets (its queries and their tree region stay
        // in cache), nearest first
        const uint32_t h = t / (uint32_t)m.nCut, round = t % (uint32_t)m.nCut;
        if (h >= nHard) break;
        const uint32_t c = sc.order[(size_t)h * 64 + round];
        if (c >= (uint32_t)m.nCut) continue;     // rankCut writes a permutation of [0, nCut)
        const int brick = (int)sc.list[h];
        int x, y, z;
        const bool valid = brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z);
        float px = (float)x, py = (float)y, pz = (float)z;
        xfApply(xf, px, py, pz);
        const float best = valid ? __uint_as_float(sc.best[(size_t)h * 64 + lane]) : -1.f;
        const CutEntry E = loadCut(m, c);
        const float cd = boxDistSq(E.mn[0], E.mn[1], E.mn[2], E.mx[0], E.mx[1], E.mx[2], px, py, pz);
        bool st;
        const float r = coopDistanceWalk<D, FILTER>(m, px, py, pz, best, E.ref, cd, &stkMem[w], 0, st);
        if (valid && r < best) atomicMin(&sc.best[(size_t)h * 64 + lane], __float_as_uint(r));
    }
}

// Third launch: the handed-over bricks' results (sign from the winding number, NaN band, store).  Every
// handed-over search is exact: the capped form hands its failed caps over uncapped.  With caps (bound),
// failCount counts the voxels whose cap (below maxDistSq) was below their distance.
template <int D>
__global__ __launch_bounds__(kBlock) void kSdfSplitFinish(MeshDev m, int dx, int dy, int zBegin, int zEnd,
                                                            int nbx, int nby, int nbz, Xf xf, DistOpts o,
                                                            SplitScratch sc, const float* __restrict__ bound,
                                                            uint32_t* __restrict__ failCount, float* __restrict__ out)
{
    __shared__ float4 stkMem[kWaves][2 * quadStack(D)];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t nHard = sc.counters[0] < sc.cap ? sc.counters[0] : sc.cap;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&sc.counters[2], 1u);
        t = (uint32_t)uniform((int)t);
        if (t >= nHard) break;
        const int brick = (int)sc.list[t];
        int x, y, z;
        const bool valid = brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z);
        float px = (float)x, py = (float)y, pz = (float)z;
        xfApply(xf, px, py, pz);
        const size_t idx = valid ? ((size_t)(z - zBegin) * dy + y) * dx + x : 0;
        const float resSq = __uint_as_float(sc.best[(size_t)t * 64 + lane]);
        if (bound) {
            const float b = valid ? bound[idx] : o.maxDistSq;
            const uint64_t f = waveBallot(valid && b < o.maxDistSq && !(resSq < b));
            if (f && lane == 0) atomicAdd(failCount, (uint32_t)__popcll(f));
        }
        sdfStore(m, px, py, pz, valid, resSq, o, reinterpret_cast<uint4*>(stkMem[w]), out, idx);
    }
}

// the listed voxels of a z-slab (ids relative to the slab), searched without any cap
template <int D>
__global__ __launch_bounds__(kBlock) void kSdfVoxelsCoop(MeshDev m, int dx, int dy, int zBegin,
                                                         const uint32_t* __restrict__ list, int64_t count,
                                                         int nLogical, Xf xf, DistOpts o,
                                                         float* __restrict__ out)
{
    __shared__ DistStack<D> stkMem[kWaves];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    const bool valid = i < count;
    if (!waveBallot(valid)) return;
    const size_t idx = valid ? list[i] : 0;
    const size_t layer = (size_t)dx * dy;
    const int z = zBegin + (int)(idx / layer);
    const int rem = (int)(idx % layer);
    float px = (float)(rem % dx), py = (float)(rem / dx), pz = (float)z;
    xfApply(xf, px, py, pz);
    const int w = threadIdx.x >> 6;
    const float resSq = coopDistance<D>(m, px, py, pz, valid ? o.maxDistSq : -1.f, &stkMem[w]);
    sdfStore(m, px, py, pz, valid, resSq, o, reinterpret_cast<uint4*>(&stkMem[w]), out, idx);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kFwnGridCoop(MeshDev m, int dx, int dy, int zBegin, int zEnd,
                                                       int nbx, int nby, int nbz, int nLogical, Xf xf,
                                                       float beta, float* __restrict__ out)
{
    __shared__ uint4 stkMem[kWaves][D];
    const int lb = gridBlock();
    if (lb >= nLogical) return;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int brick = lb * kWaves + w;
    int x, y, z;
    const bool valid = brickVoxel(brick, lane, nbx, nby, nbz, dx, dy, zBegin, zEnd, x, y, z);
    if (!waveBallot(valid)) return;
    float px = (float)x, py = (float)y, pz = (float)z;
    xfApply(xf, px, py, pz);
    const float wn = coopWinding(m, px, py, pz, valid, beta, -1, stkMem[w]);
    if (!valid) return;
    out[((size_t)(z - zBegin) * dy + y) * dx + x] = wn;
}

template <int D>
__global__ __launch_bounds__(kBlock) void kFwnPointsCoop(MeshDev m, const float* __restrict__ pts,
                                                         const uint32_t* __restrict__ perm, int64_t n,
                                                         int nLogical, float beta, int32_t skipFace,
                                                         float* __restrict__ out)
{
    __shared__ uint4 stkMem[kWaves][D];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int64_t i = (int64_t)lb * kBlock + threadIdx.x;
    const bool valid = i < n;
    if (!waveBallot(valid)) return;
    const int64_t q = valid ? (perm ? (int64_t)perm[i] : i) : 0;
    float px = 0, py = 0, pz = 0;
    if (valid) { px = pts[3 * q]; py = pts[3 * q + 1]; pz = pts[3 * q + 2]; }
    const float wn = coopWinding(m, px, py, pz, valid, beta, skipFace, stkMem[threadIdx.x >> 6]);
    if (valid) out[q] = wn;
}

template <int D>
__global__ __launch_bounds__(kBlock) void kSelfIntersectionsCoop(MeshDev m, int nLogical, float beta,
                                                                 uint8_t* __restrict__ out)
{
    __shared__ uint4 stkMem[kWaves][D];
    const int lb = logicalBlock(nLogical);
    if (lb >= nLogical) return;
    const int k = lb * kBlock + threadIdx.x;
    const bool valid = k < m.numFaces;
    if (!waveBallot(valid)) return;
    const Leaf L = loadLeaf(m.tri, valid ? k : 0);
    const float third = 1 / 3.0f;
    const float qx = third * ((L.a[0] + L.b[0]) + L.c[0]);
    const float qy = third * ((L.a[1] + L.b[1]) + L.c[1]);
    const float qz = third * ((L.a[2] + L.b[2]) + L.c[2]);
    const float wn = coopWinding(m, qx, qy, qz, valid, beta, L.face, stkMem[threadIdx.x >> 6]);
    if (valid) out[L.face] = (wn < 0 || wn > 1) ? 1 : 0;
}

// ---------------------------------------------------------------- persistent batched kernel
//
// One launch over many (mesh, grid, point set) items of uneven size.  Work units are one voxel
// brick (64 voxels) or one chunk of 64 Morton-sorted queries; each wave takes units from a ticket
// counter and runs the wave-cooperative traversal of that unit's mesh.  With xcdSplit the unit
// range is cut into 8 contiguous parts, one per XCD (blocks are dealt to XCDs round-robin), and a
// wave whose part is drained moves on to the next part: spatially adjacent units share an L2,
// and the tail is balanced.  Every wave exits once it has seen every part drained.
// CAPPED: the items' caps apply; a wave whose lanes found nothing below a cap smaller than the plain
// limit walks once more for them, uncapped (persistent waves take new units meanwhile, so the retries do
// not form a tail of their own as a separate launch over scattered queries would).
template <int D, bool CAPPED>
__global__ __launch_bounds__(kBlock) void kBatch(const BatchItem* __restrict__ items, int nItems, int64_t total,
                                                 DistOpts o, float up, unsigned long long* __restrict__ tickets,
                                                 int xcdSplit)
{
    union WaveStack {
        DistStack<D> dist;
        ProjStack<D> proj;
        uint4 fwn[D];
    };
    __shared__ WaveStack stkMem[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int parts = xcdSplit ? kXcds : 1;
    int part = xcdSplit ? (int)(blockIdx.x % kXcds) : 0;
    int drained = 0;
    for (;;) {
        const int64_t pBegin = total * part / parts, pEnd = total * (part + 1) / parts;
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(&tickets[part], 1ull);
        const uint32_t tlo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)t, 0);
        const uint32_t thi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(t >> 32), 0);
        const int64_t u = pBegin + (int64_t)(((uint64_t)thi << 32) | tlo);
        if (u >= pEnd) {
            if (++drained >= parts) break;
            part = (part + 1) % parts;
            continue;
        }
        int lo = 0, hi = nItems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].unitBegin <= u) lo = mid; else hi = mid - 1;
        }
        const BatchItem& it = items[uniform(lo)];
        const int64_t lu = u - it.unitBegin;
        if (lu < it.gridUnits) {
            int x = 0, y = 0, z = 0;
            const bool valid = brickVoxel((int)lu, lane, it.nbx, it.nby, it.nbz, it.dx, it.dy, 0, it.dz, x, y, z);
            const size_t idx = valid ? ((size_t)z * it.dy + y) * it.dx + x : 0;
            float px = (float)x, py = (float)y, pz = (float)z;
            xfApply(it.xf, px, py, pz);
            float lim = o.maxDistSq;
            if (CAPPED && it.capSdf && valid) { const float c = it.capSdf[idx]; lim = c < lim ? c : lim; }
            float resSq = coopDistance<D>(it.m, px, py, pz, valid ? lim : -1.f, &stkMem[w].dist);
            if (CAPPED) {
                const bool retry = valid && !(resSq < lim) && lim < o.maxDistSq;
                const uint64_t rm = waveBallot(retry);
                if (rm) {
                    if (lane == 0) atomicAdd(it.rerunCount, (uint32_t)__popcll(rm));
                    const float r = coopDistance<D>(it.m, px, py, pz, retry ? o.maxDistSq : -1.f, &stkMem[w].dist);
                    if (retry) resSq = r;
                }
            }
            sdfStore(it.m, px, py, pz, valid, resSq, o, stkMem[w].fwn, it.sdf, idx);
        } else {
            const int64_t i = (lu - it.gridUnits) * 64 + lane;
            const bool valid = i < it.npts;
            const int64_t q = valid ? (it.perm ? (int64_t)it.perm[i] : i) : 0;
            float px = 0, py = 0, pz = 0;
            float lim = up;
            if (valid) {
                px = it.pts[3 * q]; py = it.pts[3 * q + 1]; pz = it.pts[3 * q + 2];
                if (CAPPED && it.capProj) { const float c = it.capProj[q]; lim = c < up ? c : up; }
            }
            BestHit h;
            coopProject<D>(it.m, px, py, pz, lim, valid, &stkMem[w].proj, h);
            if (CAPPED) {
                const bool retry = valid && h.face < 0 && lim < up;
                const uint64_t rm = waveBallot(retry);
                if (rm) {
                    if (lane == 0) atomicAdd(it.rerunCount + 1, (uint32_t)__popcll(rm));
                    BestHit h2;
                    coopProject<D>(it.m, px, py, pz, up, retry, &stkMem[w].proj, h2);
                    if (retry) h = h2;
                }
            }
            if (valid) {
                float4* op = reinterpret_cast<float4*>(it.proj + q);
                op[0] = make_float4(__int_as_float(h.face), h.px, h.py, h.pz);
                op[1] = make_float4(__int_as_float(-1), h.a, h.b, h.distSq);
            }
        }
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void kMorton(const float* __restrict__ pts, int64_t n, float lx, float ly, float lz,
                        float sx, float sy, float sz, uint32_t* __restrict__ keys,
                        uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float fx = fminf(fmaxf((pts[3 * i] - lx) * sx, 0.f), 1023.f);
    const float fy = fminf(fmaxf((pts[3 * i + 1] - ly) * sy, 0.f), 1023.f);
    const float fz = fminf(fmaxf((pts[3 * i + 2] - lz) * sz, 0.f), 1023.f);
    keys[i] = spread10((uint32_t)fx) | (spread10((uint32_t)fy) << 1) | (spread10((uint32_t)fz) << 2);
    vals[i] = (uint32_t)i;
}


} // namespace

namespace {
// 4-wide record of internal node j (layout read by loadQuad): slots 0,1 = the left child's
// children, 2,3 = the right child's; a leaf child takes slot 0 / 2 itself and leaves slot 1 / 3
// empty (the point box at +inf: distance +inf, never below any limit)
__global__ void kBuildQuads(const float4* __restrict__ dnode, int32_t ni, float4* __restrict__ dquad)
{
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= ni) return;
    const float4 n0 = dnode[4 * (size_t)j], n1 = dnode[4 * (size_t)j + 1];
    const float4 n2 = dnode[4 * (size_t)j + 2], n3 = dnode[4 * (size_t)j + 3];
    float bx[4][6];
    int32_t ref[4];
    int32_t leaf[2];
    const float cb[2][6] = { { n0.x, n0.y, n0.z, n0.w, n1.x, n1.y }, { n1.z, n1.w, n2.x, n2.y, n2.z, n2.w } };
    const int32_t child[2] = { __float_as_int(n3.x), __float_as_int(n3.y) };
    for (int side = 0; side < 2; ++side) {
        const int32_t c = child[side];
        leaf[side] = c < 0 ? 1 : 0;
        if (c >= 0) {
            const float4 c0 = dnode[4 * (size_t)c], c1 = dnode[4 * (size_t)c + 1];
            const float4 c2 = dnode[4 * (size_t)c + 2], c3 = dnode[4 * (size_t)c + 3];
            const float l[6] = { c0.x, c0.y, c0.z, c0.w, c1.x, c1.y };
            const float r[6] = { c1.z, c1.w, c2.x, c2.y, c2.z, c2.w };
            for (int k = 0; k < 6; ++k) { bx[2 * side][k] = l[k]; bx[2 * side + 1][k] = r[k]; }
            ref[2 * side] = __float_as_int(c3.x);
            ref[2 * side + 1] = __float_as_int(c3.y);
        } else {
            for (int k = 0; k < 6; ++k) bx[2 * side][k] = cb[side][k];
            for (int k = 0; k < 6; ++k) bx[2 * side + 1][k] = INFINITY;
            ref[2 * side] = c;
            ref[2 * side + 1] = 0;
        }
    }
    float4* q = dquad + 8 * (size_t)j;
    q[0] = make_float4(bx[0][0], bx[0][1], bx[0][2], bx[0][3]);
    q[1] = make_float4(bx[0][4], bx[0][5], bx[1][0], bx[1][1]);
    q[2] = make_float4(bx[1][2], bx[1][3], bx[1][4], bx[1][5]);
    q[3] = make_float4(bx[2][0], bx[2][1], bx[2][2], bx[2][3]);
    q[4] = make_float4(bx[2][4], bx[2][5], bx[3][0], bx[3][1]);
    q[5] = make_float4(bx[3][2], bx[3][3], bx[3][4], bx[3][5]);
    q[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
    q[7] = make_float4(__int_as_float(leaf[0]), __int_as_float(leaf[1]), 0.f, 0.f);
}

// Subtrees at depth kCutDepth (leaves above it stand for themselves), left to right: one thread
// walks the top of the tree depth-first.
__global__ void kBuildCut(const float4* __restrict__ dnode, CutEntry root, CutEntry* __restrict__ cut,
                          int32_t* __restrict__ nCut)
{
    if (threadIdx.x != 0) return;
    CutEntry stack[kCutDepth + 2];
    int sp = 0, n = 0;
    stack[sp++] = root;
    while (sp > 0) {
        const CutEntry e = stack[--sp];
        if (e.ref < 0 || e.depth == kCutDepth) {
            cut[n++] = e;
            continue;
        }
        const float4 n0 = dnode[4 * (size_t)e.ref], n1 = dnode[4 * (size_t)e.ref + 1];
        const float4 n2 = dnode[4 * (size_t)e.ref + 2], n3 = dnode[4 * (size_t)e.ref + 3];
        CutEntry l = e, r = e;
        l.ref = __float_as_int(n3.x);
        r.ref = __float_as_int(n3.y);
        l.depth = r.depth = e.depth + 1;
        r.path = e.path | (1u << e.depth);
        l.mn[0] = n0.x; l.mn[1] = n0.y; l.mn[2] = n0.z; l.mx[0] = n0.w; l.mx[1] = n1.x; l.mx[2] = n1.y;
        r.mn[0] = n1.z; r.mn[1] = n1.w; r.mn[2] = n2.x; r.mx[0] = n2.y; r.mx[1] = n2.z; r.mx[2] = n2.w;
        stack[sp++] = r;     // left first out
        stack[sp++] = l;
    }
    *nCut = n;
}

__device__ float roundUpF(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
__device__ float roundDownF(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

// One leaf slot's plane record (leafLowerBound): normal and edge normals in FP64, rounded to float,
// then the offsets of those float planes over the three vertices in FP64 (products of floats are
// exact there), widened by r * 2^-40 and rounded outwards.  Layout, 5 float4: (n, lo), (hi, m0),
// (max m0.v, m1), (max m1.v, m2), (max m2.v, r, 0, 0).
__global__ void kBuildPrisms(const float4* __restrict__ tri, int n, float4* __restrict__ prism)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 t0 = tri[3 * (size_t)i], t1 = tri[3 * (size_t)i + 1], t2 = tri[3 * (size_t)i + 2];
    const double v[3][3] = { { t0.x, t0.y, t0.z }, { t0.w, t1.x, t1.y }, { t1.z, t1.w, t2.x } };
    double r = 0;
    for (int j = 0; j < 3; ++j) r = fmax(r, fabs(v[j][0]) + fabs(v[j][1]) + fabs(v[j][2]));
    float4* out = prism + 5 * (size_t)i;
    const double e0[3] = { v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2] };
    const double e1[3] = { v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2] };
    const double nn[3] = { e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0] };
    const double len = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    if (!(r < INFINITY) || !(len > 0) || !(len < INFINITY)) {
        for (int k = 0; k < 5; ++k) out[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const double nd[3] = { nn[0] / len, nn[1] / len, nn[2] / len };
    float pl[4][3];     // normal, then the three edge normals
    for (int k = 0; k < 3; ++k) pl[0][k] = (float)nd[k];
    for (int e = 0; e < 3; ++e) {
        const double* p0 = v[e];
        const double* p1 = v[(e + 1) % 3];
        const double d[3] = { p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2] };
        double mm[3] = { d[1] * nd[2] - d[2] * nd[1], d[2] * nd[0] - d[0] * nd[2], d[0] * nd[1] - d[1] * nd[0] };
        const double t = mm[0] * nd[0] + mm[1] * nd[1] + mm[2] * nd[2];
        for (int k = 0; k < 3; ++k) mm[k] -= t * nd[k];
        const double ml = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
        for (int k = 0; k < 3; ++k) pl[1 + e][k] = ml > 0 ? (float)(mm[k] / ml) : 0.f;
    }
    double hi[4], lo = INFINITY;
    for (int q = 0; q < 4; ++q) hi[q] = -INFINITY;
    for (int j = 0; j < 3; ++j)
        for (int q = 0; q < 4; ++q) {
            const double s = (double)pl[q][0] * v[j][0] + (double)pl[q][1] * v[j][1] + (double)pl[q][2] * v[j][2];
            hi[q] = fmax(hi[q], s);
            if (q == 0) lo = fmin(lo, s);
        }
    const double marg = r * 0x1p-40;
    out[0] = make_float4(pl[0][0], pl[0][1], pl[0][2], roundDownF(lo - marg));
    out[1] = make_float4(roundUpF(hi[0] + marg), pl[1][0], pl[1][1], pl[1][2]);
    out[2] = make_float4(roundUpF(hi[1] + marg), pl[2][0], pl[2][1], pl[2][2]);
    out[3] = make_float4(roundUpF(hi[2] + marg), pl[3][0], pl[3][1], pl[3][2]);
    out[4] = make_float4(roundUpF(hi[3] + marg), roundUpF(r), 0.f, 0.f);
}

__global__ void kFillU32(uint32_t* p, int n, uint32_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
} // namespace

hipError_t launchBatch(const BatchItem* items, int nItems, int64_t totalUnits, const DistOpts& o, float up,
                       unsigned long long* tickets, int xcdSplit, int blocks, int stackD, bool capped,
                       hipStream_t s)
{
    if (totalUnits <= 0 || nItems <= 0) return hipSuccess;
    if (o.minDistSq > 0 || blocks <= 0) return hipErrorInvalidValue;
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (capped)
            hipLaunchKernelGGL((kBatch<D, true>), dim3(blocks), dim3(kBlock), 0, s, items, nItems, totalUnits, o, up,
                               tickets, xcdSplit);
        else
            hipLaunchKernelGGL((kBatch<D, false>), dim3(blocks), dim3(kBlock), 0, s, items, nItems, totalUnits, o, up,
                               tickets, xcdSplit);
        return hipGetLastError();
    });
}

size_t splitScratchBytes(int64_t waves, int laneBytes)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return al((size_t)waves * 4) + al((size_t)waves * 64 * laneBytes) + al((size_t)waves * 64) + 256;
}

void splitScratchBind(void* base, int64_t waves, int laneBytes, int32_t budget, SplitScratch& sc)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* p = static_cast<char*>(base);
    sc.list = reinterpret_cast<uint32_t*>(p);
    p += al((size_t)waves * 4);
    sc.best = reinterpret_cast<uint32_t*>(p);
    p += al((size_t)waves * 64 * laneBytes);
    sc.order = reinterpret_cast<uint8_t*>(p);
    p += al((size_t)waves * 64);
    sc.counters = reinterpret_cast<uint32_t*>(p);
    sc.cap = (uint32_t)waves;
    sc.budget = budget;
}

int64_t projectionWaves(int64_t n)
{
    return n > 0 ? ((n + kBlock - 1) / kBlock) * kWaves : 0;
}

hipError_t buildCut(const float4* dnode, int32_t rootRef, const float rootMin[3], const float rootMax[3],
                    CutEntry* cut, int32_t* nCut, hipStream_t s)
{
    CutEntry root{};
    root.ref = rootRef;
    for (int i = 0; i < 3; ++i) { root.mn[i] = rootMin[i]; root.mx[i] = rootMax[i]; }
    hipLaunchKernelGGL(kBuildCut, dim3(1), dim3(64), 0, s, dnode, root, cut, nCut);
    return hipGetLastError();
}

hipError_t buildPrisms(const float4* tri, int32_t numLeaves, float4* prism, hipStream_t s)
{
    if (numLeaves <= 0) return hipSuccess;
    hipLaunchKernelGGL(kBuildPrisms, dim3((numLeaves + 255) / 256), dim3(256), 0, s, tri, numLeaves, prism);
    return hipGetLastError();
}

hipError_t buildQuads(const float4* dnode, int32_t numInternal, float4* dquad, hipStream_t s)
{
    if (numInternal <= 0) return hipSuccess;
    hipLaunchKernelGGL(kBuildQuads, dim3((numInternal + 255) / 256), dim3(256), 0, s, dnode, numInternal, dquad);
    return hipGetLastError();
}

hipError_t fillU32(uint32_t* p, int n, uint32_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kFillU32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

int stackDepthFor(int treeDepth)
{
    // pending far children of the ancestors of the current node: at most treeDepth entries
    if (treeDepth <= 8) return 8;
    if (treeDepth <= 16) return 16;
    if (treeDepth <= 24) return 24;
    if (treeDepth <= 32) return 32;
    return 0;
}

namespace {
std::atomic<int> g_variant{ (int)Traversal::WaveCoop };
}

void setTraversal(Traversal t) { g_variant.store((int)t); }
Traversal traversal() { return (Traversal)g_variant.load(); }

namespace {
std::atomic<int> g_leafFilter{ -1 };
}

void setLeafFilter(int mode) { g_leafFilter.store(mode); }
int leafFilter() { return g_leafFilter.load(); }

int64_t sdfGridWaves(const int32_t dims[3], int32_t zBegin, int32_t zEnd)
{
    const int64_t nbx = (dims[0] + 3) / 4, nby = (dims[1] + 3) / 4, nbz = (zEnd - zBegin + 3) / 4;
    const int64_t bricks = nbx * nby * nbz;
    return bricks > 0 ? ((bricks + 3) / 4) * kWaves : 0;
}

hipError_t launchSdfGrid(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                         int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s,
                         const float* bound, uint32_t* rerun, uint32_t* rerunCount, const SplitScratch* split)
{
    const int nbx = (dims[0] + 3) / 4, nby = (dims[1] + 3) / 4, nbz = (zEnd - zBegin + 3) / 4;
    const int64_t bricks = (int64_t)nbx * nby * nbz;
    if (bricks <= 0) return hipSuccess;
    const int nLogical = (int)((bricks + 3) / 4);
    // the wave-cooperative distance returns the minimum over candidates; with minDistSq > 0 the
    // reference stops at the first candidate <= minDistSq in its own visit order, which the keyed walk
    // reproduces (launchSdfGridKeyed; its split scratch holds 8 bytes per lane)
    if (o.minDistSq > 0 && !bound && m.traversal == (int)Traversal::WaveCoop)
        return launchSdfGridKeyed(m, dims, xf, o, zBegin, zEnd, out, stackD, s, split);
    const bool coop = (m.traversal == (int)Traversal::WaveCoop || bound) && !(o.minDistSq > 0);
    if (bound && (!coop || !rerun || !rerunCount)) return hipErrorInvalidValue;
    // The leaf filter saves instructions but adds a dependent load to every leaf visit.  It pays unless
    // a deep tree meets a small grid: then few waves walk long chains and nothing hides that load.
    // Measured on / off (bumpy spheres): k=5 64^3 0.89, k=7 128^3 0.87, k=8 96^3 0.96, k=8 256^3 0.87,
    // k=9 192^3 0.99 -- but k=8 64^3 (4 k waves) 1.11 and k=9 128^3 (32 k waves) 1.05; on where the grid
    // has more than 2^(depth - 8) waves.
    const int filterMode = m.leafFilter;
    const int64_t waves = sdfGridWaves(dims, zBegin, zEnd);
    const bool filter = filterMode < 0 ? waves > ((int64_t)1 << (m.depth > 8 ? m.depth - 8 : 0)) : filterMode != 0;
    SplitScratch sc;
    // the capped form takes the split scratch even with no step budget (its failed waves use it)
    if (coop && split && split->list && (split->budget > 0 || bound) && m.nCut > 0) {
        if ((int64_t)split->cap < sdfGridWaves(dims, zBegin, zEnd)) return hipErrorInvalidValue;
        sc = *split;
        hipError_t e = fillU32(sc.counters, 3, 0u, s);
        if (e != hipSuccess) return e;
    }
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (coop) {
            const auto gridKernel = bound ? (filter ? kSdfGridCoop<D, true, true> : kSdfGridCoop<D, false, true>)
                                          : (filter ? kSdfGridCoop<D, true, false> : kSdfGridCoop<D, false, false>);
            const auto splitKernel = filter ? kSdfSplit<D, true> : kSdfSplit<D, false>;
            hipLaunchKernelGGL(gridKernel, dim3(chunkedGrid(nLogical)), dim3(kBlock), 0, s, m, dims[0], dims[1],
                               zBegin, zEnd, nbx, nby, nbz, nLogical, xf, o, bound, rerun, rerunCount, sc, out);
            if (sc.list) {
                const int blocks = persistentBlocks();
                hipLaunchKernelGGL(splitKernel, dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 1), dims[0], dims[1], zBegin, zEnd,
                                   nbx, nby, nbz, xf, sc);
                hipLaunchKernelGGL(kSdfSplitFinish<D>, dim3(blocks), dim3(kBlock), 0, s, statLaunch(m, 2), dims[0], dims[1],
                                   zBegin, zEnd, nbx, nby, nbz, xf, o, sc, bound, rerunCount, out);
            }
        } else {
            hipLaunchKernelGGL(kSdfGrid<D>, dim3(chunkedGrid(nLogical)), dim3(kBlock), 0, s, m, dims[0],
                               dims[1], zBegin, zEnd, nbx, nby, nbz, nLogical, xf, o, out);
        }
        return hipGetLastError();
    });
}

hipError_t launchSdfVoxels(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                           int32_t zBegin, const uint32_t* list, int64_t count, float* out, int stackD,
                           hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    if (o.minDistSq > 0) return hipErrorInvalidValue;
    const int nLogical = (int)((count + kBlock - 1) / kBlock);
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        hipLaunchKernelGGL(kSdfVoxelsCoop<D>, dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m, dims[0], dims[1],
                           zBegin, list, count, nLogical, xf, o, out);
        return hipGetLastError();
    });
}

hipError_t launchFwnGrid(const MeshDev& m, const int32_t dims[3], const Xf& xf, float beta,
                         int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s)
{
    const int nbx = (dims[0] + 3) / 4, nby = (dims[1] + 3) / 4, nbz = (zEnd - zBegin + 3) / 4;
    const int64_t bricks = (int64_t)nbx * nby * nbz;
    if (bricks <= 0) return hipSuccess;
    const int nLogical = (int)((bricks + 3) / 4);
    const bool coop = m.traversal == (int)Traversal::WaveCoop;
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (coop)
            hipLaunchKernelGGL(kFwnGridCoop<D>, dim3(chunkedGrid(nLogical)), dim3(kBlock), 0, s, m, dims[0],
                               dims[1], zBegin, zEnd, nbx, nby, nbz, nLogical, xf, beta, out);
        else
            hipLaunchKernelGGL(kFwnGrid<D>, dim3(chunkedGrid(nLogical)), dim3(kBlock), 0, s, m, dims[0],
                               dims[1], zBegin, zEnd, nbx, nby, nbz, nLogical, xf, beta, out);
        return hipGetLastError();
    });
}

hipError_t launchFwnPoints(const MeshDev& m, const float* pts, const uint32_t* perm, int64_t n,
                           float beta, int32_t skipFace, float* out, int stackD, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int nLogical = (int)((n + kBlock - 1) / kBlock);
    const bool coop = m.traversal == (int)Traversal::WaveCoop;
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (coop)
            hipLaunchKernelGGL(kFwnPointsCoop<D>, dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m, pts,
                               perm, n, nLogical, beta, skipFace, out);
        else
            hipLaunchKernelGGL(kFwnPoints<D>, dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m, pts,
                               perm, n, nLogical, beta, skipFace, out);
        return hipGetLastError();
    });
}

hipError_t launchSelfIntersections(const MeshDev& m, float beta, uint8_t* out, int stackD, hipStream_t s)
{
    if (m.numFaces <= 0) return hipSuccess;
    const int nLogical = (m.numFaces + kBlock - 1) / kBlock;
    const bool coop = m.traversal == (int)Traversal::WaveCoop;
    return dispatchDepth(stackD, [&](auto Dc) {
        constexpr int D = decltype(Dc)::value;
        if (coop)
            hipLaunchKernelGGL(kSelfIntersectionsCoop<D>, dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m,
                               nLogical, beta, out);
        else
            hipLaunchKernelGGL(kSelfIntersections<D>, dim3(paddedGrid(nLogical)), dim3(kBlock), 0, s, m,
                               nLogical, beta, out);
        return hipGetLastError();
    });
}

namespace {
inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
size_t cubTempBytes(int64_t n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 30);
    return bytes;
}
} // namespace

size_t sortWorkspaceBytes(int64_t n)
{
    if (n <= 0) return 0;
    return 4 * align256((size_t)n * 4) + align256(cubTempBytes(n));
}

hipError_t sortPointsMorton(const float* pts, int64_t n, const float lo[3], const float hi[3],
                            void* ws, size_t wsBytes, const uint32_t** permOut, hipStream_t s)
{
    *permOut = nullptr;
    if (n <= 0) return hipSuccess;
    if (n > INT32_MAX || wsBytes < sortWorkspaceBytes(n)) return hipErrorInvalidValue;
    char* p = static_cast<char*>(ws);
    const size_t a = align256((size_t)n * 4);
    uint32_t* keysIn = reinterpret_cast<uint32_t*>(p);
    uint32_t* keysOut = reinterpret_cast<uint32_t*>(p + a);
    uint32_t* valsIn = reinterpret_cast<uint32_t*>(p + 2 * a);
    uint32_t* valsOut = reinterpret_cast<uint32_t*>(p + 3 * a);
    void* temp = p + 4 * a;
    size_t tempBytes = cubTempBytes(n);
    float sc[3];
    for (int i = 0; i < 3; ++i) {
        const float ext = hi[i] - lo[i];
        sc[i] = ext > 0 ? 1023.f / ext : 0.f;
    }
    const int tb = 256;
    hipLaunchKernelGGL(kMorton, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, s, pts, n, lo[0], lo[1],
                       lo[2], sc[0], sc[1], sc[2], keysIn, valsIn);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, keysIn, keysOut, valsIn, valsOut, (int)n, 0,
                                           30, s);
    if (e != hipSuccess) return e;
    *permOut = valsOut;
    return hipSuccess;
}

} // namespace mi