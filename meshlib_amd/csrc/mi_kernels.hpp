// mi_kernels.hpp -- launchers of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mi {

// device view of one mesh's flattened tree (layout in mi_tree.hpp)
// A subtree root of the cut at depth kCutDepth (or a leaf above it): where the split pass starts
// its partial walks.  path: bit l = direction taken at level l from the root (0 left, 1 right).
struct CutEntry {
    int32_t ref, depth;
    uint32_t path, pad;
    float mn[3], mx[3];
};
constexpr int kCutDepth = 6;
constexpr int kMaxCut = 1 << kCutDepth;

// Measurement hook: per-launch counters of the records the wave-cooperative walks read, added by
// lane 0 of each wave at the end of a walk when MeshDev::stats is set.  A call's launches use
// consecutive blocks of kStatStride counters (its first launch block 0, a split pass block 1, the
// split form's last launch block 2).
enum StatSlot : int {
    kStatWalks = 0,       // walks run (one per wave and walk)
    kStatQuadSteps,       // distance walk: 4-wide node visits (128-B quad record each)
    kStatPrismReads,      // distance walk: leaf visits reading the plane record (80 B each)
    kStatLeafTests,       // distance walk: leaf tests run (48-B triangle record each)
    kStatKeyedSteps,      // keyed walk: node visits (128-B quad + 48-B node record each)
    kStatKeyedTests,      // keyed walk: leaf tests run (48-B triangle record each)
    kStatFwnVisits,       // winding number: node visits (32-B dipole record each)
    kStatFwnDescents,     // winding number: descents into an internal node's children (16-B refs)
    kStatFwnLeaves,       // winding number: exact solid angles (48-B triangle record each)
    kStatSeeds,           // greedy descents seeding a walk (one root-to-leaf path per wave)
    kStatSeedLevels,      // their levels (one 64-B node record each)
    kStatCount
};
constexpr int kStatStride = 16;
constexpr int kStatLaunches = 3;

struct MeshDev {
    const float4* dnode;   // 4 per internal node
    const float4* tri;     // 3 per leaf slot
    const float4* dint;    // 2 per internal node
    const float4* dleaf;   // 2 per leaf slot
    const float4* dquad;   // 8 per internal node: its grandchildren (4-wide records of the cooperative walks)
    const CutEntry* cut;   // nCut subtree roots covering every leaf once
    const float4* prism;   // 5 per leaf slot: planes bounding the triangle (leaf filter of the walks)
    int32_t nCut;
    float rootMin[3], rootMax[3];
    int32_t rootRef;
    int32_t numFaces;
    int32_t depth;         // levels of the tree
    unsigned long long* stats;   // measurement counters (kStatStride per launch) or null
    // settings of the call (host side, read by the launchers only): resolved per call from the mesh's
    // own settings and the process-wide defaults
    int32_t traversal = 1;       // Traversal
    int32_t leafFilter = -1;     // leaf filter mode of the signed-distance walks: 0 off, 1 on, -1 auto
};

// the mesh view of a call's launch number `launch` (its own block of counters)
inline MeshDev statLaunch(const MeshDev& m, int launch)
{
    MeshDev r = m;
    if (r.stats) r.stats += (size_t)launch * kStatStride;
    return r;
}

struct Xf { float m[12]; };   // row-major A then b (AffineXf3f)

struct DistOpts { float minDistSq, maxDistSq; int32_t nullOutsideMinMax; float threshold, beta; };

struct ProjRec { int32_t face; float px, py, pz; int32_t e; float a, b; float distSq; };
static_assert(sizeof(ProjRec) == 32, "MeshProjectionResult is 32 B");

// traversal form: PerLane = one independent depth-first walk per lane (the reference's form);
// WaveCoop = one walk per wave over its 64 coherent queries (results identical, see kernels)
enum class Traversal : int { PerLane = 0, WaveCoop = 1 };
// process-wide defaults of the call settings in MeshDev
void setTraversal(Traversal t);
Traversal traversal();
// leaf filter of the signed-distance grid walks: 0 off, 1 on, -1 by the grid's wave count
void setLeafFilter(int mode);
int leafFilter();

// one item of a batched launch: a signed-distance grid (gridUnits bricks of 4x4x4 voxels) and/or a
// Morton-sorted point set (projUnits chunks of 64), on one mesh
struct BatchItem {
    MeshDev m;
    int32_t dx, dy, dz, nbx, nby, nbz;
    Xf xf;
    float* sdf;
    const float* pts;
    const uint32_t* perm;
    int64_t npts;
    ProjRec* proj;
    int64_t gridUnits, projUnits, unitBegin;
    // capped launch only: optional per-query caps (squared) on the search radius, indexed like the
    // outputs; a wave walks once more, uncapped, for its queries that found nothing below a cap smaller
    // than the plain limit, counted in rerunCount[0] (voxels) / [1] (points)
    const float* capSdf;
    const float* capProj;
    uint32_t* rerunCount;
};

// persistent launch over items[0..nItems) (device array); tickets = 8 zeroed 64-bit counters;
// blocks = number of 256-thread workgroups to keep resident; capped: the items' caps apply
hipError_t launchBatch(const BatchItem* items, int nItems, int64_t totalUnits, const DistOpts& o, float up,
                       unsigned long long* tickets, int xcdSplit, int blocks, int stackD, bool capped,
                       hipStream_t s);

// small fill kernel (used instead of hipMemsetAsync so streams stay graph-capture safe)
// waves of the wave-cooperative signed-distance launch over z-layers [zBegin, zEnd)
int64_t sdfGridWaves(const int32_t dims[3], int32_t zBegin, int32_t zEnd);
// waves of the wave-cooperative projection launch over n points
int64_t projectionWaves(int64_t n);

// 4-wide records from the binary records: internal node j's four grandchildren, a leaf child
// standing for itself in one slot with an empty partner slot
hipError_t buildQuads(const float4* dnode, int32_t numInternal, float4* dquad, hipStream_t s);
// the leaf-slot plane records of MeshDev::prism from the triangle records
hipError_t buildPrisms(const float4* tri, int32_t numLeaves, float4* prism, hipStream_t s);

// the cut of the tree at depth kCutDepth into at most kMaxCut subtrees (cut[], *nCut on the device)
hipError_t buildCut(const float4* dnode, int32_t rootRef, const float rootMin[3], const float rootMax[3],
                    CutEntry* cut, int32_t* nCut, hipStream_t s);

// Split form of the wave-cooperative walks: a wave still walking after `budget` steps stops, keeps
// its lanes' best candidates in the scratch and is finished by a second launch that splits its walk
// over the cut's subtrees (partial results merged by atomic minimum) and a third that stores the
// results.  Exact: every partial result is a real candidate and the subtrees cover every leaf.
struct SplitScratch {
    uint32_t* list = nullptr;       // ids of the waves handed over
    uint32_t* best = nullptr;       // 64 per handed-over wave; signed distance: float bits of the best
                                    // squared distance; projections (as 64-bit): distSq bits << 32 | key
    uint32_t* counters = nullptr;   // [0] waves handed over, [1] split-pass ticket, [2] finish-pass ticket
    uint8_t* order = nullptr;       // 64 per handed-over wave: cut subtrees nearest first
    uint32_t cap = 0;
    int32_t budget = 0;             // steps; 0 = never hand over
};
// laneBytes: 4 for the signed-distance walk, 8 for projections
size_t splitScratchBytes(int64_t waves, int laneBytes);
void splitScratchBind(void* base, int64_t waves, int laneBytes, int32_t budget, SplitScratch& sc);

hipError_t fillU32(uint32_t* p, int n, uint32_t v, hipStream_t s);

// stack depth template actually used for a tree of the given depth (0 if unsupported)
int stackDepthFor(int treeDepth);

// z-layer slab [zBegin, zEnd) of the grid, output x-fastest
// bound (optional): per-voxel cap on maxDistSq (indexed like out); with rerun lists given, voxels
// whose cap found nothing are appended (slab-relative ids) instead of written
hipError_t launchSdfGrid(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                         int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s,
                         const float* bound = nullptr, uint32_t* rerun = nullptr, uint32_t* rerunCount = nullptr,
                         const SplitScratch* split = nullptr);
// the same with minDistSq > 0 on the keyed walks (the reference's early exit at minDistSq reproduced
// exactly); split scratch (optional) bound with 8 bytes per lane
hipError_t launchSdfGridKeyed(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                              int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s,
                              const SplitScratch* split);
// the listed slab voxels, searched without a cap
hipError_t launchSdfVoxels(const MeshDev& m, const int32_t dims[3], const Xf& xf, const DistOpts& o,
                           int32_t zBegin, const uint32_t* list, int64_t count, float* out, int stackD,
                           hipStream_t s);
hipError_t launchFwnGrid(const MeshDev& m, const int32_t dims[3], const Xf& xf, float beta,
                         int32_t zBegin, int32_t zEnd, float* out, int stackD, hipStream_t s);
// perm may be null (identity order)
hipError_t launchFwnPoints(const MeshDev& m, const float* pts, const uint32_t* perm, int64_t n,
                           float beta, int32_t skipFace, float* out, int stackD, hipStream_t s);
hipError_t launchSelfIntersections(const MeshDev& m, float beta, uint8_t* out, int stackD, hipStream_t s);
// bound (optional): per-query cap on up (indexed by query); rerun lists as for launchSdfGrid
hipError_t launchProjections(const MeshDev& m, const float* pts, const uint32_t* perm, int64_t n,
                             const Xf* ptXf, const Xf* treeXf, float up, float lo, ProjRec* out,
                             int stackD, hipStream_t s, const float* bound = nullptr,
                             uint32_t* rerun = nullptr, uint32_t* rerunCount = nullptr,
                             const SplitScratch* split = nullptr);

// spatial ordering of query points: Morton codes over box [lo, hi] then a radix sort.
// workspace layout is private; query its size first with ws == nullptr.
size_t sortWorkspaceBytes(int64_t n);
hipError_t sortPointsMorton(const float* pts, int64_t n, const float lo[3], const float hi[3],
                            void* ws, size_t wsBytes, const uint32_t** permOut, hipStream_t s);

} // namespace mi
