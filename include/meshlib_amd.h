/*
 * meshlib_amd.h -- C-ABI of the MI355X (gfx950) distance-field path.
 *
 * Plain pointers and sizes, no C++ or torch types.  Every entry point returns
 * MI_OK (0) on success or a negative MI_E_* status; mi_last_error() returns the
 * calling thread's last message.  Host-pointer entry points behave like the
 * reference's synchronous calls (upload, compute, download, within the device
 * memory budget); they run in parts (64 MiB of voxels, 2^21 points) whose copies
 * overlap the computation of the next part.  *_device entry points take device pointers that are already
 * resident in HBM and an optional hipStream_t (NULL = default stream); they are
 * asynchronous with respect to the host.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef MESHLIB_AMD_H
#define MESHLIB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_OK 0
#define MI_E_INVALID (-1)   /* bad argument */
#define MI_E_HIP (-2)       /* HIP runtime error ("AMD GPU error: ...") */
#define MI_E_CANCELED (-3)  /* progress callback returned 0 (operation canceled) */
#define MI_E_NOMEM (-4)     /* host allocation failed */
#define MI_E_LIMIT (-5)     /* input exceeds a fixed capacity of the entry point */

/* DistanceToMeshOptions (source/MRMesh/MRDistanceToMeshOptions.h:10-33) */
typedef struct mi_dist_options {
    float minDistSq;               /* default 0 */
    float maxDistSq;               /* default FLT_MAX */
    int32_t nullOutsideMinMax;     /* default 1 */
    float windingNumberThreshold;  /* default 0.5 */
    float windingNumberBeta;       /* default 2 */
} mi_dist_options;

/* MeshProjectionResult wire record, 32 B (source/MRMesh/MRMeshProject.h:17-31,
 * GPU twin source/MRCuda/MRCudaPointsToMeshProjector.cuh:12-17).  face = -1 when
 * nothing was found below upDistLimitSq (then distSq = upDistLimitSq).  e is always
 * -1 here: edgeWithLeft(face) is topology, filled by the host adapter
 * (source/MRCuda/MRCudaPointsToMeshProjector.cpp:118-124). */
typedef struct mi_projection {
    int32_t face;
    float point[3];
    int32_t e;
    float a, b;
    float distSq;
} mi_projection;

/* AABB tree node in the reference's pre-order layout, 32 B
 * (source/MRMesh/MRAABBTreeNode.h:29-45): leaf <=> r < 0, then l = face id. */
typedef struct mi_tree_node {
    float min[3], max[3];
    int32_t l, r;
} mi_tree_node;

/* Dipole, 32 B (source/MRMesh/MRDipole.h:10-30) */
typedef struct mi_dipole {
    float pos[3];
    float area;
    float dirArea[3];
    float rr;
} mi_dipole;

/* ProgressCallback: return 0 to cancel */
typedef int (*mi_progress_fn)(float progress, void* user);

/* IFastWindingNumberByParts::GridByPartsFunc (source/MRMesh/MRFastWindingNumber.h:88):
 * receives one block of z-layers, dims = {dims.x, dims.y, layers}, starting at layer
 * zOffset; the data pointer is valid only during the call.  Return MI_OK to continue. */
typedef int (*mi_grid_part_fn)(const float* data, const int32_t dims[3], int32_t zOffset, void* user);

typedef struct mi_mesh mi_mesh;

/* ---------------------------------------------------------------- errors / config */

/* copies the calling thread's last error message into buf; returns its full length */
size_t mi_last_error(char* buf, size_t len);

/* getCudaSafeMemoryLimit (source/MRCuda/MRCudaBasic.cpp:86-92) replacement:
 * 0 = use 0.8 * (free - 128 MiB); otherwise cap the chunk budget of the host-pointer
 * entry points at this many bytes (used by tests to force multi-chunk execution) */
int mi_set_memory_limit(size_t bytes);

/* Process-wide settings.  They are atomic, so setting them from any thread is defined, but every
 * thread shares them: a thread that needs its own value sets it on its meshes with
 * mi_mesh_set_mode (below), which takes precedence over the process-wide value.
 *
 * Traversal form of the kernels, process-wide default (no counterpart in the reference; its GPU path
 * has one form, MRCudaFastWindingNumber.cu:31-133).  Results do not depend on it.
 *   MI_TRAVERSAL_PER_LANE  one depth-first walk per query, the reference's form
 *   MI_TRAVERSAL_WAVE_COOP one walk per 64-query wave with uniform node loads (default) */
#define MI_TRAVERSAL_PER_LANE 0
#define MI_TRAVERSAL_WAVE_COOP 1
int mi_set_traversal(int mode);
int mi_get_traversal(void);

/* Leaf filter of the wave-cooperative signed-distance walks, process-wide default (no counterpart in the
 * reference).  A queued leaf test is dropped for the lanes where a lower bound of the triangle's
 * distance (its plane slab and edge half-spaces, with rounding margins) already exceeds the lane's
 * best, so results do not depend on it.  MI_LEAF_FILTER_AUTO (default) uses it for grids of more
 * than 2^(depth - 8) waves of 64 voxels (depth = tree levels): on deep trees a small grid leaves too
 * few waves to hide the load it adds to each leaf visit. */
#define MI_LEAF_FILTER_OFF 0
#define MI_LEAF_FILTER_ON 1
#define MI_LEAF_FILTER_AUTO (-1)
int mi_set_leaf_filter(int mode);
int mi_get_leaf_filter(void);

/* Where mi_mesh_create builds the AABB tree and dipoles, process-wide.  The reference builds them
 * on the host inside the first call (source/MRCuda/MRCudaFastWindingNumber.cpp:70, AABBTree ctor
 * source/MRMesh/MRAABBTree.cpp:31-60, calcDipoles source/MRMesh/MRDipole.cpp:23-71); both builders
 * here produce that tree node-for-node and bit-for-bit.
 *   MI_TREE_BUILDER_DEVICE level-synchronous GPU build (default)
 *   MI_TREE_BUILDER_HOST   multithreaded host build, flattened and uploaded */
#define MI_TREE_BUILDER_DEVICE 0
#define MI_TREE_BUILDER_HOST 1
int mi_set_tree_builder(int builder);
int mi_get_tree_builder(void);

/* Step budgets of the split form of the wave-cooperative walks, process-wide defaults (no counterpart in the
 * reference).  A 64-query wave still walking after `steps` node visits hands its queries, with their
 * best candidates so far, to a second launch that walks the tree's 64 top subtrees for them in
 * parallel; results are identical for every budget.  0 = never split; MI_SPLIT_AUTO (the default)
 * picks per call from the tree depth and the number of waves (signed distance on trees of >= 18
 * levels: 3072 steps below 64 k waves, 6144 below 1 M waves, no split from 1 M waves (a 512^3 grid)
 * on; shallower trees: 512 / 1024 below / above 16 k waves; projections: 64 for trees of >= 20
 * levels, else 2048 (>= 17 levels) or 1024).  The device
 * entry points of one mesh share that mesh's workspace: calls on different streams must not overlap
 * in time. */
#define MI_SPLIT_SDF 0
#define MI_SPLIT_PROJECTIONS 1
#define MI_SPLIT_AUTO (-1)
int     mi_set_split_budget(int32_t walk, int32_t steps);
int32_t mi_get_split_budget(int32_t walk);

/* ---------------------------------------------------------------- mesh data */

/* Cuda::FastWindingNumber::prepareData_ (source/MRCuda/MRCudaFastWindingNumber.cpp:46-85)
 * and Cuda::PointsToMeshProjector::updateMeshData (source/MRCuda/MRCudaPointsToMeshProjector.cpp:31-50):
 * verts float[nv*3], tris int32[nf*3] (getTriangulation order: nf = faceSize(), a deleted face is
 * {-1,-1,-1}, source/MRMesh/MRMeshTopology.cpp:430-441).  Builds the AABB tree over the valid faces,
 * leaves keeping their face ids (source/MRMesh/MRAABBTree.cpp:31-60), and the dipoles
 * (source/MRMesh/MRDipole.cpp:23-71) with the selected builder and keeps the flattened tree on the
 * current HIP device.  A face with any negative index is a deleted face (skipped, as the reference's
 * GPU kernels do, source/MRCuda/MRCudaFastWindingNumber.cu:168); an index >= nv is MI_E_INVALID.
 * Face ids in results (projection records, skipFace, self-intersection flags) index the caller's
 * array.  At most 2^30 faces. */
int  mi_mesh_create(const float* verts, int32_t nv, const int32_t* tris, int32_t nf, mi_mesh** out);
void mi_mesh_destroy(mi_mesh* m);
/* tree nodes: 2 * (valid faces) - 1, or 0 */
int  mi_mesh_num_nodes(const mi_mesh* m);

/* Measurement hook (no counterpart in the reference): while d_counters (device, zeroed by the
 * caller, MI_STAT_LAUNCHES * MI_STAT_STRIDE uint64) is set, the wave-cooperative walks of every
 * later call on this mesh add the records they read: block k counts a call's launch k (0 the first
 * launch, 1 the split form's partial walks, 2 its last launch), slot MI_STAT_* within the block.
 * Counts are per wave (a record read once for 64 lanes counts once).  NULL turns counting off.
 * Counting costs a few scalar additions per step and one atomic per walk; results are unchanged. */
#define MI_STAT_WALKS 0          /* walks run */
#define MI_STAT_QUAD_STEPS 1     /* distance walk: 4-wide node visits (128-B quad record) */
#define MI_STAT_PRISM_READS 2    /* distance walk: leaf visits reading the plane record (80 B) */
#define MI_STAT_LEAF_TESTS 3     /* distance walk: leaf tests (48-B triangle record each) */
#define MI_STAT_KEYED_STEPS 4    /* keyed (projection) walk: node visits (128-B quad + 48-B node record) */
#define MI_STAT_KEYED_TESTS 5    /* keyed walk: leaf tests (48-B triangle record each) */
#define MI_STAT_FWN_VISITS 6     /* winding number: node visits (32-B dipole record) */
#define MI_STAT_FWN_DESCENTS 7   /* winding number: descents into children (16-B child refs) */
#define MI_STAT_FWN_LEAVES 8     /* winding number: exact solid angles (48-B triangle record) */
#define MI_STAT_SEEDS 9          /* greedy descents seeding a walk (one root-to-leaf path per wave) */
#define MI_STAT_SEED_LEVELS 10   /* their levels (one 64-B node record each) */
#define MI_STAT_COUNT 11
#define MI_STAT_STRIDE 16
#define MI_STAT_LAUNCHES 3
int  mi_mesh_set_stats(mi_mesh* m, uint64_t* d_counters);
/* exports the tree / dipoles in the reference layout (AABBTree::nodes(), Mesh::getDipoles()) */
int  mi_mesh_get_tree(const mi_mesh* m, mi_tree_node* nodes, mi_dipole* dipoles);
/* depth of the tree (longest root-to-leaf path, in edges) */
int  mi_mesh_depth(const mi_mesh* m);
/* Settings of one mesh (no counterpart in the reference, whose GPU classes carry no mode): each
 * overrides, for the calls on this mesh only, the process-wide setting of the same name above, so
 * that threads working on different meshes do not share settings.  Every setting starts at
 * MI_MODE_DEFAULT, which follows the process-wide setting at each call.  Values:
 *   MI_MODE_TRAVERSAL          MI_TRAVERSAL_*
 *   MI_MODE_LEAF_FILTER        MI_LEAF_FILTER_*
 *   MI_MODE_SPLIT_SDF          steps or MI_SPLIT_AUTO (as mi_set_split_budget(MI_SPLIT_SDF, ...))
 *   MI_MODE_SPLIT_PROJECTIONS  steps or MI_SPLIT_AUTO
 * A setting changed while a call on the mesh runs applies from the next call on.  mi_mesh_get_mode
 * returns the mesh's own value (MI_MODE_DEFAULT when it follows the process-wide setting). */
#define MI_MODE_TRAVERSAL 0
#define MI_MODE_LEAF_FILTER 1
#define MI_MODE_SPLIT_SDF 2
#define MI_MODE_SPLIT_PROJECTIONS 3
#define MI_MODE_COUNT 4
#define MI_MODE_DEFAULT (-2)
int     mi_mesh_set_mode(mi_mesh* m, int32_t mode, int32_t value);
int32_t mi_mesh_get_mode(const mi_mesh* m, int32_t mode);
/* the same host tree + dipole build without any device (AABBTree ctor + calcDipoles):
 * fills nodes[2*k-1] and dipoles[2*k-1], k = the valid faces of tris (either may be NULL), and *depth */
int  mi_build_tree_host(const float* verts, int32_t nv, const int32_t* tris, int32_t nf,
                        mi_tree_node* nodes, mi_dipole* dipoles, int32_t* depth);

/* ---------------------------------------------------------------- IFastWindingNumber */

/* IFastWindingNumber::calcFromVector (source/MRMesh/MRFastWindingNumber.h:26;
 * GPU source/MRCuda/MRCudaFastWindingNumber.cpp:87-127); skipFace = -1 for none */
int mi_fwn_calc_from_vector(mi_mesh* m, const float* pts, int64_t n, float beta, int32_t skipFace,
                            float* res, mi_progress_fn cb, void* user);
/* IFastWindingNumber::calcSelfIntersections (MRFastWindingNumber.h:31;
 * GPU MRCudaFastWindingNumber.cpp:129-173): res[f] = 1 if wn(centroid f, skip f) outside [0,1];
 * res has one entry per face of the array passed to mi_mesh_create (deleted faces: 0) */
int mi_fwn_calc_self_intersections(mi_mesh* m, float beta, uint8_t* res, mi_progress_fn cb, void* user);
/* IFastWindingNumber::calcFromGrid (MRFastWindingNumber.h:38; GPU MRCudaFastWindingNumber.cpp:175-228);
 * xf = gridToMeshXf as float[12]: row-major A then b; voxel ids x-fastest */
int mi_fwn_calc_from_grid(mi_mesh* m, const int32_t dims[3], const float xf[12], float beta,
                          float* res, mi_progress_fn cb, void* user);
/* IFastWindingNumber::calcFromGridWithDistances (MRFastWindingNumber.h:44;
 * GPU MRCudaFastWindingNumber.cpp:230-283) */
int mi_fwn_calc_from_grid_with_distances(mi_mesh* m, const int32_t dims[3], const float xf[12],
                                         const mi_dist_options* opts, float* res,
                                         mi_progress_fn cb, void* user);
/* IFastWindingNumberByParts::calcFromGridByParts / calcFromGridWithDistancesByParts
 * (MRFastWindingNumber.h:96-106; GPU MRCudaFastWindingNumber.cpp:285-429).  Parts are whole
 * z-layers, about 64 MiB of output each (at least 4 * layerOverlap layers, at most half the memory
 * budget), consecutive parts sharing layerOverlap layers; fn gets a page-locked buffer that is
 * reused once it returns, so it copies what it keeps. */
int mi_fwn_calc_from_grid_by_parts(mi_mesh* m, mi_grid_part_fn fn, void* fnUser, const int32_t dims[3],
                                   const float xf[12], float beta, int32_t layerOverlap,
                                   mi_progress_fn cb, void* user);
int mi_fwn_calc_from_grid_with_distances_by_parts(mi_mesh* m, mi_grid_part_fn fn, void* fnUser,
                                                  const int32_t dims[3], const float xf[12],
                                                  const mi_dist_options* opts, int32_t layerOverlap,
                                                  mi_progress_fn cb, void* user);

/* device-resident variants (no chunking, no host copies).  The grid variants compute the
 * z-layers [zBegin, zEnd) into d_res ((zEnd - zBegin) * dims.x * dims.y floats, x-fastest). */
int mi_fwn_vector_device(mi_mesh* m, const float* d_pts, int64_t n, float beta, int32_t skipFace,
                         float* d_res, void* stream);
int mi_fwn_grid_device(mi_mesh* m, const int32_t dims[3], const float xf[12], float beta,
                       int32_t zBegin, int32_t zEnd, float* d_res, void* stream);
int mi_sdf_grid_device(mi_mesh* m, const int32_t dims[3], const float xf[12],
                       const mi_dist_options* opts, int32_t zBegin, int32_t zEnd,
                       float* d_res, void* stream);

/* ---------------------------------------------------------------- IPointsToMeshProjector */

/* IPointsToMeshProjector::findProjections (source/MRMesh/MRPointsToMeshProjector.h:12-30;
 * GPU source/MRCuda/MRCudaPointsToMeshProjector.cpp:52-125).  objXf / refObjXf may be NULL;
 * a rigid refObjXf is folded into the point transform, a non-rigid one is applied to the
 * tree boxes and triangles (createProjectionTransforms, source/MRMesh/MRMeshProject.cpp:105-121). */
int mi_proj_find_projections(mi_mesh* m, const float* pts, int64_t n, const float* objXf,
                             const float* refObjXf, float upDistLimitSq, float loDistLimitSq,
                             mi_projection* res);
/* IPointsToMeshProjector::projectionsHeapBytes (MRCudaPointsToMeshProjector.cpp:127-130) */
size_t mi_proj_heap_bytes(size_t numProjections);
/* device-resident variant: ptXf / treeXf are the already simplified transforms (or NULL) */
int mi_proj_find_projections_device(mi_mesh* m, const float* d_pts, int64_t n, const float* ptXf,
                                    const float* treeXf, float upDistLimitSq, float loDistLimitSq,
                                    mi_projection* d_res, void* stream);

/* The same call, allocating nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops such as an ICP
 * iteration, harnesses that forbid hipMalloc).  Same arguments and results as
 * mi_proj_find_projections_device (bit-identical records), but the sort and split workspace lives
 * in a fixed 96 MiB device array that the library holds on each GPU it runs on until the process
 * ends, so it takes at most MI_NOALLOC_MAX_POINTS points per call (MI_E_LIMIT beyond).  Every call
 * on one GPU, on any mesh, shares that GPU's array: calls that overlap in time on one GPU (on
 * different streams, or from different threads without ordering) race and give wrong results. */
#define MI_NOALLOC_MAX_POINTS (1 << 21)
int mi_proj_find_projections_noalloc_device(mi_mesh* m, const float* d_pts, int64_t n, const float* ptXf,
                                            const float* treeXf, float upDistLimitSq, float loDistLimitSq,
                                            mi_projection* d_res, void* stream);

/* findProjection with a per-point upDistLimitSq: point i searches below min(d_upSq[i], upDistLimitSq),
 * the pattern of ICP's limitedProjector (source/MRMesh/MRMeshOrPoints.cpp:121-143 calling
 * findProjection, MRMeshProject.cpp:123-127 with upDistLimitSq = the previous distance). */
int mi_proj_find_projections_bounded_device(mi_mesh* m, const float* d_pts, int64_t n, const float* d_upSq,
                                            const float* ptXf, const float* treeXf, float upDistLimitSq,
                                            float loDistLimitSq, mi_projection* d_res, void* stream);

/* Exact projections accelerated by per-point caps (no counterpart in the reference; the learned
 * bound of this project feeds it): a first pass searches below min(d_capSq[i], upDistLimitSq), the
 * points that found nothing there are searched again without the cap.  The records equal
 * findProjections(upDistLimitSq, loDistLimitSq = 0) whatever the caps; *reruns returns how many
 * points needed the second pass.  Synchronises the stream once to size the second pass. */
int mi_proj_find_projections_capped_device(mi_mesh* m, const float* d_pts, int64_t n, const float* d_capSq,
                                           float upDistLimitSq, mi_projection* d_res, int64_t* reruns,
                                           void* stream);
/* The same for the signed-distance grid slab (calcFromGridWithDistances with minDistSq <= 0):
 * d_capSq holds one cap per output voxel of the slab. */
int mi_sdf_grid_capped_device(mi_mesh* m, const int32_t dims[3], const float xf[12], const mi_dist_options* opts,
                              const float* d_capSq, int32_t zBegin, int32_t zEnd, float* d_res,
                              int64_t* reruns, void* stream);

/* ---------------------------------------------------------------- batched persistent launches
 * No counterpart in the reference (it launches one kernel per call and per chunk,
 * source/MRCuda/MRCudaFastWindingNumber.cpp:260-280, MRCudaPointsToMeshProjector.cpp:219-227):
 * one persistent launch runs many items -- each a signed-distance grid (calcFromGridWithDistances)
 * and/or a point set (findProjections) on its own mesh -- with work units of 64 voxels or 64
 * queries taken from per-XCD ticket counters.  Outputs equal the per-item calls. */
typedef struct mi_batch_item {
    mi_mesh* mesh;
    int32_t dims[3];          /* grid dimensions; any 0 = no grid */
    float xf[12];             /* gridToMeshXf, row-major A then b */
    float* d_sdf;             /* dims.x*dims.y*dims.z floats, x-fastest */
    const float* d_pts;       /* numPts float3 points; 0 = no projections */
    int64_t numPts;
    mi_projection* d_proj;    /* numPts records */
} mi_batch_item;

typedef struct mi_batch mi_batch;
int  mi_batch_create(mi_batch** out);
void mi_batch_destroy(mi_batch* b);
/* opts apply to every grid (minDistSq must be <= 0); upDistLimitSq to every point set
 * (loDistLimitSq = 0) */
int  mi_batch_run_device(mi_batch* b, const mi_batch_item* items, int32_t count, const mi_dist_options* opts,
                         float upDistLimitSq, void* stream);
/* the project's autotuner (role of runProfiling, source/MRCuda/MRCudaProfiling.cuh:22-79): times the
 * batch over resident workgroups per CU {2..6} x XCD partitioning {off, on} and keeps the fastest
 * for later runs; *bestMs = its time (the outputs are computed as a side effect) */
int  mi_batch_autotune(mi_batch* b, const mi_batch_item* items, int32_t count, const mi_dist_options* opts,
                       float upDistLimitSq, void* stream, float* bestMs);
int  mi_batch_set_config(mi_batch* b, int32_t blocksPerCU, int32_t xcdSplit);
int  mi_batch_get_config(mi_batch* b, int32_t* blocksPerCU, int32_t* xcdSplit);
/* the batch with per-query caps (squared) on the search radius: d_capSdf[i] float[voxels of item i]
 * and d_capProj[i] float[numPts of item i] (device; either array or entry may be NULL = no cap).
 * One persistent launch with the caps; a wave whose queries found nothing below a cap smaller than
 * the plain limit walks once more for them, uncapped (*reruns = their count).  Any cap at or above a
 * query's true distance gives the plain result, so outputs equal mi_batch_run_device's.
 * Synchronises the stream (to read the count). */
int  mi_batch_run_capped_device(mi_batch* b, const mi_batch_item* items, int32_t count, const float* const* d_capSdf,
                                const float* const* d_capProj, const mi_dist_options* opts, float upDistLimitSq,
                                void* stream, int64_t* reruns);

/* ---------------------------------------------------------------- learned distance bound
 * No counterpart in the reference: a Fourier-feature MLP (99 features -> 4 x 128 SiLU -> softplus)
 * predicts each query's unsigned distance; the prediction caps the search radius at
 * (pred * (1 + margin) + slack)^2 (slack: 0, or the network's calibrated allowance, mi_mlp_calibrate)
 * and the capped forms above keep the outputs exact.  Grid forms predict once per 4x4x4 brick, at its
 * centre c, and cap voxel x at (pred(c) * (1 + margin) + slack + |x - c|)^2: the distance is
 * 1-Lipschitz, so this bounds x's distance whenever the brick's radius bounds c's.  Trained and
 * evaluated on the GPU with BF16 MFMA kernels accumulating in FP32. */
typedef struct mi_mlp mi_mlp;
/* network for points around `center`, coordinates normalised by `scale` (> 0); seeded Xavier init */
int  mi_mlp_create(const float center[3], float scale, uint64_t seed, mi_mlp** out);
void mi_mlp_destroy(mi_mlp* m);
int  mi_mlp_num_params(void);
/* FP32 parameter vector: W_l [4][128 out][128 in], b_l [4][128], w5 [128], b5 */
int  mi_mlp_get_params(mi_mlp* m, float* host);
int  mi_mlp_set_params(mi_mlp* m, const float* host);
/* gradient of 1/(2n) sum_i (pred_i - label_i)^2 / scale^2 over the n points, into d_grad[params] */
int  mi_mlp_gradient_device(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, float* d_grad,
                            void* stream);
/* `steps` Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) over consecutive batches of d_pts/d_labels */
int  mi_mlp_train_device(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, int32_t steps,
                         int32_t batch, float lr, void* stream);
/* predicted unsigned distance per point */
int  mi_mlp_predict_device(mi_mlp* m, const float* d_pts, int64_t n, float* d_dist, void* stream);
/* sets the network's slack to the `quantile` (0..1) of label - prediction over the n labelled points
 * (at least 0), so that about a (1 - quantile) share of such queries exceed their cap; *slack = the value.
 * Synchronises the stream.  mi_mlp_set_slack sets it directly (0 = the uncalibrated cap). */
int  mi_mlp_calibrate(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, float quantile, float* slack,
                      void* stream);
int  mi_mlp_set_slack(mi_mlp* m, float slack);
/* findProjections / calcFromGridWithDistances with the learned cap; exact (see the capped forms) */
int  mi_proj_find_projections_learned_device(mi_mesh* mesh, mi_mlp* m, float margin, const float* d_pts,
                                             int64_t n, float upDistLimitSq, mi_projection* d_res,
                                             int64_t* reruns, void* stream);
int  mi_sdf_grid_learned_device(mi_mesh* mesh, mi_mlp* m, float margin, const int32_t dims[3], const float xf[12],
                                const mi_dist_options* opts, int32_t zBegin, int32_t zEnd, float* d_res,
                                int64_t* reruns, void* stream);
/* the batch (mi_batch_run_capped_device) with the caps above from each item's own network nets[i]
 * (NULL = no cap for that item); exact */
int  mi_batch_run_learned_device(mi_batch* b, const mi_batch_item* items, int32_t count, mi_mlp* const* nets,
                                 float margin, const mi_dist_options* opts, float upDistLimitSq, void* stream,
                                 int64_t* reruns);

#ifdef __cplusplus
}
#endif
#endif
