/*
 * oracle.h -- CPU restatement of MeshLib's distance-field path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links, loads or calls it.
 *
 * It restates, in plain C with the reference's floating-point operation
 * order (built with -O2 -ffp-contract=off, no fast-math):
 *   tree build      source/MRMesh/MRAABBTree.cpp:15-60,
 *                   source/MRMesh/MRAABBTreeMaker.hpp:50-165,
 *                   source/MRMesh/MRBox.h:207-220,260-270,378-396
 *   dipoles         source/MRMesh/MRDipole.cpp:11-71 (+ MRMeshMath.cpp:119-124,140-149)
 *   winding number  source/MRMesh/MRDipole.cpp:81-128, MRDipole.h:18-27
 *   closest point   source/MRMesh/MRClosestPointInTriangle.h:27-87 (double)
 *   projection      source/MRMesh/MRMeshProject.cpp:13-127
 *   signed distance source/MRMesh/MRFastWindingNumber.cpp:23-94
 *   batch projector source/MRMesh/MRPointsToMeshProjector.cpp:17-32
 * plus full-FP64 variants and O(N*F) brute-force checkers used to pin it.
 */
#ifndef MESHLIB_AMD_ORACLE_H
#define MESHLIB_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float min[3], max[3]; } OrBox;                       /* Box3f */
typedef struct { OrBox box; int32_t l, r; } OrNode;                    /* AABBTreeNode, 32 B */
typedef struct { float pos[3]; float area; float dirArea[3]; float rr; } OrDipole; /* Dipole, 32 B */
typedef struct {                                                       /* MeshProjectionResult, 32 B */
    int32_t face; float proj[3]; int32_t e; float a, b; float distSq;
} OrProj;
typedef struct {                                                       /* DistanceToMeshOptions */
    float minDistSq, maxDistSq; int32_t nullOutsideMinMax; float threshold, beta;
} OrDistOpts;
typedef struct {               /* traversal work counters (algorithmic-bytes accounting) */
    int64_t queries;
    int64_t dist_probes;       /* child-box distance computations (incl. root) */
    int64_t dist_internal;     /* internal nodes expanded by the distance DFS */
    int64_t dist_leaves;       /* leaf closest-point tests */
    int64_t fwn_visits;        /* nodes whose dipole was tested */
    int64_t fwn_leaves;        /* exact solid-angle evaluations */
} OrCounters;

typedef struct OrMesh OrMesh;

OrMesh* or_mesh_create(const float* verts, int32_t nv, const int32_t* tris, int32_t nf);
void    or_mesh_destroy(OrMesh* m);
int32_t or_mesh_num_nodes(const OrMesh* m);
void    or_mesh_get_nodes(const OrMesh* m, OrNode* out);
void    or_mesh_get_dipoles(const OrMesh* m, OrDipole* out);

/* primitives (known-answer tests) */
float or_triangle_solid_angle(const float p[3], const float a[3], const float b[3], const float c[3]);
void  or_closest_point_in_triangle_d(const double p[3], const double a[3], const double b[3],
                                     const double c[3], double proj[3], double bary[2]);

/* reference-faithful drivers; nthreads <= 0 means all host cores */
void or_find_projections(const OrMesh* m, const float* pts, int64_t n, const float* ptXf,
                         const float* treeXf, float upDistLimitSq, float loDistLimitSq,
                         OrProj* out, int32_t nthreads, OrCounters* cnt);
void or_find_projections_bounded(const OrMesh* m, const float* pts, int64_t n, const float* upSq,
                                 float lo, OrProj* out, int32_t nthreads, OrCounters* cnt);
void or_fwn_vector(const OrMesh* m, const float* pts, int64_t n, float beta, int32_t skipFace,
                   float* out, int32_t nthreads, OrCounters* cnt);
void or_fwn_grid(const OrMesh* m, const int32_t dims[3], const float xf[12], float beta,
                 float* out, int32_t nthreads, OrCounters* cnt);
void or_sdf_grid(const OrMesh* m, const int32_t dims[3], const float xf[12], const OrDistOpts* o,
                 float* out, int32_t nthreads, OrCounters* cnt);
/* signed distance at selected voxels of the grid (ids into x-fastest order) */
void or_sdf_grid_sample(const OrMesh* m, const int32_t dims[3], const float xf[12],
                        const OrDistOpts* o, const int64_t* ids, int64_t n, float* out,
                        int32_t nthreads, OrCounters* cnt);
void or_self_intersections(const OrMesh* m, float beta, uint8_t* out, int32_t nthreads);

/* full-FP64 variants of the same algorithms (error statistics) */
void or_find_projections_f64(const OrMesh* m, const float* pts, int64_t n, double* distSq,
                             int32_t* face, int32_t nthreads);
void or_fwn_vector_f64(const OrMesh* m, const float* pts, int64_t n, double beta, double* out,
                       int32_t nthreads);

/* brute force over every triangle (pins the oracle on small meshes) */
void or_bf_nearest_f64(const float* verts, const int32_t* tris, int32_t nf, const float* pts,
                       int64_t n, double* distSq, int32_t nthreads);
void or_bf_winding_f64(const float* verts, const int32_t* tris, int32_t nf, const float* pts,
                       int64_t n, double* wn, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
