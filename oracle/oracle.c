/*
 * oracle.c -- CPU restatement of MeshLib's distance-field path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Built with
 * -O2 -ffp-contract=off and no fast-math so that every float expression is
 * evaluated with the reference CPU build's rounding, in the reference's order.
 * Each function names the reference lines it restates.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

struct OrMesh {
    int32_t nv, nf;
    float* verts;
    int32_t* tris;
    int32_t nnodes;
    OrNode* nodes;
    OrDipole* dips;
    /* FP64 dipoles for the full-double variant (same tree) */
    double* dpos; double* ddir; double* drr;
};

/* ------------------------------------------------------------------ parallel for */

typedef void (*or_body_fn)(void* ctx, int64_t begin, int64_t end, OrCounters* cnt);
typedef struct {
    or_body_fn fn; void* ctx; int64_t n; int64_t grain;
    int64_t next; pthread_mutex_t mu; OrCounters total;
} OrPool;

static void* or_worker(void* arg)
{
    OrPool* p = (OrPool*)arg;
    OrCounters local; memset(&local, 0, sizeof(local));
    for (;;) {
        int64_t b = __atomic_fetch_add(&p->next, p->grain, __ATOMIC_RELAXED);
        if (b >= p->n) break;
        int64_t e = b + p->grain < p->n ? b + p->grain : p->n;
        p->fn(p->ctx, b, e, &local);
    }
    pthread_mutex_lock(&p->mu);
    p->total.queries += local.queries;
    p->total.dist_probes += local.dist_probes;
    p->total.dist_internal += local.dist_internal;
    p->total.dist_leaves += local.dist_leaves;
    p->total.fwn_visits += local.fwn_visits;
    p->total.fwn_leaves += local.fwn_leaves;
    pthread_mutex_unlock(&p->mu);
    return NULL;
}

/* ParallelFor (source/MRMesh/MRParallelFor.h:30-50) with dynamic chunks on pthreads */
static void or_parallel_for(int64_t n, int32_t nthreads, or_body_fn fn, void* ctx, OrCounters* cnt)
{
    if (n <= 0) return;
    if (nthreads <= 0) nthreads = (int32_t)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    OrPool p; memset(&p, 0, sizeof(p));
    p.fn = fn; p.ctx = ctx; p.n = n;
    p.grain = n / (nthreads * 64) + 1;
    if (p.grain > 4096) p.grain = 4096;
    pthread_mutex_init(&p.mu, NULL);
    if (nthreads == 1) {
        or_worker(&p);
    } else {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, or_worker, &p);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
        free(th);
    }
    pthread_mutex_destroy(&p.mu);
    if (cnt) {
        cnt->queries += p.total.queries;
        cnt->dist_probes += p.total.dist_probes;
        cnt->dist_internal += p.total.dist_internal;
        cnt->dist_leaves += p.total.dist_leaves;
        cnt->fwn_visits += p.total.fwn_visits;
        cnt->fwn_leaves += p.total.fwn_leaves;
    }
}

/* ------------------------------------------------------------------ box helpers */

static void box_invalid(OrBox* b)          /* Box() : min = FLT_MAX, max = lowest (MRBox.h:42) */
{
    for (int i = 0; i < 3; ++i) { b->min[i] = FLT_MAX; b->max[i] = -FLT_MAX; }
}
static void box_include_pt(OrBox* b, const float* p)   /* MRBox.h:122-129 */
{
    for (int i = 0; i < 3; ++i) {
        if (p[i] < b->min[i]) b->min[i] = p[i];
        if (p[i] > b->max[i]) b->max[i] = p[i];
    }
}
static void box_include_box(OrBox* b, const OrBox* o)  /* MRBox.h:132-139 */
{
    for (int i = 0; i < 3; ++i) {
        if (o->min[i] < b->min[i]) b->min[i] = o->min[i];
        if (o->max[i] > b->max[i]) b->max[i] = o->max[i];
    }
}
/* Box::getDistanceSq(pt), MRBox.h:207-220: only the outside dims are summed */
static float box_dist_sq(const OrBox* b, const float* p)
{
    float res = 0.0f;
    for (int i = 0; i < 3; ++i) {
        if (p[i] < b->min[i]) { float d = p[i] - b->min[i]; res += d * d; }
        else if (p[i] > b->max[i]) { float d = p[i] - b->max[i]; res += d * d; }
    }
    return res;
}
static double box_dist_sq_d(const OrBox* b, const double* p)
{
    double res = 0.0;
    for (int i = 0; i < 3; ++i) {
        if (p[i] < b->min[i]) { double d = p[i] - b->min[i]; res += d * d; }
        else if (p[i] > b->max[i]) { double d = p[i] - b->max[i]; res += d * d; }
    }
    return res;
}

/* AffineXf3f (A rows then b) applied to a point: {dot(A.x,p), dot(A.y,p), dot(A.z,p)} + b
 * (MRMatrix3.h:127-130, MRVector3.h:190) */
static void xf_apply(const float* xf, const float* p, float* out)
{
    float r[3];
    for (int i = 0; i < 3; ++i)
        r[i] = xf[3 * i + 0] * p[0] + xf[3 * i + 1] * p[1] + xf[3 * i + 2] * p[2];
    for (int i = 0; i < 3; ++i) out[i] = r[i] + xf[9 + i];
}
/* transformed(box, xf): bbox of the 8 transformed corners (MRBox.h:336-344) */
static void xf_box(const float* xf, const OrBox* b, OrBox* out)
{
    box_invalid(out);
    for (int c = 0; c < 8; ++c) {
        float p[3] = { (c & 1) ? b->max[0] : b->min[0], (c & 2) ? b->max[1] : b->min[1],
                       (c & 4) ? b->max[2] : b->min[2] };
        float q[3];
        xf_apply(xf, p, q);
        box_include_pt(out, q);
    }
}

/* ------------------------------------------------------------------ tree build */

typedef struct { int32_t leafId; OrBox box; } BoxedLeaf;

/* findSortedBoxDims (MRBox.h:378-396): dims sorted by (size, index), smallest first */
static void sorted_box_dims(const OrBox* b, int out[3])
{
    float sz[3]; int id[3];
    for (int i = 0; i < 3; ++i) { sz[i] = b->max[i] - b->min[i]; id[i] = i; }
    for (int i = 0; i + 1 < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (sz[j] < sz[i] || (sz[j] == sz[i] && id[j] < id[i])) {
                float ts = sz[i]; sz[i] = sz[j]; sz[j] = ts;
                int ti = id[i]; id[i] = id[j]; id[j] = ti;
            }
    for (int i = 0; i < 3; ++i) out[i] = id[i];
}

/* the total order of partitionLeaves_ (MRAABBTreeMaker.hpp:55-69) */
static int leaf_less(const BoxedLeaf* a, const BoxedLeaf* b, const int sd[3])
{
    for (int i = 2; i >= 0; --i) {
        int d = sd[i];
        float ad = a->box.min[d] + a->box.max[d];
        float bd = b->box.min[d] + b->box.max[d];
        if (ad != bd) return ad < bd;
    }
    return a->leafId < b->leafId;
}

static void swap_leaf(BoxedLeaf* a, BoxedLeaf* b) { BoxedLeaf t = *a; *a = *b; *b = t; }

/* places the k-th smallest element (in the total order) at position k with every
 * smaller element before it; under a total order the resulting SET split is unique,
 * which is all that the tree depends on (std::nth_element's role at MRAABBTreeMaker.hpp:72) */
static void select_kth(BoxedLeaf* a, int64_t n, int64_t k, const int sd[3])
{
    int64_t lo = 0, hi = n - 1;
    while (hi > lo) {
        int64_t mid = lo + (hi - lo) / 2;
        if (leaf_less(&a[mid], &a[lo], sd)) swap_leaf(&a[mid], &a[lo]);
        if (leaf_less(&a[hi], &a[lo], sd)) swap_leaf(&a[hi], &a[lo]);
        if (leaf_less(&a[hi], &a[mid], sd)) swap_leaf(&a[hi], &a[mid]);
        BoxedLeaf pivot = a[mid];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (leaf_less(&a[i], &pivot, sd)) ++i;
            while (leaf_less(&pivot, &a[j], sd)) --j;
            if (i <= j) { swap_leaf(&a[i], &a[j]); ++i; --j; }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else return;
    }
}

static void tri_points(const OrMesh* m, int32_t f, float* a, float* b, float* c)
{
    const int32_t* t = m->tris + 3 * (int64_t)f;
    memcpy(a, m->verts + 3 * (int64_t)t[0], 12);
    memcpy(b, m->verts + 3 * (int64_t)t[1], 12);
    memcpy(c, m->verts + 3 * (int64_t)t[2], 12);
}

/* computeFaceBox (MRAABBTree.cpp:15-29) + insignificantlyExpanded (MRBox.h:260-270) */
static void face_box(const OrMesh* m, int32_t f, OrBox* box)
{
    float a[3], b[3], c[3];
    tri_points(m, f, a, b, c);
    box_invalid(box);
    box_include_pt(box, a); box_include_pt(box, b); box_include_pt(box, c);
    for (int i = 0; i < 3; ++i) {
        box->min[i] = nextafterf(box->min[i], -FLT_MAX);
        box->max[i] = nextafterf(box->max[i], FLT_MAX);
    }
}

typedef struct { int32_t root, first, num; } Subtree;

/* a face of the triangulation that is not deleted: getTriangulation (MRMeshTopology.cpp:430-441)
 * writes only valid faces, the others keep {-1,-1,-1}; the reference's GPU kernels test any index
 * < 0 (MRCudaFastWindingNumber.cu:168) */
static int face_valid(const OrMesh* m, int32_t f)
{
    const int32_t* t = m->tris + 3 * (int64_t)f;
    return t[0] >= 0 && t[1] >= 0 && t[2] >= 0;
}

/* AABBTree(MeshPart) + AABBTreeMaker::construct/makeNode_/makeSubtree_
 * (MRAABBTree.cpp:31-60, MRAABBTreeMaker.hpp:76-165); serial, same result.  The leaves are the
 * valid faces with their own ids (MRAABBTree.cpp:35-49, the "not packed" branch). */
static void build_tree(OrMesh* m)
{
    int32_t nf = 0;
    for (int32_t f = 0; f < m->nf; ++f) nf += face_valid(m, f);
    m->nnodes = nf > 0 ? 2 * nf - 1 : 0;
    if (nf <= 0) { m->nodes = NULL; return; }
    BoxedLeaf* leaves = (BoxedLeaf*)malloc(sizeof(BoxedLeaf) * (size_t)nf);
    for (int32_t f = 0, i = 0; f < m->nf; ++f) {
        if (!face_valid(m, f)) continue;
        leaves[i].leafId = f;
        face_box(m, f, &leaves[i].box);
        ++i;
    }
    m->nodes = (OrNode*)malloc(sizeof(OrNode) * (size_t)m->nnodes);
    Subtree* stack = (Subtree*)malloc(sizeof(Subtree) * 128);
    int sp = 0;
    stack[sp++] = (Subtree){ 0, 0, nf };
    while (sp > 0) {
        Subtree x = stack[--sp];
        OrNode* node = &m->nodes[x.root];
        if (x.num == 1) {                       /* setLeafId: l = face, r = invalid (-1) */
            node->box = leaves[x.first].box;
            node->l = leaves[x.first].leafId;
            node->r = -1;
            continue;
        }
        box_invalid(&node->box);
        for (int32_t i = 0; i < x.num; ++i) box_include_box(&node->box, &leaves[x.first + i].box);
        int sd[3];
        sorted_box_dims(&node->box, sd);
        int32_t half = x.num / 2;
        select_kth(leaves + x.first, x.num, half, sd);
        int32_t lnum = half, rnum = x.num - half;
        node->l = x.root + 1;
        node->r = x.root + 1 + (2 * lnum - 1);
        stack[sp++] = (Subtree){ node->r, x.first + half, rnum };
        stack[sp++] = (Subtree){ node->l, x.first, lnum };
    }
    free(stack);
    free(leaves);
}

/* calcDipoles (MRDipole.cpp:23-71), distToFarthestCornerSq (:11-21) */
static float dist_to_farthest_corner_sq(const OrBox* b, const float* pos)
{
    float res = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float tmin = pos[i] - b->min[i]; tmin = tmin * tmin;
        float tmax = pos[i] - b->max[i]; tmax = tmax * tmax;
        res += tmin > tmax ? tmin : tmax;       /* std::max(toMinSq, toMaxSq) */
    }
    return res;
}

static void calc_dipoles(OrMesh* m)
{
    const int32_t nn = m->nnodes;
    m->dips = (OrDipole*)calloc((size_t)(nn > 0 ? nn : 1), sizeof(OrDipole));
    for (int32_t i = 0; i < nn; ++i) {
        const OrNode* node = &m->nodes[i];
        if (node->r >= 0) continue;
        float a[3], b[3], c[3];
        tri_points(m, node->l, a, b, c);
        /* dirDblArea = cross(b - a, c - a) (MRMeshMath.cpp:140-149) */
        float ab[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
        float ac[3] = { c[0] - a[0], c[1] - a[1], c[2] - a[2] };
        float cr[3] = { ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                        ab[0] * ac[1] - ab[1] * ac[0] };
        float da[3] = { 0.5f * cr[0], 0.5f * cr[1], 0.5f * cr[2] };
        float area = sqrtf(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
        /* triCenter = (1/3.0f) * (v0 + v1 + v2) (MRMeshMath.cpp:119-124) */
        const float third = 1 / 3.0f;
        float cen[3];
        for (int k = 0; k < 3; ++k) cen[k] = third * ((a[k] + b[k]) + c[k]);
        OrDipole* d = &m->dips[i];
        for (int k = 0; k < 3; ++k) { d->pos[k] = area * cen[k]; d->dirArea[k] = da[k]; }
        d->area = area;
        d->rr = 0.0f;
    }
    for (int32_t i = nn - 1; i >= 0; --i) {     /* serial bottom-up pass, MRDipole.cpp:47-60 */
        const OrNode* node = &m->nodes[i];
        if (node->r < 0) continue;
        const OrDipole* dl = &m->dips[node->l];
        const OrDipole* dr = &m->dips[node->r];
        OrDipole* d = &m->dips[i];
        for (int k = 0; k < 3; ++k) {
            d->pos[k] = dl->pos[k] + dr->pos[k];
            d->dirArea[k] = dl->dirArea[k] + dr->dirArea[k];
        }
        d->area = dl->area + dr->area;
        d->rr = 0.0f;
    }
    for (int32_t i = 0; i < nn; ++i) {           /* MRDipole.cpp:63-70: pos /= area is pos *= 1/area */
        OrDipole* d = &m->dips[i];
        if (d->area > 0) {
            float inv = 1 / d->area;
            for (int k = 0; k < 3; ++k) d->pos[k] *= inv;
        }
        d->rr = dist_to_farthest_corner_sq(&m->nodes[i].box, d->pos);
    }
}

/* same dipoles in full double precision, for the FP64 winding-number variant */
static void calc_dipoles_f64(OrMesh* m)
{
    const int32_t nn = m->nnodes;
    size_t cnt = (size_t)(nn > 0 ? nn : 1);
    m->dpos = (double*)calloc(cnt * 4, sizeof(double));   /* pos xyz, area */
    m->ddir = (double*)calloc(cnt * 3, sizeof(double));
    m->drr = (double*)calloc(cnt, sizeof(double));
    for (int32_t i = 0; i < nn; ++i) {
        const OrNode* node = &m->nodes[i];
        if (node->r >= 0) continue;
        float af[3], bf[3], cf[3];
        tri_points(m, node->l, af, bf, cf);
        double a[3], b[3], c[3];
        for (int k = 0; k < 3; ++k) { a[k] = af[k]; b[k] = bf[k]; c[k] = cf[k]; }
        double ab[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
        double ac[3] = { c[0] - a[0], c[1] - a[1], c[2] - a[2] };
        double da[3] = { 0.5 * (ab[1] * ac[2] - ab[2] * ac[1]), 0.5 * (ab[2] * ac[0] - ab[0] * ac[2]),
                         0.5 * (ab[0] * ac[1] - ab[1] * ac[0]) };
        double area = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
        for (int k = 0; k < 3; ++k) {
            m->dpos[4 * (size_t)i + k] = area * (a[k] + b[k] + c[k]) / 3.0;
            m->ddir[3 * (size_t)i + k] = da[k];
        }
        m->dpos[4 * (size_t)i + 3] = area;
    }
    for (int32_t i = nn - 1; i >= 0; --i) {
        const OrNode* node = &m->nodes[i];
        if (node->r < 0) continue;
        for (int k = 0; k < 4; ++k)
            m->dpos[4 * (size_t)i + k] = m->dpos[4 * (size_t)node->l + k] + m->dpos[4 * (size_t)node->r + k];
        for (int k = 0; k < 3; ++k)
            m->ddir[3 * (size_t)i + k] = m->ddir[3 * (size_t)node->l + k] + m->ddir[3 * (size_t)node->r + k];
    }
    for (int32_t i = 0; i < nn; ++i) {
        double* p = m->dpos + 4 * (size_t)i;
        if (p[3] > 0) { p[0] /= p[3]; p[1] /= p[3]; p[2] /= p[3]; }
        const OrBox* b = &m->nodes[i].box;
        double rr = 0;
        for (int k = 0; k < 3; ++k) {
            double t0 = p[k] - b->min[k], t1 = p[k] - b->max[k];
            t0 *= t0; t1 *= t1;
            rr += t0 > t1 ? t0 : t1;
        }
        m->drr[i] = rr;
    }
}

OrMesh* or_mesh_create(const float* verts, int32_t nv, const int32_t* tris, int32_t nf)
{
    OrMesh* m = (OrMesh*)calloc(1, sizeof(OrMesh));
    m->nv = nv; m->nf = nf;
    m->verts = (float*)malloc(sizeof(float) * 3 * (size_t)(nv > 0 ? nv : 1));
    m->tris = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(nf > 0 ? nf : 1));
    if (nv > 0) memcpy(m->verts, verts, sizeof(float) * 3 * (size_t)nv);
    if (nf > 0) memcpy(m->tris, tris, sizeof(int32_t) * 3 * (size_t)nf);
    build_tree(m);
    calc_dipoles(m);
    calc_dipoles_f64(m);
    return m;
}

void or_mesh_destroy(OrMesh* m)
{
    if (!m) return;
    free(m->verts); free(m->tris); free(m->nodes); free(m->dips);
    free(m->dpos); free(m->ddir); free(m->drr);
    free(m);
}

int32_t or_mesh_num_nodes(const OrMesh* m) { return m->nnodes; }
void or_mesh_get_nodes(const OrMesh* m, OrNode* out)
{
    if (m->nnodes) memcpy(out, m->nodes, sizeof(OrNode) * (size_t)m->nnodes);
}
void or_mesh_get_dipoles(const OrMesh* m, OrDipole* out)
{
    if (m->nnodes) memcpy(out, m->dips, sizeof(OrDipole) * (size_t)m->nnodes);
}

/* ------------------------------------------------------------------ primitives */

/* triangleSolidAngle (MRDipole.cpp:81-92; det from MRMatrix3.h:226-232) */
float or_triangle_solid_angle(const float p[3], const float a[3], const float b[3], const float c[3])
{
    float mx[3], my[3], mz[3];
    for (int k = 0; k < 3; ++k) { mx[k] = a[k] - p[k]; my[k] = b[k] - p[k]; mz[k] = c[k] - p[k]; }
    float x = sqrtf(mx[0] * mx[0] + mx[1] * mx[1] + mx[2] * mx[2]);
    float y = sqrtf(my[0] * my[0] + my[1] * my[1] + my[2] * my[2]);
    float z = sqrtf(mz[0] * mz[0] + mz[1] * mz[1] + mz[2] * mz[2]);
    float dxy = mx[0] * my[0] + mx[1] * my[1] + mx[2] * my[2];
    float dyz = my[0] * mz[0] + my[1] * mz[1] + my[2] * mz[2];
    float dzx = mz[0] * mx[0] + mz[1] * mx[1] + mz[2] * mx[2];
    float den = x * y * z + dxy * z + dyz * x + dzx * y;
    float det = mx[0] * (my[1] * mz[2] - my[2] * mz[1])
              - mx[1] * (my[0] * mz[2] - my[2] * mz[0])
              + mx[2] * (my[0] * mz[1] - my[1] * mz[0]);
    return 2 * atan2f(det, den);
}

static double solid_angle_d(const double* p, const double* a, const double* b, const double* c)
{
    double mx[3], my[3], mz[3];
    for (int k = 0; k < 3; ++k) { mx[k] = a[k] - p[k]; my[k] = b[k] - p[k]; mz[k] = c[k] - p[k]; }
    double x = sqrt(mx[0] * mx[0] + mx[1] * mx[1] + mx[2] * mx[2]);
    double y = sqrt(my[0] * my[0] + my[1] * my[1] + my[2] * my[2]);
    double z = sqrt(mz[0] * mz[0] + mz[1] * mz[1] + mz[2] * mz[2]);
    double den = x * y * z + (mx[0] * my[0] + mx[1] * my[1] + mx[2] * my[2]) * z
               + (my[0] * mz[0] + my[1] * mz[1] + my[2] * mz[2]) * x
               + (mz[0] * mx[0] + mz[1] * mx[1] + mz[2] * mx[2]) * y;
    double det = mx[0] * (my[1] * mz[2] - my[2] * mz[1]) - mx[1] * (my[0] * mz[2] - my[2] * mz[0])
               + mx[2] * (my[0] * mz[1] - my[1] * mz[0]);
    return 2 * atan2(det, den);
}

static double dot3d(const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

/* closestPointInTriangle<double> (MRClosestPointInTriangle.h:27-87), regions #1..#6, #0 */
void or_closest_point_in_triangle_d(const double p[3], const double a[3], const double b[3],
                                    const double c[3], double proj[3], double bary[2])
{
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
    const double d1 = dot3d(ab, ap), d2 = dot3d(ac, ap);
    if (d1 <= 0 && d2 <= 0) { for (int k = 0; k < 3; ++k) proj[k] = a[k]; bary[0] = 0; bary[1] = 0; return; }
    for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
    const double d3 = dot3d(ab, bp), d4 = dot3d(ac, bp);
    if (d3 >= 0 && d4 <= d3) { for (int k = 0; k < 3; ++k) proj[k] = b[k]; bary[0] = 1; bary[1] = 0; return; }
    for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
    const double d5 = dot3d(ab, cp), d6 = dot3d(ac, cp);
    if (d6 >= 0 && d5 <= d6) { for (int k = 0; k < 3; ++k) proj[k] = c[k]; bary[0] = 0; bary[1] = 1; return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        for (int k = 0; k < 3; ++k) proj[k] = a[k] + v * ab[k];
        bary[0] = v; bary[1] = 0; return;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d6 <= 0 && d2 >= 0) {
        const double v = d2 / (d2 - d6);
        for (int k = 0; k < 3; ++k) proj[k] = a[k] + v * ac[k];
        bary[0] = 0; bary[1] = v; return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0) {
        if (d4 < d3) { for (int k = 0; k < 3; ++k) proj[k] = b[k]; bary[0] = 1; bary[1] = 0; return; }
        if (d5 < d6) { for (int k = 0; k < 3; ++k) proj[k] = c[k]; bary[0] = 0; bary[1] = 1; return; }
        const double v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; ++k) proj[k] = b[k] + v * (c[k] - b[k]);
        bary[0] = 1 - v; bary[1] = v; return;
    }
    const double denom = 1 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; ++k) proj[k] = a[k] + v * ab[k] + w * ac[k];
    bary[0] = v; bary[1] = w;
}

/* ------------------------------------------------------------------ traversals */

/* findProjectionSubtree (MRMeshProject.cpp:13-103) without predicates/regions */
static OrProj find_projection(const OrMesh* m, const float* pt, float up, const float* treeXf,
                              float lo, OrCounters* cnt)
{
    OrProj res;
    memset(&res, 0, sizeof(res));
    res.face = -1; res.e = -1;
    res.distSq = up;
    if (m->nnodes == 0) return res;
    typedef struct { int32_t n; float distSq; } SubTask;
    SubTask st[32];
    int sp = 0;
    int64_t probes = 0, internal = 0, leaves = 0;
#define OR_BOXDIST(nid, outd) do { \
        ++probes; \
        if (treeXf) { OrBox tb_; xf_box(treeXf, &m->nodes[nid].box, &tb_); outd = box_dist_sq(&tb_, pt); } \
        else outd = box_dist_sq(&m->nodes[nid].box, pt); } while (0)
    {
        float d0; OR_BOXDIST(0, d0);
        if (d0 < res.distSq) { st[sp].n = 0; st[sp].distSq = d0; ++sp; }
    }
    while (sp > 0) {
        SubTask s = st[--sp];
        const OrNode* node = &m->nodes[s.n];
        if (s.distSq >= res.distSq) continue;
        if (node->r < 0) {
            ++leaves;
            const int32_t face = node->l;
            float a[3], b[3], c[3];
            tri_points(m, face, a, b, c);
            if (treeXf) {
                float t[3];
                xf_apply(treeXf, a, t); memcpy(a, t, 12);
                xf_apply(treeXf, b, t); memcpy(b, t, 12);
                xf_apply(treeXf, c, t); memcpy(c, t, 12);
            }
            double pd[3] = { pt[0], pt[1], pt[2] }, ad[3] = { a[0], a[1], a[2] },
                   bd[3] = { b[0], b[1], b[2] }, cd[3] = { c[0], c[1], c[2] };
            double projD[3], baryD[2];
            or_closest_point_in_triangle_d(pd, ad, bd, cd, projD, baryD);
            float proj[3] = { (float)projD[0], (float)projD[1], (float)projD[2] };
            float dx = proj[0] - pt[0], dy = proj[1] - pt[1], dz = proj[2] - pt[2];
            float distSq = dx * dx + dy * dy + dz * dz;
            if (distSq < res.distSq) {
                res.face = face;
                memcpy(res.proj, proj, 12);
                res.e = -1;   /* edgeWithLeft(face) is topology-dependent; filled by the caller */
                res.a = (float)baryD[0];
                res.b = (float)baryD[1];
                res.distSq = distSq;
                if (res.distSq <= lo) break;
            }
            continue;
        }
        ++internal;
        float d1, d2;
        OR_BOXDIST(node->l, d1);
        OR_BOXDIST(node->r, d2);
        /* push the farther child first, so the nearer is explored first; equal -> right first */
        if (d1 < d2) {
            if (d2 < res.distSq) { st[sp].n = node->r; st[sp].distSq = d2; ++sp; }
            if (d1 < res.distSq) { st[sp].n = node->l; st[sp].distSq = d1; ++sp; }
        } else {
            if (d1 < res.distSq) { st[sp].n = node->l; st[sp].distSq = d1; ++sp; }
            if (d2 < res.distSq) { st[sp].n = node->r; st[sp].distSq = d2; ++sp; }
        }
    }
#undef OR_BOXDIST
    if (cnt) {
        cnt->dist_probes += probes; cnt->dist_internal += internal; cnt->dist_leaves += leaves;
    }
    return res;
}

/* calcFastWindingNumber (MRDipole.cpp:94-128) + Dipole::addIfGoodApprox (MRDipole.h:18-27) */
static float calc_fwn(const OrMesh* m, const float* q, float beta, int32_t skipFace, OrCounters* cnt)
{
    if (m->nnodes == 0) return 0;
    const float betaSq = beta * beta;
    int32_t st[32];
    int sp = 0;
    st[sp++] = 0;
    float res = 0;
    int64_t visits = 0, leaves = 0;
    while (sp > 0) {
        const int32_t i = st[--sp];
        const OrNode* node = &m->nodes[i];
        const OrDipole* d = &m->dips[i];
        ++visits;
        float dp[3] = { d->pos[0] - q[0], d->pos[1] - q[1], d->pos[2] - q[2] };
        float dd = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
        if (!(dd <= betaSq * d->rr)) {
            float dist = sqrtf(dd);
            if (dist > 0)
                res += (dp[0] * d->dirArea[0] + dp[1] * d->dirArea[1] + dp[2] * d->dirArea[2]) / (dist * dd);
            continue;
        }
        if (node->r >= 0) {
            st[sp++] = node->r;
            st[sp++] = node->l;
            continue;
        }
        if (node->l != skipFace) {
            ++leaves;
            float a[3], b[3], c[3];
            tri_points(m, node->l, a, b, c);
            res += or_triangle_solid_angle(q, a, b, c);
        }
    }
    if (cnt) { cnt->fwn_visits += visits; cnt->fwn_leaves += leaves; }
    const float INV_4PI = 1.0f / (4 * 3.14159265358979f);
    return INV_4PI * res;
}

/* FastWindingNumber::calcWithDistances (MRFastWindingNumber.cpp:69-76) */
static float calc_with_distances(const OrMesh* m, const float* p, const OrDistOpts* o, OrCounters* cnt)
{
    float resSq = find_projection(m, p, o->maxDistSq, NULL, o->minDistSq, cnt).distSq;
    if (o->nullOutsideMinMax && (resSq < o->minDistSq || resSq >= o->maxDistSq))
        return NAN;
    const float sign = calc_fwn(m, p, o->beta, -1, cnt) > o->threshold ? -1.f : +1.f;
    return sign * sqrtf(resSq);
}

/* VolumeIndexer::toPos (MRVolumeIndexer.h:131-139) then gridToMeshXf */
static void grid_point(const int32_t* dims, const float* xf, int64_t id, float* out)
{
    int64_t sxy = (int64_t)dims[0] * dims[1];
    int z = (int)(id / sxy);
    int sumZ = (int)(id % sxy);
    int y = sumZ / dims[0];
    int x = sumZ % dims[0];
    float p[3] = { (float)x, (float)y, (float)z };
    xf_apply(xf, p, out);
}

/* ------------------------------------------------------------------ drivers */

typedef struct {
    const OrMesh* m; const float* pts; const float* ptXf; const float* treeXf;
    float up, lo; OrProj* out;
} ProjCtx;
static void proj_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    ProjCtx* x = (ProjCtx*)c;
    for (int64_t i = b; i < e; ++i) {
        float p[3];
        if (x->ptXf) xf_apply(x->ptXf, x->pts + 3 * i, p);
        else memcpy(p, x->pts + 3 * i, 12);
        x->out[i] = find_projection(x->m, p, x->up, x->treeXf, x->lo, cnt);
    }
    cnt->queries += e - b;
}
/* PointsToMeshProjector::findProjections (MRPointsToMeshProjector.cpp:17-32); the caller
 * passes the already simplified transforms (createProjectionTransforms, MRMeshProject.cpp:105-121) */
void or_find_projections(const OrMesh* m, const float* pts, int64_t n, const float* ptXf,
                         const float* treeXf, float up, float lo, OrProj* out, int32_t nthreads,
                         OrCounters* cnt)
{
    ProjCtx c = { m, pts, ptXf, treeXf, up, lo, out };
    or_parallel_for(n, nthreads, proj_body, &c, cnt);
}


typedef struct { const OrMesh* m; const float* pts; const float* ups; float lo; OrProj* out; } BProjCtx;
static void bproj_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    BProjCtx* x = (BProjCtx*)c;
    for (int64_t i = b; i < e; ++i)
        x->out[i] = find_projection(x->m, x->pts + 3 * i, x->ups[i], NULL, x->lo, cnt);
    cnt->queries += e - b;
}
/* findProjection with a per-point upDistLimitSq, as ICP's limitedProjector feeds it
 * (source/MRMesh/MRMeshOrPoints.cpp:121-143) */
void or_find_projections_bounded(const OrMesh* m, const float* pts, int64_t n, const float* upSq,
                                 float lo, OrProj* out, int32_t nthreads, OrCounters* cnt)
{
    BProjCtx c = { m, pts, upSq, lo, out };
    or_parallel_for(n, nthreads, bproj_body, &c, cnt);
}

typedef struct {
    const OrMesh* m; const float* pts; float beta; int32_t skip; float* out;
    const int32_t* dims; const float* xf; const OrDistOpts* o; const int64_t* ids;
} FwnCtx;
static void fwn_vec_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    FwnCtx* x = (FwnCtx*)c;
    for (int64_t i = b; i < e; ++i) x->out[i] = calc_fwn(x->m, x->pts + 3 * i, x->beta, x->skip, cnt);
    cnt->queries += e - b;
}
/* FastWindingNumber::calcFromVector (MRFastWindingNumber.cpp:28-38) */
void or_fwn_vector(const OrMesh* m, const float* pts, int64_t n, float beta, int32_t skipFace,
                   float* out, int32_t nthreads, OrCounters* cnt)
{
    FwnCtx c; memset(&c, 0, sizeof(c));
    c.m = m; c.pts = pts; c.beta = beta; c.skip = skipFace; c.out = out;
    or_parallel_for(n, nthreads, fwn_vec_body, &c, cnt);
}

static void fwn_grid_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    FwnCtx* x = (FwnCtx*)c;
    for (int64_t i = b; i < e; ++i) {
        float p[3];
        grid_point(x->dims, x->xf, i, p);
        x->out[i] = calc_fwn(x->m, p, x->beta, -1, cnt);
    }
    cnt->queries += e - b;
}
/* FastWindingNumber::calcFromGrid (MRFastWindingNumber.cpp:54-67) */
void or_fwn_grid(const OrMesh* m, const int32_t dims[3], const float xf[12], float beta,
                 float* out, int32_t nthreads, OrCounters* cnt)
{
    FwnCtx c; memset(&c, 0, sizeof(c));
    c.m = m; c.beta = beta; c.out = out; c.dims = dims; c.xf = xf;
    or_parallel_for((int64_t)dims[0] * dims[1] * dims[2], nthreads, fwn_grid_body, &c, cnt);
}

static void sdf_grid_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    FwnCtx* x = (FwnCtx*)c;
    for (int64_t i = b; i < e; ++i) {
        int64_t id = x->ids ? x->ids[i] : i;
        float p[3];
        grid_point(x->dims, x->xf, id, p);
        x->out[i] = calc_with_distances(x->m, p, x->o, cnt);
    }
    cnt->queries += e - b;
}
/* FastWindingNumber::calcFromGridWithDistances (MRFastWindingNumber.cpp:78-94) */
void or_sdf_grid(const OrMesh* m, const int32_t dims[3], const float xf[12], const OrDistOpts* o,
                 float* out, int32_t nthreads, OrCounters* cnt)
{
    FwnCtx c; memset(&c, 0, sizeof(c));
    c.m = m; c.out = out; c.dims = dims; c.xf = xf; c.o = o;
    or_parallel_for((int64_t)dims[0] * dims[1] * dims[2], nthreads, sdf_grid_body, &c, cnt);
}
void or_sdf_grid_sample(const OrMesh* m, const int32_t dims[3], const float xf[12],
                        const OrDistOpts* o, const int64_t* ids, int64_t n, float* out,
                        int32_t nthreads, OrCounters* cnt)
{
    FwnCtx c; memset(&c, 0, sizeof(c));
    c.m = m; c.out = out; c.dims = dims; c.xf = xf; c.o = o; c.ids = ids;
    or_parallel_for(n, nthreads, sdf_grid_body, &c, cnt);
}

typedef struct { const OrMesh* m; float beta; uint8_t* out; } SelfCtx;
static void self_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    SelfCtx* x = (SelfCtx*)c;
    const float third = 1 / 3.0f;
    for (int64_t f = b; f < e; ++f) {
        float a[3], bb[3], cc[3], q[3];
        x->out[f] = 0;
        if (!face_valid(x->m, (int32_t)f)) continue;     /* BitSetParallelFor over valid faces */
        tri_points(x->m, (int32_t)f, a, bb, cc);
        for (int k = 0; k < 3; ++k) q[k] = third * ((a[k] + bb[k]) + cc[k]);
        float wn = calc_fwn(x->m, q, x->beta, (int32_t)f, cnt);
        x->out[f] = (wn < 0 || wn > 1) ? 1 : 0;
    }
}
/* FastWindingNumber::calcSelfIntersections (MRFastWindingNumber.cpp:40-52) */
void or_self_intersections(const OrMesh* m, float beta, uint8_t* out, int32_t nthreads)
{
    SelfCtx c = { m, beta, out };
    or_parallel_for(m->nf, nthreads, self_body, &c, NULL);
}

/* ------------------------------------------------------------------ FP64 variants */

typedef struct { const OrMesh* m; const float* pts; double* dist; int32_t* face; double beta; double* wn; } F64Ctx;

static void proj_f64_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    F64Ctx* x = (F64Ctx*)c;
    const OrMesh* m = x->m;
    (void)cnt;
    for (int64_t qi = b; qi < e; ++qi) {
        double pt[3] = { x->pts[3 * qi], x->pts[3 * qi + 1], x->pts[3 * qi + 2] };
        double best = INFINITY;
        int32_t bestFace = -1;
        if (m->nnodes) {
            typedef struct { int32_t n; double d; } T;
            T st[32]; int sp = 0;
            st[sp].n = 0; st[sp].d = box_dist_sq_d(&m->nodes[0].box, pt); ++sp;
            while (sp > 0) {
                T s = st[--sp];
                if (s.d >= best) continue;
                const OrNode* node = &m->nodes[s.n];
                if (node->r < 0) {
                    float af[3], bf[3], cf[3];
                    tri_points(m, node->l, af, bf, cf);
                    double a[3] = { af[0], af[1], af[2] }, bb[3] = { bf[0], bf[1], bf[2] },
                           cc[3] = { cf[0], cf[1], cf[2] }, pr[3], ba[2];
                    or_closest_point_in_triangle_d(pt, a, bb, cc, pr, ba);
                    double d = (pr[0] - pt[0]) * (pr[0] - pt[0]) + (pr[1] - pt[1]) * (pr[1] - pt[1])
                             + (pr[2] - pt[2]) * (pr[2] - pt[2]);
                    if (d < best) { best = d; bestFace = node->l; }
                    continue;
                }
                double d1 = box_dist_sq_d(&m->nodes[node->l].box, pt);
                double d2 = box_dist_sq_d(&m->nodes[node->r].box, pt);
                if (d1 < d2) {
                    if (d2 < best) { st[sp].n = node->r; st[sp].d = d2; ++sp; }
                    if (d1 < best) { st[sp].n = node->l; st[sp].d = d1; ++sp; }
                } else {
                    if (d1 < best) { st[sp].n = node->l; st[sp].d = d1; ++sp; }
                    if (d2 < best) { st[sp].n = node->r; st[sp].d = d2; ++sp; }
                }
            }
        }
        x->dist[qi] = best;
        if (x->face) x->face[qi] = bestFace;
    }
}
void or_find_projections_f64(const OrMesh* m, const float* pts, int64_t n, double* distSq,
                             int32_t* face, int32_t nthreads)
{
    F64Ctx c; memset(&c, 0, sizeof(c));
    c.m = m; c.pts = pts; c.dist = distSq; c.face = face;
    or_parallel_for(n, nthreads, proj_f64_body, &c, NULL);
}

static void fwn_f64_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    F64Ctx* x = (F64Ctx*)c;
    const OrMesh* m = x->m;
    (void)cnt;
    const double betaSq = x->beta * x->beta;
    for (int64_t qi = b; qi < e; ++qi) {
        double q[3] = { x->pts[3 * qi], x->pts[3 * qi + 1], x->pts[3 * qi + 2] };
        double res = 0;
        if (m->nnodes) {
            int32_t st[32]; int sp = 0;
            st[sp++] = 0;
            while (sp > 0) {
                int32_t i = st[--sp];
                const OrNode* node = &m->nodes[i];
                const double* pos = m->dpos + 4 * (size_t)i;
                const double* dir = m->ddir + 3 * (size_t)i;
                double dp[3] = { pos[0] - q[0], pos[1] - q[1], pos[2] - q[2] };
                double dd = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                if (!(dd <= betaSq * m->drr[i])) {
                    double dist = sqrt(dd);
                    if (dist > 0) res += (dp[0] * dir[0] + dp[1] * dir[1] + dp[2] * dir[2]) / (dist * dd);
                    continue;
                }
                if (node->r >= 0) { st[sp++] = node->r; st[sp++] = node->l; continue; }
                float af[3], bf[3], cf[3];
                tri_points(m, node->l, af, bf, cf);
                double a[3] = { af[0], af[1], af[2] }, bb[3] = { bf[0], bf[1], bf[2] },
                       cc[3] = { cf[0], cf[1], cf[2] };
                res += solid_angle_d(q, a, bb, cc);
            }
        }
        x->wn[qi] = res / (4.0 * 3.14159265358979323846);
    }
}
void or_fwn_vector_f64(const OrMesh* m, const float* pts, int64_t n, double beta, double* out,
                       int32_t nthreads)
{
    F64Ctx c; memset(&c, 0, sizeof(c));
    c.m = m; c.pts = pts; c.beta = beta; c.wn = out;
    or_parallel_for(n, nthreads, fwn_f64_body, &c, NULL);
}

/* ------------------------------------------------------------------ brute force */

typedef struct { const float* verts; const int32_t* tris; int32_t nf; const float* pts; double* out; } BfCtx;

static void bf_nearest_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    BfCtx* x = (BfCtx*)c;
    (void)cnt;
    for (int64_t qi = b; qi < e; ++qi) {
        double p[3] = { x->pts[3 * qi], x->pts[3 * qi + 1], x->pts[3 * qi + 2] };
        double best = INFINITY;
        for (int32_t f = 0; f < x->nf; ++f) {
            const int32_t* t = x->tris + 3 * (int64_t)f;
            if (t[0] < 0 || t[1] < 0 || t[2] < 0) continue;
            double a[3], bb[3], cc[3], pr[3], ba[2];
            for (int k = 0; k < 3; ++k) {
                a[k] = x->verts[3 * (int64_t)t[0] + k];
                bb[k] = x->verts[3 * (int64_t)t[1] + k];
                cc[k] = x->verts[3 * (int64_t)t[2] + k];
            }
            or_closest_point_in_triangle_d(p, a, bb, cc, pr, ba);
            double d = (pr[0] - p[0]) * (pr[0] - p[0]) + (pr[1] - p[1]) * (pr[1] - p[1])
                     + (pr[2] - p[2]) * (pr[2] - p[2]);
            if (d < best) best = d;
        }
        x->out[qi] = best;
    }
}
void or_bf_nearest_f64(const float* verts, const int32_t* tris, int32_t nf, const float* pts,
                       int64_t n, double* distSq, int32_t nthreads)
{
    BfCtx c = { verts, tris, nf, pts, distSq };
    or_parallel_for(n, nthreads, bf_nearest_body, &c, NULL);
}

static void bf_winding_body(void* c, int64_t b, int64_t e, OrCounters* cnt)
{
    BfCtx* x = (BfCtx*)c;
    (void)cnt;
    for (int64_t qi = b; qi < e; ++qi) {
        double p[3] = { x->pts[3 * qi], x->pts[3 * qi + 1], x->pts[3 * qi + 2] };
        double s = 0;
        for (int32_t f = 0; f < x->nf; ++f) {
            const int32_t* t = x->tris + 3 * (int64_t)f;
            if (t[0] < 0 || t[1] < 0 || t[2] < 0) continue;
            double a[3], bb[3], cc[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = x->verts[3 * (int64_t)t[0] + k];
                bb[k] = x->verts[3 * (int64_t)t[1] + k];
                cc[k] = x->verts[3 * (int64_t)t[2] + k];
            }
            s += solid_angle_d(p, a, bb, cc);
        }
        x->out[qi] = s / (4.0 * 3.14159265358979323846);
    }
}
void or_bf_winding_f64(const float* verts, const int32_t* tris, int32_t nf, const float* pts,
                       int64_t n, double* wn, int32_t nthreads)
{
    BfCtx c = { verts, tris, nf, pts, wn };
    or_parallel_for(n, nthreads, bf_winding_body, &c, NULL);
}
