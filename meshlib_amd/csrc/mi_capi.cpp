// mi_capi.cpp -- the C-ABI (include/meshlib_amd.h): mesh upload, chunked host-pointer
// entry points mirroring the reference's synchronous GPU classes, device-resident entry
// points, error strings and the memory budget.
#include "meshlib_amd.h"

#include "mi_kernels.hpp"
#include "mi_noalloc.hpp"
#include "mi_tree.hpp"
#include "mi_tree_gpu.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

static_assert(mi::kNoAllocMaxPoints == MI_NOALLOC_MAX_POINTS, "no-allocation capacity of the header");

namespace {

thread_local std::string g_err;
std::atomic<size_t> g_memLimit{ 0 };
std::atomic<int> g_treeBuilder{ MI_TREE_BUILDER_DEVICE };
std::atomic<int> g_splitBudget{ MI_SPLIT_AUTO };       // signed-distance walks
std::atomic<int> g_splitBudgetProj{ MI_SPLIT_AUTO };   // projection walks

// Automatic split budgets, from sweeps on MI355X (bumpy spheres of 20 k to 5.2 M triangles, tree
// depths 15 to 23).  Signed distance on trees of >= 18 levels: 3072 steps below 64 k waves, 6144
// below 1 M (k=8 256^3: 99 -> 94.5 ms), none from 1 M waves on (512^3: k=8 521 -> 477 ms, k=9
// 1031 -> 856 ms: enough waves to spread the long ones); shallower trees walk less, so the 64 subtree items of a
// hand-over weigh more and an earlier hand-over spreads the few long waves (512 steps below 16 k
// waves, 1024 above; 20 k triangles at 64^3: 2.7 -> 1.6 ms).  Projections: on trees of >= 20
// levels nearly every wave is long and 64 steps is best (flat to within noise up to 2048 at 2^22
// points); on shallower trees most waves finish alone given 1-2 k steps (2^22 points on 20 k /
// 82 k / 328 k triangles: 36.7 / 50.8 / 54.1 ms at 64 steps, 5.0 / 10.0 / 24.0 ms at 1024-2048).
int autoSdfBudget(int depth, int64_t waves)
{
    if (depth >= 18) return waves < 65536 ? 3072 : waves < (1 << 20) ? 6144 : 0;
    return waves < 16384 ? 512 : 1024;
}
int autoProjBudget(int depth)
{
    if (depth >= 20) return 64;
    return depth >= 17 ? 2048 : 1024;
}

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define MI_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(MI_E_HIP, std::string("AMD GPU error: ") + hipGetErrorString(e_) + " (" + \
                                      #expr + ")");                                               \
    } while (0)

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// getCudaSafeMemoryLimit (source/MRCuda/MRCudaBasic.cpp:75-92): 80 % of (free - 128 MiB),
// optionally capped by mi_set_memory_limit
size_t memBudget()
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    const size_t reserve = size_t(128) << 20;
    const size_t avail = fr > reserve ? fr - reserve : 0;
    size_t lim = (size_t)((float)avail * 0.80f);
    const size_t u = g_memLimit.load();
    if (u && u < lim) lim = u;
    return lim;
}

// RAII device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b)
    {
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (b == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Page-locked host staging buffers, kept for reuse across calls: pinning costs far more than the
// copy it speeds up (about 0.2 ms per MiB, against 0.02 ms per MiB for the copy). The pool keeps
// at most kPinnedKeep buffers; it lives until the process ends (never destroyed, so no HIP call
// runs from a static destructor).
constexpr size_t kPinnedKeep = 4;

struct PinnedPool {
    std::mutex mu;
    std::vector<std::pair<void*, size_t>> idle;

    // a buffer of at least `bytes`; *got = its allocated size
    hipError_t take(size_t bytes, void** p, size_t* got)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < idle.size(); ++i)
                if (idle[i].second >= bytes) {
                    *p = idle[i].first;
                    *got = idle[i].second;
                    idle.erase(idle.begin() + (ptrdiff_t)i);
                    return hipSuccess;
                }
        }
        *got = bytes;
        return hipHostMalloc(p, bytes, hipHostMallocDefault);
    }
    // `bytes`: the allocated size that take reported
    void give(void* p, size_t bytes)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu);
        idle.emplace_back(p, bytes);
        if (idle.size() > kPinnedKeep) {   // drop the smallest
            auto s = std::min_element(idle.begin(), idle.end(),
                                      [](const auto& a, const auto& b) { return a.second < b.second; });
            (void)hipHostFree(s->first);
            idle.erase(s);
        }
    }
};

PinnedPool& pinnedPool()
{
    static PinnedPool* pool = new PinnedPool;
    return *pool;
}

// RAII lease of a pooled pinned buffer
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { pinnedPool().give(p, bytes); }
    hipError_t take(size_t b)
    {
        hipError_t e = pinnedPool().take(b, &p, &bytes);
        if (e != hipSuccess) { p = nullptr; bytes = 0; }
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

bool report(mi_progress_fn cb, void* user, float v) { return !cb || cb(v, user) != 0; }

// ---------------------------------------------------------------- affine helpers (host)

struct M3 { float r[3][3]; };

float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// isRigid (source/MRMesh/MRMatrix3Decompose.h:27-39): Gram-Schmidt QR of the columns
// (MRMatrix3.h:272-295), rigid when |r_ii| are within 10 eps of 1
bool isRigid(const float* xf)
{
    float a0[3] = { xf[0], xf[3], xf[6] }, a1[3] = { xf[1], xf[4], xf[7] }, a2[3] = { xf[2], xf[5], xf[8] };
    auto len = [](const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    auto scaled = [](const float* v, float d, float* o) {
        const float inv = 1 / d;
        for (int i = 0; i < 3; ++i) o[i] = v[i] * inv;
    };
    const float r00 = len(a0);
    float e0[3] = { 0, 0, 0 };
    if (r00 > 0) scaled(a0, r00, e0);
    const float r01 = dot3(e0, a1), r02 = dot3(e0, a2);
    for (int i = 0; i < 3; ++i) a1[i] -= r01 * e0[i];
    const float r11 = len(a1);
    float e1[3] = { 0, 0, 0 };
    if (r11 > 0) scaled(a1, r11, e1);
    const float r12 = dot3(e1, a2);
    for (int i = 0; i < 3; ++i) a2[i] -= r02 * e0[i] + r12 * e1[i];
    const float r22 = len(a2);
    const float eps = 10 * FLT_EPSILON;
    return std::fabs(std::fabs(r00) - 1) <= eps && std::fabs(std::fabs(r11) - 1) <= eps &&
           std::fabs(std::fabs(r22) - 1) <= eps;
}

// AffineXf::inverse (source/MRMesh/MRAffineXf.h:51-57), Matrix3::inverse (MRMatrix3.h:236-247)
void xfInverse(const float* xf, float* out)
{
    const float* x = xf; const float* y = xf + 3; const float* z = xf + 6;
    const float det = x[0] * (y[1] * z[2] - y[2] * z[1]) - x[1] * (y[0] * z[2] - y[2] * z[0]) +
                      x[2] * (y[0] * z[1] - y[1] * z[0]);
    float A[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    if (det != 0) {
        const float adj[9] = { y[1] * z[2] - y[2] * z[1], x[2] * z[1] - x[1] * z[2], x[1] * y[2] - x[2] * y[1],
                               y[2] * z[0] - y[0] * z[2], x[0] * z[2] - x[2] * z[0], x[2] * y[0] - x[0] * y[2],
                               y[0] * z[1] - y[1] * z[0], x[1] * z[0] - x[0] * z[1], x[0] * y[1] - x[1] * y[0] };
        const float inv = 1 / det;
        for (int i = 0; i < 9; ++i) A[i] = inv * adj[i];
    }
    for (int i = 0; i < 9; ++i) out[i] = A[i];
    for (int i = 0; i < 3; ++i) out[9 + i] = -dot3(A + 3 * i, xf + 9);
}

// u * v = { u.A * v.A, u.A * v.b + u.b } (source/MRMesh/MRAffineXf.h:61-62, MRMatrix3.h:133-140)
void xfCompose(const float* u, const float* v, float* out)
{
    float r[12];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const float col[3] = { v[j], v[3 + j], v[6 + j] };
            r[3 * i + j] = dot3(u + 3 * i, col);
        }
    for (int i = 0; i < 3; ++i) r[9 + i] = dot3(u + 3 * i, v + 9) + u[9 + i];
    std::memcpy(out, r, sizeof(r));
}

mi::Xf toXf(const float* xf)
{
    mi::Xf x;
    std::memcpy(x.m, xf, sizeof(x.m));
    return x;
}

} // namespace

// error reporting shared with the other translation units of the library
int mi_internal_fail(int code, const char* msg) { return fail(code, msg); }

struct mi_mesh {
    int device = 0;
    int32_t nv = 0, nf = 0;       // nf: faces in the tree (the valid ones)
    int32_t faceSize = 0;         // entries of the caller's triangle array, deleted faces included
    int32_t numNodes = 0, depth = 0;
    mi::HostTree tree;     // host build: filled at creation; device build: downloaded on request
    DevBuf treeDev;        // device build: nodes + dipoles in the reference layout
    int stackD = 8;
    DevBuf arena;
    mi::MeshDev dev{};
    float sortLo[3] = { 0, 0, 0 }, sortHi[3] = { 0, 0, 0 };
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevBuf ws;   // sort workspace of the device-resident entry points (grow-only)
    // the mesh's own settings (MI_MODE_*), MI_MODE_DEFAULT = follow the process-wide setting
    std::atomic<int32_t> modes[MI_MODE_COUNT] = { MI_MODE_DEFAULT, MI_MODE_DEFAULT, MI_MODE_DEFAULT,
                                                  MI_MODE_DEFAULT };
    ~mi_mesh() { if (stream) (void)hipStreamDestroy(stream); }
};

namespace {

int32_t meshMode(const mi_mesh* m, int32_t mode, int32_t processDefault)
{
    const int32_t v = m->modes[mode].load();
    return v == MI_MODE_DEFAULT ? processDefault : v;
}

// the mesh's device view with this call's settings resolved
mi::MeshDev callView(const mi_mesh* m)
{
    mi::MeshDev d = m->dev;
    d.traversal = meshMode(m, MI_MODE_TRAVERSAL, static_cast<int32_t>(mi::traversal()));
    d.leafFilter = meshMode(m, MI_MODE_LEAF_FILTER, mi::leafFilter());
    return d;
}

int ensureWorkspace(mi_mesh* m, size_t bytes)
{
    if (m->ws.bytes >= bytes) return MI_OK;
    MI_HIP(m->ws.alloc(bytes));
    return MI_OK;
}

// bytes per lane of the signed-distance split form: the distance walk's float best, or the keyed walk's
// (order value, key) for minDistSq > 0
int sdfLaneBytes(const mi_dist_options* o) { return o && o->minDistSq > 0 ? 8 : 4; }

// device bytes per voxel of a signed-distance chunk: the output, and the split form's scratch
// (lane bytes + the cut-order byte + the unit list)
size_t sdfBytesPerVoxel(const mi_dist_options* o) { return o ? sizeof(float) + sdfLaneBytes(o) + 2 : sizeof(float); }

// waves of a signed-distance launch over `layers` z-layers and the split budget it runs with
int64_t sdfSplitPlan(const mi_mesh* m, const int32_t dims[3], int32_t layers, int32_t* budget)
{
    *budget = 0;
    if (layers <= 0) return 0;
    const int32_t d[3] = { dims[0], dims[1], layers };
    const int64_t waves = mi::sdfGridWaves(d, 0, layers);
    const int setting = meshMode(m, MI_MODE_SPLIT_SDF, g_splitBudget.load());
    *budget = setting >= 0 ? setting : autoSdfBudget(m->depth, waves);
    return waves;
}

// scratch of the split form for grids of up to `layers` z-layers, in `buf` (grown if needed)
int splitScratch(mi_mesh* m, DevBuf& buf, const int32_t dims[3], int32_t layers, int laneBytes, mi::SplitScratch& sc)
{
    sc = mi::SplitScratch{};
    int32_t budget = 0;
    const int64_t waves = sdfSplitPlan(m, dims, layers, &budget);
    if (budget <= 0) return MI_OK;
    const size_t bytes = mi::splitScratchBytes(waves, laneBytes);
    if (buf.bytes < bytes) MI_HIP(buf.alloc(bytes));
    mi::splitScratchBind(buf.p, waves, laneBytes, budget, sc);
    return MI_OK;
}

// split scratch of a projection launch over n points, placed at `base` (caller-sized with
// projSplitBytes); no split when the budget is 0
int projBudget(const mi_mesh* m)
{
    const int setting = meshMode(m, MI_MODE_SPLIT_PROJECTIONS, g_splitBudgetProj.load());
    return setting >= 0 ? setting : autoProjBudget(m->depth);
}

size_t projSplitBytes(const mi_mesh* m, int64_t n)
{
    return projBudget(m) > 0 ? mi::splitScratchBytes(mi::projectionWaves(n), 8) : 0;
}

void projSplit(const mi_mesh* m, void* base, int64_t n, mi::SplitScratch& sc)
{
    sc = mi::SplitScratch{};
    const int budget = projBudget(m);
    if (budget <= 0 || !base) return;
    mi::splitScratchBind(base, mi::projectionWaves(n), 8, budget, sc);
}

// device-resident entry points run on the caller's stream; NULL is the HIP null (default) stream
hipStream_t pick(mi_mesh*, void* s) { return static_cast<hipStream_t>(s); }

int checkGridArgs(const int32_t* dims, const float* xf)
{
    if (!dims || !xf) return fail(MI_E_INVALID, "dims and xf are required");
    if (dims[0] < 0 || dims[1] < 0 || dims[2] < 0) return fail(MI_E_INVALID, "negative grid dims");
    if ((int64_t)dims[0] * dims[1] > INT32_MAX) return fail(MI_E_LIMIT, "grid layer too large");
    return MI_OK;
}

} // namespace

extern "C" {

size_t mi_last_error(char* buf, size_t len)
{
    if (buf && len) {
        const size_t n = std::min(len - 1, g_err.size());
        std::memcpy(buf, g_err.data(), n);
        buf[n] = 0;
    }
    return g_err.size();
}

int mi_set_memory_limit(size_t bytes)
{
    g_memLimit.store(bytes);
    return MI_OK;
}

int mi_set_traversal(int mode)
{
    if (mode != MI_TRAVERSAL_PER_LANE && mode != MI_TRAVERSAL_WAVE_COOP)
        return fail(MI_E_INVALID, "unknown traversal mode");
    mi::setTraversal(static_cast<mi::Traversal>(mode));
    return MI_OK;
}

int mi_get_traversal(void) { return static_cast<int>(mi::traversal()); }

int mi_set_leaf_filter(int mode)
{
    if (mode != MI_LEAF_FILTER_OFF && mode != MI_LEAF_FILTER_ON && mode != MI_LEAF_FILTER_AUTO)
        return fail(MI_E_INVALID, "unknown leaf filter mode");
    mi::setLeafFilter(mode);
    return MI_OK;
}

int mi_get_leaf_filter(void) { return mi::leafFilter(); }

int mi_set_tree_builder(int builder)
{
    if (builder != MI_TREE_BUILDER_DEVICE && builder != MI_TREE_BUILDER_HOST)
        return fail(MI_E_INVALID, "unknown tree builder");
    g_treeBuilder.store(builder);
    return MI_OK;
}

int mi_get_tree_builder(void) { return g_treeBuilder.load(); }

int mi_set_split_budget(int32_t walk, int32_t steps)
{
    if (steps < MI_SPLIT_AUTO) return fail(MI_E_INVALID, "negative split budget");
    if (walk == MI_SPLIT_SDF) g_splitBudget.store(steps);
    else if (walk == MI_SPLIT_PROJECTIONS) g_splitBudgetProj.store(steps);
    else return fail(MI_E_INVALID, "unknown walk kind");
    return MI_OK;
}

int32_t mi_get_split_budget(int32_t walk)
{
    return walk == MI_SPLIT_PROJECTIONS ? g_splitBudgetProj.load() : walk == MI_SPLIT_SDF ? g_splitBudget.load() : -1;
}

} // extern "C"

namespace {

struct ArenaLayout {
    size_t dnode, tri, dint, dleaf, dquad, cut, prism;
    size_t total() const { return dnode + tri + dint + dleaf + dquad + cut + prism; }
};

ArenaLayout arenaLayout(int32_t nf)
{
    const size_t ni = nf > 0 ? (size_t)nf - 1 : 0, nl = nf > 0 ? (size_t)nf : 0;
    return { align256(ni * 64), align256(nl * 48), align256(ni * 32), align256(nl * 32), align256(ni * 128),
             nf > 0 ? align256(mi::kMaxCut * sizeof(mi::CutEntry)) + 256 : 0, align256(nl * 80) };
}

void bindArena(mi_mesh* m, const ArenaLayout& L)
{
    char* base = m->arena.as<char>();
    m->dev.dnode = reinterpret_cast<const float4*>(base);
    m->dev.tri = reinterpret_cast<const float4*>(base + L.dnode);
    m->dev.dint = reinterpret_cast<const float4*>(base + L.dnode + L.tri);
    m->dev.dleaf = reinterpret_cast<const float4*>(base + L.dnode + L.tri + L.dint);
    m->dev.dquad = reinterpret_cast<const float4*>(base + L.dnode + L.tri + L.dint + L.dleaf);
    m->dev.prism = reinterpret_cast<const float4*>(base + L.dnode + L.tri + L.dint + L.dleaf + L.dquad + L.cut);
}

// 4-wide records of the cooperative walks and the cut of the split form, from the binary records
// already in the arena (needs the root set)
int finishArena(mi_mesh* m, const ArenaLayout& L)
{
    if (m->nf < 1) return MI_OK;
    char* base = m->arena.as<char>();
    if (m->nf >= 2)
        MI_HIP(mi::buildQuads(m->dev.dnode, m->nf - 1,
                              reinterpret_cast<float4*>(base + L.dnode + L.tri + L.dint + L.dleaf), m->stream));
    MI_HIP(mi::buildPrisms(m->dev.tri, m->nf,
                           reinterpret_cast<float4*>(base + L.dnode + L.tri + L.dint + L.dleaf + L.dquad + L.cut),
                           m->stream));
    char* cutBase = base + L.dnode + L.tri + L.dint + L.dleaf + L.dquad;
    mi::CutEntry* cut = reinterpret_cast<mi::CutEntry*>(cutBase);
    int32_t* nCutDev = reinterpret_cast<int32_t*>(cutBase + align256(mi::kMaxCut * sizeof(mi::CutEntry)));
    MI_HIP(mi::buildCut(m->dev.dnode, m->dev.rootRef, m->dev.rootMin, m->dev.rootMax, cut, nCutDev, m->stream));
    int32_t nCut = 0;
    MI_HIP(hipMemcpyAsync(&nCut, nCutDev, 4, hipMemcpyDeviceToHost, m->stream));
    MI_HIP(hipStreamSynchronize(m->stream));
    if (nCut < 1 || nCut > mi::kMaxCut) return fail(MI_E_HIP, "AMD GPU error: bad tree cut");
    m->dev.cut = cut;
    m->dev.nCut = nCut;
    return MI_OK;
}

void setRoot(mi_mesh* m, const mi::Box3& rootBox, int32_t rootRef)
{
    for (int i = 0; i < 3; ++i) {
        m->dev.rootMin[i] = rootBox.min[i];
        m->dev.rootMax[i] = rootBox.max[i];
        const float c = 0.5f * (rootBox.min[i] + rootBox.max[i]);
        const float h = 0.75f * (rootBox.max[i] - rootBox.min[i]);
        m->sortLo[i] = c - h;
        m->sortHi[i] = c + h;
    }
    m->dev.rootRef = rootRef;
}

// The valid faces of a triangle array in getTriangulation order.  MeshTopology::getTriangulation
// (source/MRMesh/MRMeshTopology.cpp:430-441) returns faceSize() entries with {-1,-1,-1} for a deleted
// face, and the reference's tree covers the valid faces only, keeping their FaceIds as leaf ids
// (source/MRMesh/MRAABBTree.cpp:31-49).  A face with any negative index counts as deleted, as in the
// reference's GPU kernels (source/MRCuda/MRCudaFastWindingNumber.cu:168); an index >= nv is an error.
// `packed` / `ids` stay empty when every face is valid.
struct ValidFaces {
    std::vector<int32_t> packed, ids;
    int32_t count = 0;
};

int validFaces(const int32_t* tris, int32_t nf, int32_t nv, ValidFaces& out)
{
    out = ValidFaces{};
    int64_t bad = 0, deleted = 0;
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t* t = tris + 3 * (int64_t)f;
        if (t[0] < 0 || t[1] < 0 || t[2] < 0) ++deleted;
        else if (t[0] >= nv || t[1] >= nv || t[2] >= nv) ++bad;
    }
    if (bad) return fail(MI_E_INVALID, "triangle vertex index out of range");
    out.count = nf - (int32_t)deleted;
    if (!deleted) return MI_OK;
    try {
        out.packed.reserve(3 * (size_t)out.count);
        out.ids.reserve((size_t)out.count);
        for (int32_t f = 0; f < nf; ++f) {
            const int32_t* t = tris + 3 * (int64_t)f;
            if (t[0] < 0 || t[1] < 0 || t[2] < 0) continue;
            out.packed.insert(out.packed.end(), t, t + 3);
            out.ids.push_back(f);
        }
    } catch (const std::bad_alloc&) {
        return fail(MI_E_NOMEM, "out of host memory packing the valid faces");
    }
    return MI_OK;
}

// tree built by the host builder, flattened on the host and uploaded
int createHostBuilt(mi_mesh* m, const float* verts, const int32_t* tris, const ValidFaces& vf)
{
    try {
        mi::buildHostTree(verts, m->nv, vf.ids.empty() ? tris : vf.packed.data(), m->nf, m->tree);
    } catch (const std::bad_alloc&) {
        return fail(MI_E_NOMEM, "out of host memory building the tree");
    }
    // leaves name packed indices; the flattened records then read the caller's array by face id
    if (!vf.ids.empty())
        for (mi::TreeNode& n : m->tree.nodes)
            if (n.r < 0) n.l = vf.ids[(size_t)n.l];
    m->depth = m->tree.depth;
    m->stackD = mi::stackDepthFor(m->depth);
    if (m->stackD == 0) return fail(MI_E_LIMIT, "tree deeper than 32 levels");
    mi::FlatTree ft;
    mi::flattenTree(m->tree, verts, tris, ft);
    const ArenaLayout L = arenaLayout(m->nf);
    if (L.total()) {
        MI_HIP(m->arena.alloc(L.total()));
        bindArena(m, L);
        char* base = m->arena.as<char>();
        if (!ft.dnode.empty()) MI_HIP(hipMemcpy(base, ft.dnode.data(), ft.dnode.size() * 4, hipMemcpyHostToDevice));
        MI_HIP(hipMemcpy(base + L.dnode, ft.tri.data(), ft.tri.size() * 4, hipMemcpyHostToDevice));
        if (!ft.dint.empty())
            MI_HIP(hipMemcpy(base + L.dnode + L.tri, ft.dint.data(), ft.dint.size() * 4, hipMemcpyHostToDevice));
        MI_HIP(hipMemcpy(base + L.dnode + L.tri + L.dint, ft.dleaf.data(), ft.dleaf.size() * 4,
                         hipMemcpyHostToDevice));
    }
    if (m->nf > 0) {
        setRoot(m, ft.rootBox, ft.rootRef);
        if (int rc = finishArena(m, L)) return rc;
    }
    return MI_OK;
}

// tree, dipoles and the flattened layout built by the device builder from uploaded mesh arrays
int createDeviceBuilt(mi_mesh* m, const float* verts, const int32_t* tris, const ValidFaces& vf)
{
    if (m->nf == 0) return MI_OK;
    const size_t nn = 2 * (size_t)m->nf - 1;
    const ArenaLayout L = arenaLayout(m->nf);
    MI_HIP(m->arena.alloc(L.total()));
    bindArena(m, L);
    MI_HIP(m->treeDev.alloc(2 * align256(nn * sizeof(mi::TreeNode))));
    DevBuf dv, dt, scratch, ids;
    MI_HIP(dv.alloc(std::max<size_t>((size_t)m->nv * 12, 16)));
    MI_HIP(dt.alloc((size_t)m->nf * 12));
    MI_HIP(scratch.alloc(mi::deviceTreeScratchBytes(m->nf)));
    if (m->nv > 0) MI_HIP(hipMemcpyAsync(dv.p, verts, (size_t)m->nv * 12, hipMemcpyHostToDevice, m->stream));
    MI_HIP(hipMemcpyAsync(dt.p, vf.ids.empty() ? tris : vf.packed.data(), (size_t)m->nf * 12, hipMemcpyHostToDevice,
                          m->stream));
    if (!vf.ids.empty()) {
        MI_HIP(ids.alloc((size_t)m->nf * 4));
        MI_HIP(hipMemcpyAsync(ids.p, vf.ids.data(), (size_t)m->nf * 4, hipMemcpyHostToDevice, m->stream));
    }
    char* base = m->arena.as<char>();
    mi::DeviceTreeOut o;
    o.nodes = m->treeDev.as<mi::TreeNode>();
    o.dipoles = reinterpret_cast<mi::Dipole*>(m->treeDev.as<char>() + align256(nn * sizeof(mi::TreeNode)));
    o.dnode = reinterpret_cast<float4*>(base);
    o.tri = reinterpret_cast<float4*>(base + L.dnode);
    o.dint = reinterpret_cast<float4*>(base + L.dnode + L.tri);
    o.dleaf = reinterpret_cast<float4*>(base + L.dnode + L.tri + L.dint);
    mi::DeviceTreeInfo info;
    const hipError_t e = mi::buildDeviceTree(dv.as<float>(), m->nv, dt.as<int32_t>(), m->nf, scratch.p, scratch.bytes,
                                             o, info, m->stream, ids.as<int32_t>());
    if (e == hipErrorInvalidValue) return fail(MI_E_INVALID, "triangle vertex index out of range");
    MI_HIP(e);
    m->depth = info.depth;
    m->stackD = mi::stackDepthFor(m->depth);
    if (m->stackD == 0) return fail(MI_E_LIMIT, "tree deeper than 32 levels");
    setRoot(m, info.rootBox, info.rootRef);
    if (int rc = finishArena(m, L)) return rc;
    return MI_OK;
}

} // namespace

extern "C" {

int mi_mesh_create(const float* verts, int32_t nv, const int32_t* tris, int32_t nf, mi_mesh** out)
{
    if (!out) return fail(MI_E_INVALID, "out is null");
    *out = nullptr;
    if (nv < 0 || nf < 0 || (nv > 0 && !verts) || (nf > 0 && !tris))
        return fail(MI_E_INVALID, "invalid mesh arrays");
    if (nf > (1 << 30)) return fail(MI_E_LIMIT, "more than 2^30 faces");
    ValidFaces vf;
    if (int rc = validFaces(tris, nf, nv, vf)) return rc;
    mi_mesh* m = new (std::nothrow) mi_mesh();
    if (!m) return fail(MI_E_NOMEM, "out of host memory");
    m->nv = nv;
    m->nf = vf.count;
    m->faceSize = nf;
    m->numNodes = vf.count > 0 ? 2 * vf.count - 1 : 0;
    if (hipGetDevice(&m->device) != hipSuccess) { delete m; return fail(MI_E_HIP, "AMD GPU error: no device"); }
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return fail(MI_E_HIP, "AMD GPU error: stream creation failed");
    }
    const int rc = g_treeBuilder.load() == MI_TREE_BUILDER_HOST ? createHostBuilt(m, verts, tris, vf)
                                                                 : createDeviceBuilt(m, verts, tris, vf);
    if (rc != MI_OK) {
        delete m;
        return rc;
    }
    m->dev.numFaces = m->nf;
    m->dev.depth = m->depth;
    *out = m;
    return MI_OK;
}

void mi_mesh_destroy(mi_mesh* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    delete m;
}

int mi_mesh_num_nodes(const mi_mesh* m) { return m ? m->numNodes : 0; }

int mi_mesh_set_stats(mi_mesh* m, uint64_t* d_counters)
{
    if (!m) return fail(MI_E_INVALID, "mesh is null");
    static_assert(MI_STAT_STRIDE == mi::kStatStride && MI_STAT_LAUNCHES == mi::kStatLaunches, "counter layout");
    static_assert(MI_STAT_COUNT == mi::kStatCount, "counter layout");
    std::lock_guard<std::mutex> lk(m->mu);
    m->dev.stats = reinterpret_cast<unsigned long long*>(d_counters);
    return MI_OK;
}

int mi_mesh_depth(const mi_mesh* m) { return m ? m->depth : 0; }

int mi_mesh_set_mode(mi_mesh* m, int32_t mode, int32_t value)
{
    if (!m) return fail(MI_E_INVALID, "mesh is null");
    if (mode < 0 || mode >= MI_MODE_COUNT) return fail(MI_E_INVALID, "unknown mesh setting");
    bool ok = value == MI_MODE_DEFAULT;
    if (mode == MI_MODE_TRAVERSAL)
        ok = ok || value == MI_TRAVERSAL_PER_LANE || value == MI_TRAVERSAL_WAVE_COOP;
    else if (mode == MI_MODE_LEAF_FILTER)
        ok = ok || value == MI_LEAF_FILTER_OFF || value == MI_LEAF_FILTER_ON || value == MI_LEAF_FILTER_AUTO;
    else
        ok = ok || value >= MI_SPLIT_AUTO;
    if (!ok) return fail(MI_E_INVALID, "invalid value of a mesh setting");
    m->modes[mode].store(value);
    return MI_OK;
}

int32_t mi_mesh_get_mode(const mi_mesh* m, int32_t mode)
{
    if (!m || mode < 0 || mode >= MI_MODE_COUNT) return MI_MODE_DEFAULT;
    return m->modes[mode].load();
}

int mi_build_tree_host(const float* verts, int32_t nv, const int32_t* tris, int32_t nf, mi_tree_node* nodes,
                       mi_dipole* dipoles, int32_t* depth)
{
    if (nv < 0 || nf < 0 || (nv > 0 && !verts) || (nf > 0 && !tris))
        return fail(MI_E_INVALID, "invalid mesh arrays");
    ValidFaces vf;
    if (int rc = validFaces(tris, nf, nv, vf)) return rc;
    mi::HostTree t;
    try {
        mi::buildHostTree(verts, nv, vf.ids.empty() ? tris : vf.packed.data(), vf.count, t);
    } catch (const std::bad_alloc&) {
        return fail(MI_E_NOMEM, "out of host memory building the tree");
    }
    if (!vf.ids.empty())
        for (mi::TreeNode& n : t.nodes)
            if (n.r < 0) n.l = vf.ids[(size_t)n.l];
    if (nodes && !t.nodes.empty()) std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(mi_tree_node));
    if (dipoles && !t.dipoles.empty())
        std::memcpy(dipoles, t.dipoles.data(), t.dipoles.size() * sizeof(mi_dipole));
    if (depth) *depth = t.depth;
    return MI_OK;
}

int mi_mesh_get_tree(const mi_mesh* m, mi_tree_node* nodes, mi_dipole* dipoles)
{
    if (!m) return fail(MI_E_INVALID, "mesh is null");
    static_assert(sizeof(mi_tree_node) == sizeof(mi::TreeNode), "node layout");
    static_assert(sizeof(mi_dipole) == sizeof(mi::Dipole), "dipole layout");
    if (m->treeDev.p) {
        const size_t nb = (size_t)m->numNodes * sizeof(mi_tree_node);
        MI_HIP(hipSetDevice(m->device));
        if (nodes) MI_HIP(hipMemcpy(nodes, m->treeDev.p, nb, hipMemcpyDeviceToHost));
        if (dipoles) MI_HIP(hipMemcpy(dipoles, m->treeDev.as<char>() + align256(nb), nb, hipMemcpyDeviceToHost));
        return MI_OK;
    }
    if (nodes && !m->tree.nodes.empty())
        std::memcpy(nodes, m->tree.nodes.data(), m->tree.nodes.size() * sizeof(mi_tree_node));
    if (dipoles && !m->tree.dipoles.empty())
        std::memcpy(dipoles, m->tree.dipoles.data(), m->tree.dipoles.size() * sizeof(mi_dipole));
    return MI_OK;
}

// ---------------------------------------------------------------- device-resident entry points

int mi_sdf_grid_device(mi_mesh* m, const int32_t dims[3], const float xf[12], const mi_dist_options* opts,
                       int32_t zBegin, int32_t zEnd, float* d_res, void* stream)
{
    if (!m || !opts) return fail(MI_E_INVALID, "mesh and options are required");
    if (int rc = checkGridArgs(dims, xf)) return rc;
    if (zBegin < 0 || zEnd > dims[2] || zBegin > zEnd) return fail(MI_E_INVALID, "bad z-layer range");
    if (zEnd > zBegin && (int64_t)dims[0] * dims[1] > 0 && !d_res) return fail(MI_E_INVALID, "output is null");
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    mi::DistOpts o{ opts->minDistSq, opts->maxDistSq, opts->nullOutsideMinMax, opts->windingNumberThreshold,
                    opts->windingNumberBeta };
    mi::SplitScratch sc;
    if (int rc = splitScratch(m, m->ws, dims, zEnd - zBegin, sdfLaneBytes(opts), sc)) return rc;
    MI_HIP(mi::launchSdfGrid(callView(m), dims, toXf(xf), o, zBegin, zEnd, d_res, m->stackD, pick(m, stream), nullptr,
                             nullptr, nullptr, &sc));
    return MI_OK;
}

int mi_fwn_grid_device(mi_mesh* m, const int32_t dims[3], const float xf[12], float beta, int32_t zBegin,
                       int32_t zEnd, float* d_res, void* stream)
{
    if (!m) return fail(MI_E_INVALID, "mesh is null");
    if (int rc = checkGridArgs(dims, xf)) return rc;
    if (zBegin < 0 || zEnd > dims[2] || zBegin > zEnd) return fail(MI_E_INVALID, "bad z-layer range");
    if (zEnd > zBegin && (int64_t)dims[0] * dims[1] > 0 && !d_res) return fail(MI_E_INVALID, "output is null");
    MI_HIP(hipSetDevice(m->device));
    MI_HIP(mi::launchFwnGrid(callView(m), dims, toXf(xf), beta, zBegin, zEnd, d_res, m->stackD, pick(m, stream)));
    return MI_OK;
}

int mi_fwn_vector_device(mi_mesh* m, const float* d_pts, int64_t n, float beta, int32_t skipFace, float* d_res,
                         void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_res))) return fail(MI_E_INVALID, "invalid arguments");
    if (n > INT32_MAX) return fail(MI_E_LIMIT, "more than 2^31-1 points in one call");
    if (n == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    if (int rc = ensureWorkspace(m, mi::sortWorkspaceBytes(n))) return rc;
    const uint32_t* perm = nullptr;
    MI_HIP(mi::sortPointsMorton(d_pts, n, m->sortLo, m->sortHi, m->ws.p, m->ws.bytes, &perm, s));
    MI_HIP(mi::launchFwnPoints(callView(m), d_pts, perm, n, beta, skipFace, d_res, m->stackD, s));
    return MI_OK;
}

int mi_proj_find_projections_device(mi_mesh* m, const float* d_pts, int64_t n, const float* ptXf,
                                    const float* treeXf, float up, float lo, mi_projection* d_res, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_res))) return fail(MI_E_INVALID, "invalid arguments");
    if (n > INT32_MAX) return fail(MI_E_LIMIT, "more than 2^31-1 points in one call");
    if (n == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    const size_t sortBytes = align256(mi::sortWorkspaceBytes(n));
    if (int rc = ensureWorkspace(m, sortBytes + projSplitBytes(m, n))) return rc;
    mi::SplitScratch sc;
    projSplit(m, m->ws.as<char>() + sortBytes, n, sc);
    const uint32_t* perm = nullptr;
    MI_HIP(mi::sortPointsMorton(d_pts, n, m->sortLo, m->sortHi, m->ws.p, m->ws.bytes, &perm, s));
    mi::Xf px, tx;
    if (ptXf) px = toXf(ptXf);
    if (treeXf) tx = toXf(treeXf);
    MI_HIP(mi::launchProjections(callView(m), d_pts, perm, n, ptXf ? &px : nullptr, treeXf ? &tx : nullptr, up, lo,
                                 reinterpret_cast<mi::ProjRec*>(d_res), m->stackD, s, nullptr, nullptr, nullptr, &sc));
    return MI_OK;
}

int mi_proj_find_projections_noalloc_device(mi_mesh* m, const float* d_pts, int64_t n, const float* ptXf,
                                            const float* treeXf, float up, float lo, mi_projection* d_res,
                                            void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_res))) return fail(MI_E_INVALID, "invalid arguments");
    if (n > mi::kNoAllocMaxPoints)
        return fail(MI_E_LIMIT, "more than MI_NOALLOC_MAX_POINTS points in one call of the no-allocation entry point");
    if (n == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    void* scratch = nullptr;
    MI_HIP(mi::noAllocScratch(m->device, &scratch));
    // the same workspace layout as mi_proj_find_projections_device, in the fixed scratch
    const size_t sortBytes = align256(mi::sortWorkspaceBytes(n));
    if (sortBytes + projSplitBytes(m, n) > mi::kNoAllocScratchBytes)
        return fail(MI_E_LIMIT, "the no-allocation scratch is too small for this call");
    mi::SplitScratch sc;
    projSplit(m, static_cast<char*>(scratch) + sortBytes, n, sc);
    const uint32_t* perm = nullptr;
    MI_HIP(mi::sortPointsMorton(d_pts, n, m->sortLo, m->sortHi, scratch, sortBytes, &perm, s));
    mi::Xf px, tx;
    if (ptXf) px = toXf(ptXf);
    if (treeXf) tx = toXf(treeXf);
    MI_HIP(mi::launchProjections(callView(m), d_pts, perm, n, ptXf ? &px : nullptr, treeXf ? &tx : nullptr, up, lo,
                                 reinterpret_cast<mi::ProjRec*>(d_res), m->stackD, s, nullptr, nullptr, nullptr, &sc));
    return MI_OK;
}

int mi_proj_find_projections_bounded_device(mi_mesh* m, const float* d_pts, int64_t n, const float* d_upSq,
                                            const float* ptXf, const float* treeXf, float up, float lo,
                                            mi_projection* d_res, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_res || !d_upSq))) return fail(MI_E_INVALID, "invalid arguments");
    if (n > INT32_MAX) return fail(MI_E_LIMIT, "more than 2^31-1 points in one call");
    if (n == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    const size_t sortBytes = align256(mi::sortWorkspaceBytes(n));
    if (int rc = ensureWorkspace(m, sortBytes + projSplitBytes(m, n))) return rc;
    mi::SplitScratch sc;
    projSplit(m, m->ws.as<char>() + sortBytes, n, sc);
    const uint32_t* perm = nullptr;
    MI_HIP(mi::sortPointsMorton(d_pts, n, m->sortLo, m->sortHi, m->ws.p, m->ws.bytes, &perm, s));
    mi::Xf px, tx;
    if (ptXf) px = toXf(ptXf);
    if (treeXf) tx = toXf(treeXf);
    MI_HIP(mi::launchProjections(callView(m), d_pts, perm, n, ptXf ? &px : nullptr, treeXf ? &tx : nullptr, up, lo,
                                 reinterpret_cast<mi::ProjRec*>(d_res), m->stackD, s, d_upSq, nullptr, nullptr, &sc));
    return MI_OK;
}

int mi_proj_find_projections_capped_device(mi_mesh* m, const float* d_pts, int64_t n, const float* d_capSq,
                                           float up, mi_projection* d_res, int64_t* reruns, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_res || !d_capSq))) return fail(MI_E_INVALID, "invalid arguments");
    if (n > INT32_MAX) return fail(MI_E_LIMIT, "more than 2^31-1 points in one call");
    if (reruns) *reruns = 0;
    if (n == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    // workspace: sort, rerun list, its counter, split scratch.  The split scratch is bound even when the
    // step budget is 0: the waves with a failed cap hand their queries, uncapped, to the split pass, so no
    // query is searched again after it (a separate rerun launch over scattered queries ends on its
    // slowest single walk).  Only a tree too small to cut uses the rerun list.
    const size_t sortBytes = align256(mi::sortWorkspaceBytes(n));
    const size_t listBytes = align256((size_t)n * 4) + 256;
    const int64_t waves = mi::projectionWaves(n);
    const size_t splitBytes = mi::splitScratchBytes(waves, 8);
    if (int rc = ensureWorkspace(m, sortBytes + listBytes + splitBytes)) return rc;
    char* base = m->ws.as<char>();
    uint32_t* list = reinterpret_cast<uint32_t*>(base + sortBytes);
    uint32_t* count = reinterpret_cast<uint32_t*>(base + sortBytes + align256((size_t)n * 4));
    mi::SplitScratch sc;
    mi::splitScratchBind(base + sortBytes + listBytes, waves, 8, projBudget(m), sc);
    const uint32_t* perm = nullptr;
    MI_HIP(mi::sortPointsMorton(d_pts, n, m->sortLo, m->sortHi, m->ws.p, sortBytes, &perm, s));
    MI_HIP(mi::fillU32(count, 1, 0u, s));
    MI_HIP(mi::launchProjections(callView(m), d_pts, perm, n, nullptr, nullptr, up, 0.f,
                                 reinterpret_cast<mi::ProjRec*>(d_res), m->stackD, s, d_capSq, list, count, &sc));
    uint32_t h = 0;
    MI_HIP(hipMemcpyAsync(&h, count, 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (h && m->dev.nCut <= 0)
        MI_HIP(mi::launchProjections(callView(m), d_pts, list, (int64_t)h, nullptr, nullptr, up, 0.f,
                                     reinterpret_cast<mi::ProjRec*>(d_res), m->stackD, s));
    if (reruns) *reruns = h;
    return MI_OK;
}

int mi_sdf_grid_capped_device(mi_mesh* m, const int32_t dims[3], const float xf[12], const mi_dist_options* opts,
                              const float* d_capSq, int32_t zBegin, int32_t zEnd, float* d_res, int64_t* reruns,
                              void* stream)
{
    if (!m || !opts || !d_capSq) return fail(MI_E_INVALID, "mesh, options and caps are required");
    if (int rc = checkGridArgs(dims, xf)) return rc;
    if (zBegin < 0 || zEnd > dims[2] || zBegin > zEnd) return fail(MI_E_INVALID, "bad z-layer range");
    if (opts->minDistSq > 0) return fail(MI_E_INVALID, "capped signed distance needs minDistSq <= 0");
    if (reruns) *reruns = 0;
    const size_t nvox = (size_t)dims[0] * dims[1] * (size_t)(zEnd - zBegin);
    if (nvox == 0) return MI_OK;
    if (!d_res) return fail(MI_E_INVALID, "output is null");
    if (nvox > UINT32_MAX) return fail(MI_E_LIMIT, "slab larger than 2^32 voxels");
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    hipStream_t s = pick(m, stream);
    // workspace: rerun list, its counter, split scratch (bound even with a step budget of 0, as for the
    // capped projections: failed caps are searched again in the split pass; only a tree too small to cut
    // uses the rerun list)
    const size_t listBytes = align256(nvox * 4) + 256;
    int32_t budget = 0;
    const int64_t waves = sdfSplitPlan(m, dims, zEnd - zBegin, &budget);
    const size_t splitBytes = mi::splitScratchBytes(waves, sdfLaneBytes(opts));
    if (int rc = ensureWorkspace(m, listBytes + splitBytes)) return rc;
    uint32_t* list = m->ws.as<uint32_t>();
    uint32_t* count = reinterpret_cast<uint32_t*>(m->ws.as<char>() + align256(nvox * 4));
    mi::SplitScratch sc;
    mi::splitScratchBind(m->ws.as<char>() + listBytes, waves, sdfLaneBytes(opts), budget, sc);
    mi::DistOpts o{ opts->minDistSq, opts->maxDistSq, opts->nullOutsideMinMax, opts->windingNumberThreshold,
                    opts->windingNumberBeta };
    MI_HIP(mi::fillU32(count, 1, 0u, s));
    MI_HIP(mi::launchSdfGrid(callView(m), dims, toXf(xf), o, zBegin, zEnd, d_res, m->stackD, s, d_capSq, list, count,
                             &sc));
    uint32_t h = 0;
    MI_HIP(hipMemcpyAsync(&h, count, 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (h && m->dev.nCut <= 0)
        MI_HIP(mi::launchSdfVoxels(callView(m), dims, toXf(xf), o, zBegin, list, (int64_t)h, d_res, m->stackD, s));
    if (reruns) *reruns = h;
    return MI_OK;
}

// ---------------------------------------------------------------- host-pointer entry points

// z-layers per part of the host-pointer and by-parts pipelines: parts of about kPartBytes of
// output, so that computing one part overlaps copying (and consuming) the one before it
constexpr size_t kPartBytes = size_t(64) << 20;

static size_t partLayers(size_t layer, size_t nz)
{
    return std::min(std::max<size_t>(1, kPartBytes / (layer * sizeof(float))), nz);
}

// one grid launch over layers [z0, z1) into `out` (layer z0 at out[0])
static int gridLaunch(mi_mesh* m, const int32_t dims[3], const mi::Xf& x, const mi_dist_options* opts, float beta,
                      int32_t z0, int32_t z1, float* out, hipStream_t s, const mi::SplitScratch* sc)
{
    if (opts) {
        mi::DistOpts o{ opts->minDistSq, opts->maxDistSq, opts->nullOutsideMinMax, opts->windingNumberThreshold,
                        opts->windingNumberBeta };
        MI_HIP(mi::launchSdfGrid(callView(m), dims, x, o, z0, z1, out, m->stackD, s, nullptr, nullptr, nullptr, sc));
    } else {
        MI_HIP(mi::launchFwnGrid(callView(m), dims, x, beta, z0, z1, out, m->stackD, s));
    }
    return MI_OK;
}

// The two streams of a part pipeline (the mesh's stream and a second one) and its events; both
// streams drain before the second stream and the events are released.
struct PartStreams {
    hipStream_t st[2] = { nullptr, nullptr };
    std::vector<hipEvent_t> ev;
    ~PartStreams()
    {
        if (st[0]) (void)hipStreamSynchronize(st[0]);
        if (st[1]) { (void)hipStreamSynchronize(st[1]); (void)hipStreamDestroy(st[1]); }
        for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
    }
    hipError_t init(hipStream_t mainStream, size_t events = 0)
    {
        st[0] = mainStream;
        hipError_t e = hipStreamCreateWithFlags(&st[1], hipStreamNonBlocking);
        if (e != hipSuccess) { st[1] = nullptr; return e; }
        ev.assign(events, nullptr);
        for (hipEvent_t& x : ev)
            if ((e = hipEventCreateWithFlags(&x, hipEventDisableTiming)) != hipSuccess) return e;
        return hipSuccess;
    }
};

// The volume is computed in parts of whole z-layers, all queued on the mesh's stream into a ring of
// device slots (as many as the memory budget holds, maxBufferSizeAlignedByBlock,
// source/MRCuda/MRCudaBasic.cpp:103-107); a second stream copies each finished part to the caller's
// array while the later parts compute, and a slot takes its next part once its copy is done.
static int gridHost(mi_mesh* m, const int32_t dims[3], const float xf[12], const mi_dist_options* opts, float beta,
                    float* res, mi_progress_fn cb, void* user)
{
    if (!m) return fail(MI_E_INVALID, "mesh is null");
    if (int rc = checkGridArgs(dims, xf)) return rc;
    const size_t layer = (size_t)dims[0] * dims[1];
    const size_t total = layer * dims[2];
    if (total && !res) return fail(MI_E_INVALID, "output is null");
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    if (!report(cb, user, 0.5f)) return fail(MI_E_CANCELED, "Operation was canceled");
    if (total == 0) return MI_OK;
    const size_t nz = (size_t)dims[2];
    const size_t perVoxel = sdfBytesPerVoxel(opts);
    const size_t budgetLayers = std::min(memBudget() / perVoxel / layer, nz);
    if (budgetLayers == 0) return fail(MI_E_LIMIT, "one z-layer does not fit in the device memory budget");
    const size_t part = std::min(budgetLayers, partLayers(layer, nz));
    const size_t nParts = (nz + part - 1) / part;
    const size_t slots = std::min(budgetLayers / part, nParts);
    DevBuf out, splitBuf;
    MI_HIP(out.alloc(slots * part * layer * sizeof(float)));
    mi::SplitScratch sc;
    if (opts)
        if (int rc = splitScratch(m, splitBuf, dims, (int32_t)part, sdfLaneBytes(opts), sc)) return rc;
    PartStreams ps;   // st[0] computes, st[1] copies; events: computed / copied per slot
    MI_HIP(ps.init(m->stream, 2 * slots));
    hipEvent_t* const computed = ps.ev.data();
    hipEvent_t* const copied = ps.ev.data() + slots;
    if (!report(cb, user, 0.6f)) return fail(MI_E_CANCELED, "Operation was canceled");
    const mi::Xf x = toXf(xf);
    float* const ring = out.as<float>();
    auto launch = [&](size_t k) -> int {
        const size_t slot = k % slots;
        if (k >= slots) MI_HIP(hipStreamWaitEvent(ps.st[0], copied[slot], 0));   // the slot's last copy
        const int32_t z0 = (int32_t)(k * part), z1 = (int32_t)std::min(nz, (k + 1) * part);
        if (int rc = gridLaunch(m, dims, x, opts, beta, z0, z1, ring + slot * part * layer, ps.st[0], &sc)) return rc;
        MI_HIP(hipEventRecord(computed[slot], ps.st[0]));
        return MI_OK;
    };
    for (size_t k = 0; k < slots; ++k)
        if (int rc = launch(k)) return rc;
    for (size_t k = 0; k < nParts; ++k) {
        const size_t slot = k % slots;
        const size_t z0 = k * part, z1 = std::min(nz, (k + 1) * part);
        MI_HIP(hipStreamWaitEvent(ps.st[1], computed[slot], 0));
        MI_HIP(hipMemcpyAsync(res + z0 * layer, ring + slot * part * layer, (z1 - z0) * layer * sizeof(float),
                              hipMemcpyDeviceToHost, ps.st[1]));
        MI_HIP(hipEventRecord(copied[slot], ps.st[1]));
        if (k + slots < nParts)
            if (int rc = launch(k + slots)) return rc;
        MI_HIP(hipStreamSynchronize(ps.st[1]));
        if (!report(cb, user, 0.6f + 0.4f * (float)(k + 1) / (float)nParts))
            return fail(MI_E_CANCELED, "Operation was canceled");
    }
    return MI_OK;
}

int mi_fwn_calc_from_grid(mi_mesh* m, const int32_t dims[3], const float xf[12], float beta, float* res,
                          mi_progress_fn cb, void* user)
{
    return gridHost(m, dims, xf, nullptr, beta, res, cb, user);
}

int mi_fwn_calc_from_grid_with_distances(mi_mesh* m, const int32_t dims[3], const float xf[12],
                                         const mi_dist_options* opts, float* res, mi_progress_fn cb, void* user)
{
    if (!opts) return fail(MI_E_INVALID, "options are required");
    return gridHost(m, dims, xf, opts, 0.f, res, cb, user);
}

// by-parts: the GPU computes parts k+1 (and k+2) while the caller's callback consumes part k
// (cudaPipeline, source/MRCuda/MRCudaPipeline.hpp:9-48, on two HIP streams). Parts are about
// kPartBytes of output (at least 4 overlaps long, at most half the memory budget); their boundaries
// follow splitByChunks (source/MRMesh/MRChunkIterator.cpp:29-53) in units of z-layers.
static int gridByParts(mi_mesh* m, mi_grid_part_fn fn, void* fnUser, const int32_t dims[3], const float xf[12],
                       const mi_dist_options* opts, float beta, int32_t layerOverlap, mi_progress_fn cb, void* user)
{
    if (!m || !fn) return fail(MI_E_INVALID, "mesh and callback are required");
    if (int rc = checkGridArgs(dims, xf)) return rc;
    if (layerOverlap < 0) return fail(MI_E_INVALID, "negative layer overlap");
    const size_t layer = (size_t)dims[0] * dims[1];
    const size_t total = layer * dims[2];
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    if (!report(cb, user, 0.5f)) return fail(MI_E_CANCELED, "Operation was canceled");
    if (total == 0) return MI_OK;
    const size_t nz = (size_t)dims[2];
    // two device buffers (+ split scratch) of `chunkLayers` whole z-layers within the budget
    const size_t perVoxel = sdfBytesPerVoxel(opts);
    const size_t budgetLayers = std::min(memBudget() / perVoxel / layer / 2, nz);
    if (budgetLayers == 0) return fail(MI_E_LIMIT, "one z-layer does not fit in the device memory budget");
    const size_t chunkLayers =
        std::min(budgetLayers, std::max(partLayers(layer, nz), std::min(nz, 4 * (size_t)layerOverlap)));
    size_t step = chunkLayers, nChunks = 1;
    if (nz > chunkLayers) {
        if ((size_t)layerOverlap >= chunkLayers)
            return fail(MI_E_LIMIT, "layer overlap does not fit in the device memory budget");
        step = chunkLayers - (size_t)layerOverlap;
        nChunks = (nz - (size_t)layerOverlap + step - 1) / step;
    }
    const size_t slots = std::min<size_t>(2, nChunks);
    DevBuf dbuf[2], splitBuf[2];
    mi::SplitScratch sc[2];
    PinnedBuf hbuf[2];
    for (size_t i = 0; i < slots; ++i) {
        MI_HIP(dbuf[i].alloc(chunkLayers * layer * sizeof(float)));
        MI_HIP(hbuf[i].take(chunkLayers * layer * sizeof(float)));
        if (opts)
            if (int rc = splitScratch(m, splitBuf[i], dims, (int32_t)chunkLayers, sdfLaneBytes(opts), sc[i])) return rc;
    }
    PartStreams ps;
    MI_HIP(ps.init(m->stream));
    if (!report(cb, user, 0.6f)) return fail(MI_E_CANCELED, "Operation was canceled");
    const mi::Xf x = toXf(xf);
    auto chunkRange = [&](size_t c, int32_t& z0, int32_t& z1) {
        z0 = (int32_t)(c * step);
        z1 = (int32_t)std::min(nz, (size_t)z0 + chunkLayers);
    };
    // part c into slot c % 2: compute, then copy to the slot's page-locked buffer, on the slot's stream
    auto launch = [&](size_t c) -> int {
        const size_t b = c % slots;
        int32_t z0, z1;
        chunkRange(c, z0, z1);
        if (int rc = gridLaunch(m, dims, x, opts, beta, z0, z1, dbuf[b].as<float>(), ps.st[b], &sc[b])) return rc;
        MI_HIP(hipMemcpyAsync(hbuf[b].p, dbuf[b].p, (size_t)(z1 - z0) * layer * sizeof(float),
                              hipMemcpyDeviceToHost, ps.st[b]));
        return MI_OK;
    };
    for (size_t c = 0; c < slots; ++c)
        if (int rc = launch(c)) return rc;
    for (size_t c = 0; c < nChunks; ++c) {
        const size_t b = c % slots;
        MI_HIP(hipStreamSynchronize(ps.st[b]));
        int32_t z0, z1;
        chunkRange(c, z0, z1);
        const int32_t pd[3] = { dims[0], dims[1], z1 - z0 };
        const int rc = fn(hbuf[b].as<float>(), pd, z0, fnUser);
        if (rc != MI_OK) return fail(rc, "grid part callback failed");
        if (c + slots < nChunks)   // the slot's buffers are free again
            if (int rc2 = launch(c + slots)) return rc2;
        if (!report(cb, user, 0.6f + 0.4f * (float)(c + 1) / (float)nChunks))
            return fail(MI_E_CANCELED, "Operation was canceled");
    }
    return MI_OK;
}

int mi_fwn_calc_from_grid_by_parts(mi_mesh* m, mi_grid_part_fn fn, void* fnUser, const int32_t dims[3],
                                   const float xf[12], float beta, int32_t layerOverlap, mi_progress_fn cb,
                                   void* user)
{
    return gridByParts(m, fn, fnUser, dims, xf, nullptr, beta, layerOverlap, cb, user);
}

int mi_fwn_calc_from_grid_with_distances_by_parts(mi_mesh* m, mi_grid_part_fn fn, void* fnUser,
                                                  const int32_t dims[3], const float xf[12],
                                                  const mi_dist_options* opts, int32_t layerOverlap,
                                                  mi_progress_fn cb, void* user)
{
    if (!opts) return fail(MI_E_INVALID, "options are required");
    return gridByParts(m, fn, fnUser, dims, xf, opts, 0.f, layerOverlap, cb, user);
}

} // extern "C"

// Per-point pipeline shared by calcFromVector and findProjections: points go in parts of up to
// kPartPoints (within the memory budget, maxBufferSize, source/MRCuda/MRCudaBasic.cpp:94-97, with
// the sort workspace counted), two in flight on two streams; each part is uploaded, sorted, walked
// and downloaded on its slot's stream while the other slot's part runs.
constexpr size_t kPartPoints = size_t(1) << 21;

template <class LaunchFn>
static int pointsHost(mi_mesh* m, const float* pts, int64_t n, size_t outElem, void* res, LaunchFn&& launchFn,
                      mi_progress_fn cb, void* user, bool split = false)
{
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    if (!report(cb, user, 0.5f)) return fail(MI_E_CANCELED, "Operation was canceled");
    if (n == 0) return MI_OK;
    // device bytes per point: coordinates, result, sort workspace (+ 9 B of the projection walk's
    // split scratch)
    const size_t perElem = 12 + outElem + 20 + (split ? 9 : 0);
    const size_t budget = memBudget() / perElem;
    if (budget == 0) return fail(MI_E_LIMIT, "no point fits in the device memory budget");
    const size_t part = std::min({ kPartPoints, (size_t)n, std::max<size_t>(1, budget / 2) });
    const size_t nParts = ((size_t)n + part - 1) / part;
    const size_t slots = std::min({ (size_t)2, budget / part, nParts });
    struct Slot {
        DevBuf pts, res, ws, split;
        mi::SplitScratch sc;
    } sl[2];
    for (size_t i = 0; i < slots; ++i) {
        MI_HIP(sl[i].pts.alloc(part * 12));
        MI_HIP(sl[i].res.alloc(part * outElem));
        MI_HIP(sl[i].ws.alloc(mi::sortWorkspaceBytes((int64_t)part)));
        if (split && projSplitBytes(m, (int64_t)part)) {
            MI_HIP(sl[i].split.alloc(projSplitBytes(m, (int64_t)part)));
            projSplit(m, sl[i].split.p, (int64_t)part, sl[i].sc);
        }
    }
    PartStreams ps;
    MI_HIP(ps.init(m->stream));
    if (!report(cb, user, 0.6f)) return fail(MI_E_CANCELED, "Operation was canceled");
    auto range = [&](size_t k, size_t& off, size_t& sz) {
        off = k * part;
        sz = std::min(part, (size_t)n - off);
    };
    auto start = [&](size_t k) -> int {
        Slot& s = sl[k % slots];
        const hipStream_t st = ps.st[k % slots];
        size_t off, sz;
        range(k, off, sz);
        MI_HIP(hipMemcpyAsync(s.pts.p, pts + 3 * off, sz * 12, hipMemcpyHostToDevice, st));
        float* const dp = static_cast<float*>(s.pts.p);
        const uint32_t* perm = nullptr;
        MI_HIP(mi::sortPointsMorton(dp, (int64_t)sz, m->sortLo, m->sortHi, s.ws.p, s.ws.bytes, &perm, st));
        MI_HIP(launchFn(dp, perm, (int64_t)sz, s.res.p, &s.sc, st));
        return MI_OK;
    };
    for (size_t k = 0; k < slots; ++k)
        if (int rc = start(k)) return rc;
    for (size_t k = 0; k < nParts; ++k) {
        size_t off, sz;
        range(k, off, sz);
        const hipStream_t st = ps.st[k % slots];
        MI_HIP(hipMemcpyAsync(static_cast<char*>(res) + off * outElem, sl[k % slots].res.p, sz * outElem,
                              hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
        if (k + slots < nParts)
            if (int rc = start(k + slots)) return rc;
        if (!report(cb, user, 0.6f + 0.4f * (float)(k + 1) / (float)nParts))
            return fail(MI_E_CANCELED, "Operation was canceled");
    }
    return MI_OK;
}

extern "C" {

int mi_fwn_calc_from_vector(mi_mesh* m, const float* pts, int64_t n, float beta, int32_t skipFace, float* res,
                            mi_progress_fn cb, void* user)
{
    if (!m || n < 0 || (n > 0 && (!pts || !res))) return fail(MI_E_INVALID, "invalid arguments");
    return pointsHost(m, pts, n, sizeof(float), res,
                      [&](const float* dp, const uint32_t* perm, int64_t sz, void* dr, const mi::SplitScratch*,
                          hipStream_t st) {
                          return mi::launchFwnPoints(callView(m), dp, perm, sz, beta, skipFace, static_cast<float*>(dr),
                                                     m->stackD, st);
                      },
                      cb, user);
}

int mi_fwn_calc_self_intersections(mi_mesh* m, float beta, uint8_t* res, mi_progress_fn cb, void* user)
{
    if (!m || (m->faceSize > 0 && !res)) return fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MI_HIP(hipSetDevice(m->device));
    if (!report(cb, user, 0.5f)) return fail(MI_E_CANCELED, "Operation was canceled");
    if (m->faceSize == 0) return MI_OK;
    // one flag per entry of the caller's triangle array; deleted faces stay 0 (the reference sets bits
    // of valid faces only, source/MRMesh/MRFastWindingNumber.cpp:40-52)
    const int words = (int)(((size_t)m->faceSize + 3) / 4);
    DevBuf out;
    MI_HIP(out.alloc((size_t)words * 4));
    MI_HIP(mi::fillU32(out.as<uint32_t>(), words, 0u, m->stream));
    MI_HIP(mi::launchSelfIntersections(callView(m), beta, out.as<uint8_t>(), m->stackD, m->stream));
    MI_HIP(hipMemcpyAsync(res, out.p, (size_t)m->faceSize, hipMemcpyDeviceToHost, m->stream));
    MI_HIP(hipStreamSynchronize(m->stream));
    if (!report(cb, user, 1.0f)) return fail(MI_E_CANCELED, "Operation was canceled");
    return MI_OK;
}

int mi_proj_find_projections(mi_mesh* m, const float* pts, int64_t n, const float* objXf, const float* refObjXf,
                             float up, float lo, mi_projection* res)
{
    if (!m || n < 0 || (n > 0 && (!pts || !res))) return fail(MI_E_INVALID, "invalid arguments");
    // createProjectionTransforms (source/MRMesh/MRMeshProject.cpp:105-121)
    const float* treeXf = (refObjXf && !isRigid(refObjXf)) ? refObjXf : nullptr;
    float folded[12];
    const float* ptXf = nullptr;
    if (treeXf || !refObjXf) {
        ptXf = objXf;
    } else {
        xfInverse(refObjXf, folded);
        if (objXf) xfCompose(folded, objXf, folded);
        ptXf = folded;
    }
    mi::Xf px, tx;
    if (ptXf) px = toXf(ptXf);
    if (treeXf) tx = toXf(treeXf);
    return pointsHost(m, pts, n, sizeof(mi_projection), res,
                      [&](const float* dp, const uint32_t* perm, int64_t sz, void* dr, const mi::SplitScratch* sc,
                          hipStream_t st) {
                          return mi::launchProjections(callView(m), dp, perm, sz, ptXf ? &px : nullptr,
                                                       treeXf ? &tx : nullptr, up, lo,
                                                       static_cast<mi::ProjRec*>(dr), m->stackD, st,
                                                       nullptr, nullptr, nullptr, sc);
                      },
                      nullptr, nullptr, true);
}

size_t mi_proj_heap_bytes(size_t numProjections)
{
    return numProjections * (12 + sizeof(mi_projection) + 20 + 9);
}

} // extern "C"

// ---------------------------------------------------------------- batched persistent launches

struct mi_batch {
    int device = 0;
    int numCUs = 0;
    int blocksPerCU = 4;
    int xcdSplit = 1;
    DevBuf items;        // BatchItem[]
    DevBuf tickets;      // 8 x u64
    DevBuf sortWs;       // per-item Morton sort workspaces
    DevBuf rerunWs;      // capped form: per-item counters of failed caps
    std::vector<mi::BatchItem> host;     // launch order (see prepareBatch)
    std::vector<int32_t> slotOf;          // caller's item index -> position in `host`
    std::mutex mu;
};

namespace {

int uploadItems(mi_batch* b, int32_t count, hipStream_t s)
{
    if ((size_t)count * sizeof(mi::BatchItem) > b->items.bytes)
        MI_HIP(b->items.alloc((size_t)count * sizeof(mi::BatchItem)));
    // pageable source: the copy is complete when hipMemcpyAsync returns
    MI_HIP(hipMemcpyAsync(b->items.p, b->host.data(), (size_t)count * sizeof(mi::BatchItem), hipMemcpyHostToDevice, s));
    return MI_OK;
}

// builds the device descriptors (sorting every item's points) and returns the unit count
int prepareBatch(mi_batch* b, const mi_batch_item* items, int32_t count, hipStream_t s, int64_t* total, int* stackD,
                 bool upload = true)
{
    *total = 0;
    *stackD = 8;
    b->host.assign((size_t)count, mi::BatchItem{});
    size_t wsTotal = 0;
    for (int32_t i = 0; i < count; ++i) {
        const mi_batch_item& it = items[i];
        if (!it.mesh) return fail(MI_E_INVALID, "batch item without a mesh");
        if (it.numPts < 0 || it.numPts > INT32_MAX || (it.numPts > 0 && (!it.d_pts || !it.d_proj)))
            return fail(MI_E_INVALID, "invalid batch item points");
        const int64_t nvox = (int64_t)it.dims[0] * it.dims[1] * it.dims[2];
        if (it.dims[0] < 0 || it.dims[1] < 0 || it.dims[2] < 0 || (nvox > 0 && !it.d_sdf))
            return fail(MI_E_INVALID, "invalid batch item grid");
        if (it.mesh->device != b->device) return fail(MI_E_INVALID, "batch items must live on the batch's device");
        wsTotal += align256(mi::sortWorkspaceBytes(it.numPts));
        *stackD = std::max(*stackD, it.mesh->stackD);
    }
    if (wsTotal && b->sortWs.bytes < wsTotal) MI_HIP(b->sortWs.alloc(wsTotal));
    // launch order: items on deeper trees first (their work units are the long ones), larger first
    // among equals, so that the persistent launch ends on short units instead of waiting for a
    // long one taken last
    std::vector<int32_t> order((size_t)count);
    for (int32_t i = 0; i < count; ++i) order[(size_t)i] = i;
    auto units = [&](int32_t i) {
        const mi_batch_item& it = items[i];
        return (int64_t)it.dims[0] * it.dims[1] * it.dims[2] / 64 + it.numPts / 64;
    };
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
        if (items[a].mesh->depth != items[c].mesh->depth) return items[a].mesh->depth > items[c].mesh->depth;
        return units(a) > units(c);
    });
    b->slotOf.assign((size_t)count, 0);
    for (int32_t k = 0; k < count; ++k) b->slotOf[(size_t)order[(size_t)k]] = k;
    size_t off = 0;
    int64_t unit = 0;
    for (int32_t k = 0; k < count; ++k) {
        const int32_t i = order[(size_t)k];
        const mi_batch_item& it = items[i];
        mi::BatchItem& d = b->host[(size_t)k];
        d.m = it.mesh->dev;
        d.dx = it.dims[0]; d.dy = it.dims[1]; d.dz = it.dims[2];
        d.nbx = (d.dx + 3) / 4; d.nby = (d.dy + 3) / 4; d.nbz = (d.dz + 3) / 4;
        std::memcpy(d.xf.m, it.xf, sizeof(d.xf.m));
        d.sdf = it.d_sdf;
        d.gridUnits = (int64_t)d.dx * d.dy * d.dz > 0 ? (int64_t)d.nbx * d.nby * d.nbz : 0;
        d.pts = it.d_pts;
        d.npts = it.numPts;
        d.proj = reinterpret_cast<mi::ProjRec*>(it.d_proj);
        d.projUnits = (it.numPts + 63) / 64;
        d.perm = nullptr;
        if (it.numPts > 0) {
            const size_t wsb = align256(mi::sortWorkspaceBytes(it.numPts));
            const uint32_t* perm = nullptr;
            MI_HIP(mi::sortPointsMorton(it.d_pts, it.numPts, it.mesh->sortLo, it.mesh->sortHi,
                                        b->sortWs.as<char>() + off, wsb, &perm, s));
            d.perm = perm;
            off += wsb;
        }
        d.unitBegin = unit;
        unit += d.gridUnits + d.projUnits;
    }
    *total = unit;
    return upload ? uploadItems(b, count, s) : MI_OK;
}

int launchPrepared(mi_batch* b, int32_t count, int64_t total, int stackD, const mi_dist_options* opts, float up,
                   int blocksPerCU, int xcdSplit, hipStream_t s, bool capped = false)
{
    mi::DistOpts o{ opts->minDistSq, opts->maxDistSq, opts->nullOutsideMinMax, opts->windingNumberThreshold,
                    opts->windingNumberBeta };
    MI_HIP(mi::fillU32(b->tickets.as<uint32_t>(), 16, 0u, s));
    MI_HIP(mi::launchBatch(b->items.as<mi::BatchItem>(), count, total, o, up,
                           b->tickets.as<unsigned long long>(), xcdSplit, b->numCUs * blocksPerCU, stackD, capped, s));
    return MI_OK;
}

int checkBatch(mi_batch* b, const mi_batch_item* items, int32_t count, const mi_dist_options* opts)
{
    if (!b || count < 0 || (count > 0 && !items) || !opts) return fail(MI_E_INVALID, "invalid batch arguments");
    if (opts->minDistSq > 0) return fail(MI_E_INVALID, "batched signed distance needs minDistSq <= 0");
    return MI_OK;
}

} // namespace

extern "C" {

int mi_batch_create(mi_batch** out)
{
    if (!out) return fail(MI_E_INVALID, "out is null");
    *out = nullptr;
    mi_batch* b = new (std::nothrow) mi_batch();
    if (!b) return fail(MI_E_NOMEM, "out of host memory");
    hipDeviceProp_t prop;
    if (hipGetDevice(&b->device) != hipSuccess || hipGetDeviceProperties(&prop, b->device) != hipSuccess ||
        b->tickets.alloc(16 * sizeof(uint32_t)) != hipSuccess) {
        delete b;
        return fail(MI_E_HIP, "AMD GPU error: cannot set up the batch");
    }
    b->numCUs = prop.multiProcessorCount;
    *out = b;
    return MI_OK;
}

void mi_batch_destroy(mi_batch* b) { delete b; }

int mi_batch_set_config(mi_batch* b, int32_t blocksPerCU, int32_t xcdSplit)
{
    if (!b || blocksPerCU < 1 || blocksPerCU > 16) return fail(MI_E_INVALID, "invalid batch config");
    b->blocksPerCU = blocksPerCU;
    b->xcdSplit = xcdSplit ? 1 : 0;
    return MI_OK;
}

int mi_batch_get_config(mi_batch* b, int32_t* blocksPerCU, int32_t* xcdSplit)
{
    if (!b) return fail(MI_E_INVALID, "batch is null");
    if (blocksPerCU) *blocksPerCU = b->blocksPerCU;
    if (xcdSplit) *xcdSplit = b->xcdSplit;
    return MI_OK;
}

int mi_batch_run_device(mi_batch* b, const mi_batch_item* items, int32_t count, const mi_dist_options* opts,
                        float upDistLimitSq, void* stream)
{
    if (int rc = checkBatch(b, items, count, opts)) return rc;
    std::lock_guard<std::mutex> lk(b->mu);
    MI_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t total = 0;
    int stackD = 8;
    if (int rc = prepareBatch(b, items, count, s, &total, &stackD)) return rc;
    return launchPrepared(b, count, total, stackD, opts, upDistLimitSq, b->blocksPerCU, b->xcdSplit, s);
}

int mi_batch_run_capped_device(mi_batch* b, const mi_batch_item* items, int32_t count, const float* const* d_capSdf,
                               const float* const* d_capProj, const mi_dist_options* opts, float upDistLimitSq,
                               void* stream, int64_t* reruns)
{
    if (int rc = checkBatch(b, items, count, opts)) return rc;
    if (reruns) *reruns = 0;
    if (count == 0) return MI_OK;
    std::lock_guard<std::mutex> lk(b->mu);
    MI_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t total = 0;
    int stackD = 8;
    if (int rc = prepareBatch(b, items, count, s, &total, &stackD, false)) return rc;
    // two counters per item: the voxels and points whose cap failed (each searched again, uncapped, by its
    // own wave inside the launch)
    const size_t cntBytes = align256((size_t)count * 8);
    if (b->rerunWs.bytes < cntBytes) MI_HIP(b->rerunWs.alloc(cntBytes));
    uint32_t* counts = b->rerunWs.as<uint32_t>();
    for (int32_t i = 0; i < count; ++i) {
        const int32_t k = b->slotOf[(size_t)i];
        mi::BatchItem& d = b->host[(size_t)k];
        d.capSdf = d_capSdf ? d_capSdf[i] : nullptr;
        d.capProj = d_capProj ? d_capProj[i] : nullptr;
        d.rerunCount = counts + 2 * k;
    }
    MI_HIP(mi::fillU32(counts, 2 * count, 0u, s));
    if (int rc = uploadItems(b, count, s)) return rc;
    if (int rc = launchPrepared(b, count, total, stackD, opts, upDistLimitSq, b->blocksPerCU, b->xcdSplit, s, true))
        return rc;
    std::vector<uint32_t> hc((size_t)count * 2);
    MI_HIP(hipMemcpyAsync(hc.data(), counts, hc.size() * 4, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    int64_t again = 0;
    for (uint32_t c : hc) again += c;
    if (reruns) *reruns = again;
    // the items' descriptors go back to the plain form for later calls
    for (int32_t k = 0; k < count; ++k) {
        mi::BatchItem& d = b->host[(size_t)k];
        d.capSdf = d.capProj = nullptr;
        d.rerunCount = nullptr;
    }
    return MI_OK;
}

int mi_batch_autotune(mi_batch* b, const mi_batch_item* items, int32_t count, const mi_dist_options* opts,
                      float upDistLimitSq, void* stream, float* bestMs)
{
    if (int rc = checkBatch(b, items, count, opts)) return rc;
    std::lock_guard<std::mutex> lk(b->mu);
    MI_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t total = 0;
    int stackD = 8;
    if (int rc = prepareBatch(b, items, count, s, &total, &stackD)) return rc;
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    float best = -1.f;
    int bestBlocks = b->blocksPerCU, bestSplit = b->xcdSplit;
    const int blocksCand[] = { 2, 3, 4, 5, 6 };
    int rc = MI_OK;
    // warm-up (caches, code object load)
    rc = launchPrepared(b, count, total, stackD, opts, upDistLimitSq, b->blocksPerCU, b->xcdSplit, s);
    for (int split = 0; split < 2 && rc == MI_OK; ++split) {
        for (int bpc : blocksCand) {
            if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(MI_E_HIP, "AMD GPU error: event"); break; }
            rc = launchPrepared(b, count, total, stackD, opts, upDistLimitSq, bpc, split, s);
            if (rc != MI_OK) break;
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(MI_E_HIP, "AMD GPU error: timing the batch");
                break;
            }
            if (best < 0 || ms < best) { best = ms; bestBlocks = bpc; bestSplit = split; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != MI_OK) return rc;
    b->blocksPerCU = bestBlocks;
    b->xcdSplit = bestSplit;
    if (bestMs) *bestMs = best;
    return MI_OK;
}

} // extern "C"
