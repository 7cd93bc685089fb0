// mi_mlp_capi.cpp -- C-ABI of the learned distance bound (include/meshlib_amd.h, mi_mlp_*):
// network state on the device, GPU training, inference, and the exact learned-bound forms of
// findProjections / calcFromGridWithDistances (prediction -> per-query cap -> capped search).
#include "meshlib_amd.h"

#include "mi_mlp.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

// shared with mi_capi.cpp
int mi_internal_fail(int code, const char* msg);

namespace {

#define MLP_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return mi_internal_fail(MI_E_HIP, (std::string("AMD GPU error: ") + hipGetErrorString(e_) + \
                                               " (" #expr ")").c_str());                                \
    } while (0)

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t ensure(size_t b)
    {
        if (bytes >= b) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// splitmix64 (seeded initialisation)
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

} // namespace

struct mi_mlp {
    int device = 0;
    Buf theta, mom, vel, grad, Wb, WbT, w5b;
    Buf trainPartial;
    Buf work;      // predictions / caps / grid points of the learned-bound forms
    int64_t trainRows = 0;
    int step = 0;
    mi::mlp::Params params{};
    std::mutex mu;
};

namespace {

hipError_t setupParams(mi_mlp* m, const float center[3], float scale)
{
    using namespace mi::mlp;
    hipError_t e;
    if ((e = m->theta.ensure(kParamsPadded * 4)) != hipSuccess) return e;
    if ((e = m->mom.ensure(kParamsPadded * 4)) != hipSuccess) return e;
    if ((e = m->vel.ensure(kParamsPadded * 4)) != hipSuccess) return e;
    if ((e = m->grad.ensure(kParamsPadded * 4)) != hipSuccess) return e;
    if ((e = m->Wb.ensure((size_t)kLayers * kHidden * kHidden * 2)) != hipSuccess) return e;
    if ((e = m->WbT.ensure((size_t)kLayers * kHidden * kHidden * 2)) != hipSuccess) return e;
    if ((e = m->w5b.ensure(kHidden * 2)) != hipSuccess) return e;
    Params& p = m->params;
    p.theta = m->theta.as<float>();
    p.thetaMut = m->theta.as<float>();
    p.Wb = m->Wb.as<uint16_t>();
    p.WbMut = m->Wb.as<uint16_t>();
    p.WbT = m->WbT.as<uint16_t>();
    p.WbTMut = m->WbT.as<uint16_t>();
    p.w5b = m->w5b.as<uint16_t>();
    p.w5bMut = m->w5b.as<uint16_t>();
    for (int i = 0; i < 3; ++i) p.center[i] = center[i];
    p.scale = scale;
    p.invScale = 1.f / scale;
    for (int k = 0; k < kBands; ++k) p.piFreq[k] = (float)(M_PI * std::pow(2.0, k / 3.0));
    p.slack = 0.f;
    return hipSuccess;
}

} // namespace

extern "C" {

int mi_mlp_num_params(void) { return mi::mlp::kParams; }

int mi_mlp_create(const float center[3], float scale, uint64_t seed, mi_mlp** out)
{
    using namespace mi::mlp;
    if (!out || !center || !(scale > 0)) return mi_internal_fail(MI_E_INVALID, "invalid network frame");
    *out = nullptr;
    mi_mlp* m = new (std::nothrow) mi_mlp();
    if (!m) return mi_internal_fail(MI_E_NOMEM, "out of host memory");
    if (hipGetDevice(&m->device) != hipSuccess || setupParams(m, center, scale) != hipSuccess) {
        delete m;
        return mi_internal_fail(MI_E_HIP, "AMD GPU error: cannot allocate the network");
    }
    // Xavier-uniform weights, zero biases; layer 0 only has kFeatures real inputs
    std::vector<float> th(kParamsPadded, 0.f);
    Rng rng{ seed };
    for (int l = 0; l < kLayers; ++l) {
        const int fanIn = l == 0 ? kFeatures : kHidden;
        const double a = std::sqrt(6.0 / (fanIn + kHidden));
        for (int o = 0; o < kHidden; ++o)
            for (int i = 0; i < fanIn; ++i)
                th[kOffW + (size_t)l * kHidden * kHidden + o * kHidden + i] = (float)((2 * rng.uniform() - 1) * a);
    }
    const double a5 = std::sqrt(6.0 / (kHidden + 1));
    for (int c = 0; c < kHidden; ++c) th[kOffW5 + c] = (float)((2 * rng.uniform() - 1) * a5);
    hipError_t e = hipMemcpy(m->theta.p, th.data(), kParamsPadded * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->mom.p, 0, kParamsPadded * 4);
    if (e == hipSuccess) e = hipMemset(m->vel.p, 0, kParamsPadded * 4);
    if (e == hipSuccess) e = mi::mlp::refreshCopies(m->params, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        delete m;
        return mi_internal_fail(MI_E_HIP, "AMD GPU error: cannot initialise the network");
    }
    *out = m;
    return MI_OK;
}

void mi_mlp_destroy(mi_mlp* m) { delete m; }

int mi_mlp_get_params(mi_mlp* m, float* host)
{
    if (!m || !host) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipMemcpy(host, m->theta.p, mi::mlp::kParams * 4, hipMemcpyDeviceToHost));
    return MI_OK;
}

int mi_mlp_set_params(mi_mlp* m, const float* host)
{
    if (!m || !host) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(hipMemcpy(m->theta.p, host, mi::mlp::kParams * 4, hipMemcpyHostToDevice));
    MLP_HIP(mi::mlp::refreshCopies(m->params, nullptr));
    MLP_HIP(hipDeviceSynchronize());
    m->step = 0;
    return MI_OK;
}

static int ensureTrain(mi_mlp* m, int64_t rows)
{
    using namespace mi::mlp;
    if (rows <= m->trainRows) return MI_OK;
    const int64_t tiles = (rows + kTileRows - 1) / kTileRows;
    MLP_HIP(m->trainPartial.ensure((size_t)tiles * kParamsPadded * 4));
    m->trainRows = rows;
    return MI_OK;
}

static mi::mlp::TrainBuffers trainBuffers(mi_mlp* m)
{
    return { m->trainPartial.as<float>() };
}

int mi_mlp_gradient_device(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, float* d_grad,
                           void* stream)
{
    if (!m || n <= 0 || !d_pts || !d_labels || !d_grad) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    if (int rc = ensureTrain(m, n)) return rc;
    MLP_HIP(mi::mlp::gradient(m->params, d_pts, d_labels, n, trainBuffers(m), d_grad,
                              static_cast<hipStream_t>(stream)));
    return MI_OK;
}

int mi_mlp_train_device(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, int32_t steps,
                        int32_t batch, float lr, void* stream)
{
    if (!m || n <= 0 || !d_pts || !d_labels || steps < 0 || batch <= 0 || batch > n)
        return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    if (int rc = ensureTrain(m, batch)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nBatches = n / batch;
    for (int32_t it = 0; it < steps; ++it) {
        const int64_t off = (int64_t)(it % nBatches) * batch;
        MLP_HIP(mi::mlp::trainStep(m->params, d_pts + 3 * off, d_labels + off, batch, trainBuffers(m),
                                   m->mom.as<float>(), m->vel.as<float>(), ++m->step, lr, s));
    }
    return MI_OK;
}

int mi_mlp_predict_device(mi_mlp* m, const float* d_pts, int64_t n, float* d_dist, void* stream)
{
    if (!m || n < 0 || (n > 0 && (!d_pts || !d_dist))) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(mi::mlp::forward(m->params, d_pts, n, d_dist, -1.f, static_cast<hipStream_t>(stream)));
    return MI_OK;
}

int mi_mlp_calibrate(mi_mlp* m, const float* d_pts, const float* d_labels, int64_t n, float quantile, float* slack,
                     void* stream)
{
    if (!m || n <= 0 || !d_pts || !d_labels || !(quantile >= 0.f && quantile <= 1.f))
        return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    if (n > (int64_t)1 << 26) return mi_internal_fail(MI_E_LIMIT, "more than 2^26 calibration samples");
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    MLP_HIP(m->work.ensure((size_t)n * 4));
    MLP_HIP(mi::mlp::forward(m->params, d_pts, n, m->work.as<float>(), -1.f, s));
    std::vector<float> pred((size_t)n), lab((size_t)n);
    MLP_HIP(hipMemcpyAsync(pred.data(), m->work.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    MLP_HIP(hipMemcpyAsync(lab.data(), d_labels, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    MLP_HIP(hipStreamSynchronize(s));
    // how far the true distance exceeds the prediction, at the requested quantile (never below 0)
    std::vector<float> under((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const float u = lab[(size_t)i] - pred[(size_t)i];
        under[(size_t)i] = u == u ? u : 0.f;
    }
    const size_t k = std::min((size_t)n - 1, (size_t)std::floor((double)quantile * (double)(n - 1) + 0.5));
    std::nth_element(under.begin(), under.begin() + (ptrdiff_t)k, under.end());
    m->params.slack = std::max(0.f, under[k]);
    if (slack) *slack = m->params.slack;
    return MI_OK;
}

int mi_mlp_set_slack(mi_mlp* m, float slack)
{
    if (!m || !(slack >= 0.f)) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    m->params.slack = slack;
    return MI_OK;
}

int mi_proj_find_projections_learned_device(mi_mesh* mesh, mi_mlp* m, float margin, const float* d_pts, int64_t n,
                                            float upDistLimitSq, mi_projection* d_res, int64_t* reruns,
                                            void* stream)
{
    if (!mesh || !m || n < 0 || !(margin >= 0) || (n > 0 && (!d_pts || !d_res)))
        return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    if (reruns) *reruns = 0;
    if (n == 0) return MI_OK;
    // the network's lock is held through the capped search, which reads the caps from its workspace
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    MLP_HIP(m->work.ensure((size_t)n * 4));
    MLP_HIP(mi::mlp::forward(m->params, d_pts, n, m->work.as<float>(), margin, static_cast<hipStream_t>(stream)));
    return mi_proj_find_projections_capped_device(mesh, d_pts, n, m->work.as<float>(), upDistLimitSq, d_res, reruns,
                                                  stream);
}

int mi_sdf_grid_learned_device(mi_mesh* mesh, mi_mlp* m, float margin, const int32_t dims[3], const float xf[12],
                               const mi_dist_options* opts, int32_t zBegin, int32_t zEnd, float* d_res,
                               int64_t* reruns, void* stream)
{
    if (!mesh || !m || !dims || !xf || !opts || !(margin >= 0)) return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    if (zBegin < 0 || zEnd > dims[2] || zBegin > zEnd || dims[0] < 0 || dims[1] < 0)
        return mi_internal_fail(MI_E_INVALID, "bad grid or z-layer range");
    if (reruns) *reruns = 0;
    const size_t nvox = (size_t)dims[0] * dims[1] * (size_t)(zEnd - zBegin);
    if (nvox == 0) return MI_OK;
    // the network's lock is held through the capped search, which reads the caps from its workspace
    std::lock_guard<std::mutex> lk(m->mu);
    MLP_HIP(hipSetDevice(m->device));
    const size_t capBytes = align256(nvox * 4);
    MLP_HIP(m->work.ensure(capBytes + (size_t)mi::mlp::gridCapBricks(dims, zBegin, zEnd) * 4));
    float* caps = m->work.as<float>();
    float* radius = reinterpret_cast<float*>(m->work.as<char>() + capBytes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    MLP_HIP(mi::mlp::gridCaps(m->params, dims, xf, zBegin, zEnd, margin, radius, caps, s));
    return mi_sdf_grid_capped_device(mesh, dims, xf, opts, m->work.as<float>(), zBegin, zEnd, d_res, reruns, stream);
}

int mi_batch_run_learned_device(mi_batch* b, const mi_batch_item* items, int32_t count, mi_mlp* const* nets,
                                float margin, const mi_dist_options* opts, float upDistLimitSq, void* stream,
                                int64_t* reruns)
{
    if (!b || count < 0 || (count > 0 && (!items || !nets)) || !opts || !(margin >= 0))
        return mi_internal_fail(MI_E_INVALID, "invalid arguments");
    if (reruns) *reruns = 0;
    if (count == 0) return MI_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // caps of every item (squared, from its own network), then the capped two-pass batch
    size_t need = 0, radiusBytes = 0;
    for (int32_t i = 0; i < count; ++i) {
        const mi_batch_item& it = items[i];
        if (it.dims[0] < 0 || it.dims[1] < 0 || it.dims[2] < 0 || it.numPts < 0)
            return mi_internal_fail(MI_E_INVALID, "invalid batch item");
        const size_t nvox = (size_t)it.dims[0] * it.dims[1] * it.dims[2];
        need += align256(nvox * 4) + align256((size_t)it.numPts * 4);
        radiusBytes += align256((size_t)mi::mlp::gridCapBricks(it.dims, 0, it.dims[2]) * 4);
    }
    Buf caps;
    MLP_HIP(caps.ensure(need + radiusBytes + 256));
    char* base = caps.as<char>();
    char* radiusBase = base + need;
    std::vector<const float*> capSdf((size_t)count, nullptr), capProj((size_t)count, nullptr);
    size_t off = 0;
    for (int32_t i = 0; i < count; ++i) {
        const mi_batch_item& it = items[i];
        const size_t nvox = (size_t)it.dims[0] * it.dims[1] * it.dims[2];
        float* cs = reinterpret_cast<float*>(base + off);
        off += align256(nvox * 4);
        float* cp = reinterpret_cast<float*>(base + off);
        off += align256((size_t)it.numPts * 4);
        mi_mlp* net = nets[i];
        if (!net) continue;
        std::lock_guard<std::mutex> lk(net->mu);
        MLP_HIP(hipSetDevice(net->device));
        float* radius = reinterpret_cast<float*>(radiusBase);
        radiusBase += align256((size_t)mi::mlp::gridCapBricks(it.dims, 0, it.dims[2]) * 4);
        if (nvox) {
            MLP_HIP(mi::mlp::gridCaps(net->params, it.dims, it.xf, 0, it.dims[2], margin, radius, cs, s));
            capSdf[i] = cs;
        }
        if (it.numPts) {
            MLP_HIP(mi::mlp::forward(net->params, it.d_pts, it.numPts, cp, margin, s));
            capProj[i] = cp;
        }
    }
    const int rc = mi_batch_run_capped_device(b, items, count, capSdf.data(), capProj.data(), opts, upDistLimitSq,
                                              stream, reruns);
    // the caps are read by the launches queued above: keep them until those have run
    (void)hipStreamSynchronize(s);
    return rc;
}

} // extern "C"
