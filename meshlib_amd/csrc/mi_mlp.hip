// mi_mlp.hip -- the learned distance bound: a Fourier-feature MLP trained and evaluated on gfx950
// with hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Network: x^ = (p - centre) / scale; features [x^ (3), sin(pi f_k x^_d), cos(pi f_k x^_d)] for
// 16 bands f_k = 2^(k/3) (99 dims, zero-padded to 128) -> 4 x (Linear 128 + SiLU) -> Linear 1 ->
// softplus = distance / scale.  Parameters live in one FP32 vector (master copy, Adam state) with
// BF16 copies W (out-major) and W^T (in-major) that the MFMA kernels read.
//
// Tile shape: a workgroup of 4 waves owns 64 rows; wave w computes rows [16w, 16w+16) of every
// layer as 8 output tiles of 16x16, K = 128 in 4 MFMA steps of 32.  Activations stay in LDS between
// layers; each layer's weights are staged once per workgroup into LDS (rows padded to 136 BF16 so
// the 16-byte fragment reads of 16 rows fall in distinct banks).
//
// MFMA 16x16x32 BF16 fragments (lane l): A row l%16, K elements 8*(l/16)..+7; B column l%16, the
// same K elements; accumulator element j sits at row 4*(l/16)+j, column l%16.
#include "mi_mlp.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace mi {
namespace mlp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int LDP = kHidden + 8;   // padded LDS row (BF16 elements) for 128-wide rows

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);     // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 frag(const uint16_t* p)
{
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }
__device__ __forceinline__ float siluf(float z) { return z / (1.f + expf(-z)); }
__device__ __forceinline__ float softplusf(float u) { return u > 20.f ? u : log1pf(expf(u)); }

// feature `i` (0..127) of the normalised point x
__device__ __forceinline__ float feature(int i, const float x[3], const float* piFreq)
{
    if (i < 3) return x[i];
    if (i >= kFeatures) return 0.f;
    const int j = i - 3, k = j / 6, r = j % 6;
    const float a = piFreq[k] * x[r % 3];
    return r < 3 ? sinf(a) : cosf(a);
}

__device__ __forceinline__ void normalised(const Params& p, const float* pt, float x[3])
{
    x[0] = (pt[0] - p.center[0]) * p.invScale;
    x[1] = (pt[1] - p.center[1]) * p.invScale;
    x[2] = (pt[2] - p.center[2]) * p.invScale;
}

// stages layer l's BF16 weights (128 x 128, rows of `src`) into LDS rows padded to LDP
__device__ __forceinline__ void stageWeights(const uint16_t* __restrict__ src, uint16_t* Ws)
{
    const int tid = threadIdx.x;
    for (int it = 0; it < 8; ++it) {          // 128 x 128 BF16 = 2048 chunks of 16 B
        const int c = it * 256 + tid, row = c >> 4, c8 = (c & 15) * 8;
        *reinterpret_cast<uint4*>(&Ws[row * LDP + c8]) = *reinterpret_cast<const uint4*>(&src[row * kHidden + c8]);
    }
}

// 8 BF16 of column `col` in rows r .. r + 7 of a row-major LDS array (a K-major MFMA fragment)
__device__ __forceinline__ bf16x8 colFrag(const uint16_t* A, int r, int col)
{
    uint16_t h[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) h[e] = A[(r + e) * LDP + col];
    const uint4 v = make_uint4((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                               (uint32_t)h[4] | ((uint32_t)h[5] << 16), (uint32_t)h[6] | ((uint32_t)h[7] << 16));
    return __builtin_bit_cast(bf16x8, v);
}

// One training step's forward and backward pass over a 64-row tile, in one launch.  loss =
// 1/(2B) sum (pred/scale - label/scale)^2 over the batch; writes the tile's partial gradient.
// The BF16 activations H_{-1} (features) .. H_3 stay in LDS; each layer's pre-activations Z are computed
// again in the backward pass from the same MFMA inputs in the same order (so the same values) instead
// of being stored; dH and dZ stay in registers in the accumulator layout (wave w: rows 16w + 4(l/16) + j,
// column 16 nt + l%16), dZ also in LDS as BF16 for the two products dW = dZ^T H and dH = dZ W.
// Rounding points as the oracle: features, weights, activations and dZ (for the products) in BF16,
// FP32 accumulation, the bias gradients from FP32 dZ.
__global__ __launch_bounds__(256) void kTrainTile(Params p, const float* __restrict__ pts,
                                                  const float* __restrict__ labels, int64_t n, float invB,
                                                  float* __restrict__ partial)
{
    __shared__ __attribute__((aligned(16))) uint16_t Act[kLayers + 1][kTileRows * LDP];
    __shared__ __attribute__((aligned(16))) uint16_t Ws[kHidden * LDP];
    __shared__ __attribute__((aligned(16))) uint16_t dZr[kTileRows * LDP];
    __shared__ float duS[kTileRows];
    __shared__ float red[4][kHidden];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int r0 = 16 * w;
    const int64_t row0 = (int64_t)blockIdx.x * kTileRows;
    float* part = partial + (size_t)blockIdx.x * kParamsPadded;

    {   // features (rows past n: the features of the origin; their loss weight is 0)
        const int r = tid >> 2, q4 = tid & 3;
        const int64_t gi = row0 + r;
        float x[3] = { 0.f, 0.f, 0.f };
        if (gi < n) normalised(p, pts + 3 * gi, x);
        for (int j = 0; j < 32; ++j) {
            const int fi = 32 * q4 + j;
            Act[0][r * LDP + fi] = f2bf(feature(fi, x, p.piFreq));
        }
    }
    // forward: H_l = BF16(SiLU(H_{l-1} W_l^T + b_l))
    for (int l = 0; l < kLayers; ++l) {
        __syncthreads();
        stageWeights(p.Wb + (size_t)l * kHidden * kHidden, Ws);
        __syncthreads();
        bf16x8 a[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) a[kc] = frag(&Act[l][(r0 + lr) * LDP + kc * 32 + lk * 8]);
        const float* bias = p.theta + kOffB + l * kHidden;
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            f32x4 acc = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) acc = mfma(a[kc], frag(&Ws[(nt * 16 + lr) * LDP + kc * 32 + lk * 8]), acc);
            const int col = nt * 16 + lr;
            const float bc = bias[col];
#pragma unroll
            for (int j = 0; j < 4; ++j) Act[l + 1][(r0 + 4 * lk + j) * LDP + col] = f2bf(siluf(acc[j] + bc));
        }
    }
    __syncthreads();
    {   // output layer u = sum_c h3[c] w5[c] + b5 and du = dloss/du (four partial sums of 32 per row)
        const int r = tid >> 2, q4 = tid & 3;
        float sum = 0.f;
        for (int j = 0; j < 32; ++j) {
            const int c = 32 * q4 + j;
            sum += bf2f(Act[kLayers][r * LDP + c]) * bf2f(p.w5b[c]);
        }
        sum += __shfl_xor(sum, 1);
        sum += __shfl_xor(sum, 2);
        if (q4 == 0) {
            const int64_t gi = row0 + r;
            float du = 0.f;
            if (gi < n) {
                const float u = sum + p.theta[kOffB5];
                du = (softplusf(u) - labels[gi] * p.invScale) * invB * sigmoidf(u);
            }
            duS[r] = du;
        }
    }
    __syncthreads();
    {   // dw5[c] = sum_r du[r] h3[r][c], db5 = sum_r du[r]
        const int c = tid & 127, half = tid >> 7;
        float sum = 0.f;
        for (int rr = 0; rr < 32; ++rr) {
            const int r = half * 32 + rr;
            sum += duS[r] * bf2f(Act[kLayers][r * LDP + c]);
        }
        red[half][c] = sum;
    }
    __syncthreads();
    if (tid < kHidden) part[kOffW5 + tid] = red[0][tid] + red[1][tid];
    if (tid == 0) {
        float sum = 0.f;
        for (int r = 0; r < kTileRows; ++r) sum += duS[r];
        part[kOffB5] = sum;
    }
    // dH_3 = du w5^T, in registers
    float dH[8][4];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) dH[nt][j] = duS[r0 + 4 * lk + j] * bf2f(p.w5b[nt * 16 + lr]);
    for (int l = kLayers - 1; l >= 0; --l) {
        __syncthreads();      // the previous layer's readers of Ws, dZr and red are done
        stageWeights(p.Wb + (size_t)l * kHidden * kHidden, Ws);
        __syncthreads();
        // Z_l again, then dZ = dH * SiLU'(Z): FP32 for the bias gradient, BF16 in dZr for the products
        bf16x8 a[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) a[kc] = frag(&Act[l][(r0 + lr) * LDP + kc * 32 + lk * 8]);
        const float* bias = p.theta + kOffB + l * kHidden;
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            f32x4 acc = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) acc = mfma(a[kc], frag(&Ws[(nt * 16 + lr) * LDP + kc * 32 + lk * 8]), acc);
            const int col = nt * 16 + lr;
            const float bc = bias[col];
            float colSum = 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float z = acc[j] + bc;
                const float sg = sigmoidf(z);
                const float dz = dH[nt][j] * (sg * (1.f + z * (1.f - sg)));
                dZr[(r0 + 4 * lk + j) * LDP + col] = f2bf(dz);
                colSum += dz;
            }
            // column sums over the wave's 16 rows (lanes l, l^16, l^32, l^48 hold the same column)
            colSum += __shfl_xor(colSum, 16);
            colSum += __shfl_xor(colSum, 32);
            if (lk == 0) red[w][col] = colSum;
        }
        __syncthreads();
        if (tid < kHidden) part[kOffB + l * kHidden + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
        {   // dW_l (128 out x 128 in) = dZ^T (out x rows) . H_{l-1} (rows x in); wave w: out rows [32w, 32w+32)
            f32x4 acc[2][8];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 8; ++nt) acc[mt][nt] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                const bf16x8 a0 = colFrag(dZr, 32 * kc + 8 * lk, 32 * w + lr);
                const bf16x8 a1 = colFrag(dZr, 32 * kc + 8 * lk, 32 * w + 16 + lr);
#pragma unroll
                for (int nt = 0; nt < 8; ++nt) {
                    const bf16x8 b = colFrag(Act[l], 32 * kc + 8 * lk, 16 * nt + lr);
                    acc[0][nt] = mfma(a0, b, acc[0][nt]);
                    acc[1][nt] = mfma(a1, b, acc[1][nt]);
                }
            }
            float* dW = part + kOffW + l * kHidden * kHidden;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 8; ++nt)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        dW[(32 * w + 16 * mt + 4 * lk + j) * kHidden + 16 * nt + lr] = acc[mt][nt][j];
        }
        if (l == 0) break;
        // dH_{l-1} (rows x in) = dZ (rows x out) . W_l (out x in); B(k = out, n = in) = W^T[in][out].  Every
        // wave finished reading Ws (W_l) before the barrier above.
        stageWeights(p.WbT + (size_t)l * kHidden * kHidden, Ws);
        __syncthreads();
        bf16x8 d[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) d[kc] = frag(&dZr[(r0 + lr) * LDP + kc * 32 + lk * 8]);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            f32x4 acc = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) acc = mfma(d[kc], frag(&Ws[(nt * 16 + lr) * LDP + kc * 32 + lk * 8]), acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) dH[nt][j] = acc[j];
        }
    }
}

// ---------------------------------------------------------------- inference

// Inference kernel: persistent workgroups of 8 waves hold all four hidden layers' BF16 weights in LDS
// (staged once per workgroup, in MFMA fragment order: one 1-KB block per (layer, 16-row tile, 32-wide K
// chunk), lane l's 16 bytes at 16 l), and each wave carries 64 points (4 groups of 16) through the
// layers with the activations in registers.  The layers are computed transposed, Z^T = W . H^T: the B
// operand is the activations (lane l: point l%16 of a group, 8 K elements of a chunk), and an output
// tile leaves lane l holding 4 features 16 mt + 4(l/16) + j of point l%16.  Two output tiles 2c, 2c+1 are
// one K chunk of the next layer with its K elements in the order (32c + 4g + 0..3, 32c + 16 + 4g + 0..3)
// for lane group g = l/16; the staged weights of layers 1..3 are permuted to that order, so no
// activation leaves its lane.  Layer 0's chunks take the features in natural order (lane group g:
// features 32c + 8g + 0..7), computed in-lane with the hardware sine / cosine (v_sin_f32, v_cos_f32 take
// revolutions: sin(pi f x) = v_sin(f x / 2); within 2e-4 of sinf / cosf, far below the features' BF16
// rounding).  Rounding points as in kForward: features, weights and hidden activations in BF16, FP32
// accumulation, the output layer on BF16 activations and weights.
constexpr int kPredWaves = 8;
constexpr int kPredGroups = 4;                        // 16-point groups per wave
constexpr int kPredPoints = 16 * kPredGroups;         // points per wave and tile
constexpr int kFragBlocks = kLayers * 8 * 4;          // (layer, tile, chunk) fragment blocks

struct Xf3 { float m[9]; };   // linear part of a grid transform (row-major)

// query source of the inference: free points, or the centres of the 4x4x4 bricks of a grid slab (brick
// index (bz nby + by) nbx + bx, centre at grid coordinates 4 b + 1.5, z counted from zBegin)
struct PredSource {
    const float* pts;      // null: brick centres
    int nbx, nby, zBegin;
    float xf[12];
    __device__ __forceinline__ void point(int64_t i, float p[3]) const
    {
        if (pts) {
            p[0] = pts[3 * i]; p[1] = pts[3 * i + 1]; p[2] = pts[3 * i + 2];
            return;
        }
        const int64_t layer = (int64_t)nbx * nby;
        const float z = (float)zBegin + 4.f * (float)(int)(i / layer) + 1.5f;
        const int rem = (int)(i % layer);
        const float x = 4.f * (float)(rem % nbx) + 1.5f, y = 4.f * (float)(rem / nbx) + 1.5f;
        p[0] = (xf[0] * x + xf[1] * y + xf[2] * z) + xf[9];
        p[1] = (xf[3] * x + xf[4] * y + xf[5] * z) + xf[10];
        p[2] = (xf[6] * x + xf[7] * y + xf[8] * z) + xf[11];
    }
};

__device__ __forceinline__ float siluFast(float z) { return z * __builtin_amdgcn_rcpf(1.f + __expf(-z)); }

// feature f of the normalised point with the hardware sine / cosine; halfFreq[k] = f_k / 2
__device__ __forceinline__ float featureFast(int f, const float x[3], const float* halfFreq)
{
    if (f < 3) return x[f];
    if (f >= kFeatures) return 0.f;
    const int j = f - 3, k = j / 6, r = j % 6;
    const float a = halfFreq[k] * x[r < 3 ? r : r - 3];
    return r < 3 ? __builtin_amdgcn_sinf(a) : __builtin_amdgcn_cosf(a);
}

__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

// out[i]: the predicted distance (capMargin < 0), else the cap radius r = pred (1 + capMargin) + slack, squared
// when capSquared
__global__ __launch_bounds__(64 * kPredWaves) void kPredict(Params p, PredSource src, int64_t n,
                                                             float* __restrict__ out, float capMargin, int capSquared)
{
    __shared__ __attribute__((aligned(16))) uint4 Wf[kFragBlocks * 64];
    __shared__ float bias[kLayers * kHidden];
    __shared__ float w5[kHidden];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    // stage the weights in fragment order (layers 1.. with the permuted K order)
    for (int e = tid; e < kFragBlocks * 64; e += 64 * kPredWaves) {
        const int ln = e & 63, blk = e >> 6;
        const int kc = blk & 3, mt = (blk >> 2) & 7, l = blk >> 5;
        const int o = 16 * mt + (ln & 15), g = ln >> 4;
        const uint16_t* row = p.Wb + ((size_t)l * kHidden + o) * kHidden;
        uint4 v;
        if (l == 0) {
            v = *reinterpret_cast<const uint4*>(row + 32 * kc + 8 * g);
        } else {
            const uint2 a = *reinterpret_cast<const uint2*>(row + 32 * kc + 4 * g);
            const uint2 b = *reinterpret_cast<const uint2*>(row + 32 * kc + 16 + 4 * g);
            v = make_uint4(a.x, a.y, b.x, b.y);
        }
        Wf[e] = v;
    }
    for (int e = tid; e < kLayers * kHidden; e += 64 * kPredWaves) bias[e] = p.theta[kOffB + e];
    for (int e = tid; e < kHidden; e += 64 * kPredWaves) w5[e] = bf2f(p.w5b[e]);
    __syncthreads();
    float halfFreq[kBands];
#pragma unroll
    for (int k = 0; k < kBands; ++k) halfFreq[k] = p.piFreq[k] * (float)(0.5 / M_PI);
    const float b5 = p.theta[kOffB5];
    const int64_t tiles = (n + kPredPoints - 1) / kPredPoints;
    for (int64_t t = (int64_t)blockIdx.x * kPredWaves + wv; t < tiles; t += (int64_t)gridDim.x * kPredWaves) {
        const int64_t base = t * kPredPoints;
        // layer 0 inputs: this lane's 32 features of its point in each group
        bf16x8 cur[kPredGroups][4];
#pragma unroll
        for (int g = 0; g < kPredGroups; ++g) {
            const int64_t i = base + 16 * g + lr;
            float x[3] = { 0.f, 0.f, 0.f };
            if (i < n) {
                float pt[3];
                src.point(i, pt);
                normalised(p, pt, x);
            }
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                uint16_t h[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) h[e] = f2bf(featureFast(32 * kc + 8 * lg + e, x, halfFreq));
                const uint4 v = make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7]));
                cur[g][kc] = __builtin_bit_cast(bf16x8, v);
            }
        }
        for (int l = 0; l < kLayers; ++l) {
            bf16x8 nxt[kPredGroups][4];
            const uint4* Wl = Wf + (size_t)l * 32 * 64;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                f32x4 acc[kPredGroups][2];
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int mt = 2 * c + half;
#pragma unroll
                    for (int g = 0; g < kPredGroups; ++g) acc[g][half] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
                    for (int kc = 0; kc < 4; ++kc) {
                        const bf16x8 a = __builtin_bit_cast(bf16x8, Wl[(mt * 4 + kc) * 64 + lane]);
#pragma unroll
                        for (int g = 0; g < kPredGroups; ++g) acc[g][half] = mfma(a, cur[g][kc], acc[g][half]);
                    }
                }
                // bias + SiLU, BF16: tile 2c gives K slots 0..3 of chunk c, tile 2c+1 slots 4..7
                float bb[2][4];
#pragma unroll
                for (int half = 0; half < 2; ++half)
#pragma unroll
                    for (int j = 0; j < 4; ++j) bb[half][j] = bias[l * kHidden + 16 * (2 * c + half) + 4 * lg + j];
#pragma unroll
                for (int g = 0; g < kPredGroups; ++g) {
                    uint16_t h[8];
#pragma unroll
                    for (int half = 0; half < 2; ++half)
#pragma unroll
                        for (int j = 0; j < 4; ++j) h[4 * half + j] = f2bf(siluFast(acc[g][half][j] + bb[half][j]));
                    const uint4 v = make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]),
                                               pack2(h[6], h[7]));
                    nxt[g][c] = __builtin_bit_cast(bf16x8, v);
                }
            }
#pragma unroll
            for (int g = 0; g < kPredGroups; ++g)
#pragma unroll
                for (int c = 0; c < 4; ++c) cur[g][c] = nxt[g][c];
        }
        // output layer: this lane's 32 features (chunk c, slot 4 half + j = feature 32c + 16 half + 4lg + j),
        // then the sum over the four lanes of the point
#pragma unroll
        for (int g = 0; g < kPredGroups; ++g) {
            float s = 0.f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint4 v = __builtin_bit_cast(uint4, cur[g][c]);
                const uint32_t wds[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint16_t hb = (uint16_t)(wds[e >> 1] >> (16 * (e & 1)));
                    const int f = 32 * c + 16 * (e >> 2) + 4 * lg + (e & 3);
                    s += bf2f(hb) * w5[f];
                }
            }
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            const int64_t i = base + 16 * g + lr;
            if (lg == 0 && i < n) {
                const float d = softplusf(s + b5) * p.scale;
                if (capMargin >= 0.f) {
                    const float r = d * (1.f + capMargin) + p.slack;
                    out[i] = capSquared ? r * r : r;
                } else {
                    out[i] = d;
                }
            }
        }
    }
}

// squared caps of the slab's voxels from their bricks' radii: (r_b + |x - c_b|)^2
__global__ void kBrickCaps(int dx, int dy, int64_t count, int nbx, int nby, Xf3 A, const float* __restrict__ radius,
                           float* __restrict__ capSq)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t layer = (int64_t)dx * dy;
    const int z = (int)(i / layer), rem = (int)(i % layer);
    const int x = rem % dx, y = rem / dx;
    const int64_t b = ((int64_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> 2);
    // offset from the brick centre in grid steps, then in space
    const float gx = (float)(x & 3) - 1.5f, gy = (float)(y & 3) - 1.5f, gz = (float)(z & 3) - 1.5f;
    const float wx = A.m[0] * gx + A.m[1] * gy + A.m[2] * gz;
    const float wy = A.m[3] * gx + A.m[4] * gy + A.m[5] * gz;
    const float wz = A.m[6] * gx + A.m[7] * gy + A.m[8] * gz;
    const float r = radius[b] + sqrtf(wx * wx + wy * wy + wz * wz);
    capSq[i] = r * r;
}

__global__ void kReduce(const float* __restrict__ partial, int nBlocks, float* __restrict__ grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kParamsPadded) return;
    float s = 0.f;
    for (int b = 0; b < nBlocks; ++b) s += partial[(size_t)b * kParamsPadded + j];
    grad[j] = s;
}

// refreshes the BF16 copies (W, W^T, w5) from the FP32 master
__device__ __forceinline__ void refresh(const Params& p, int j, float v)
{
    if (j < kOffB) {
        const int l = j / (kHidden * kHidden), rem = j % (kHidden * kHidden);
        const int o = rem / kHidden, i = rem % kHidden;
        const uint16_t b = f2bf(v);
        p.WbMut[j] = b;
        p.WbTMut[(size_t)l * kHidden * kHidden + i * kHidden + o] = b;
    } else if (j >= kOffW5 && j < kOffW5 + kHidden) {
        p.w5bMut[j - kOffW5] = f2bf(v);
    }
}

// one Adam step straight from the backward pass's per-tile partial gradients (their sum in four
// interleaved accumulators, then the update): no separate reduction launch
__global__ void kAdamPartials(Params p, const float* __restrict__ partial, int nBlocks, float* __restrict__ mom,
                              float* __restrict__ vel, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kParams) return;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int b = 0;
    for (; b + 4 <= nBlocks; b += 4) {
        s0 += partial[(size_t)b * kParamsPadded + j];
        s1 += partial[(size_t)(b + 1) * kParamsPadded + j];
        s2 += partial[(size_t)(b + 2) * kParamsPadded + j];
        s3 += partial[(size_t)(b + 3) * kParamsPadded + j];
    }
    for (; b < nBlocks; ++b) s0 += partial[(size_t)b * kParamsPadded + j];
    const float g = (s0 + s1) + (s2 + s3);
    const float m = b1 * mom[j] + (1.f - b1) * g;
    const float v = b2 * vel[j] + (1.f - b2) * g * g;
    mom[j] = m;
    vel[j] = v;
    const float th = p.thetaMut[j] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    p.thetaMut[j] = th;
    refresh(p, j, th);
}

__global__ void kRefreshAll(Params p)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kParams) return;
    refresh(p, j, p.theta[j]);
}

} // namespace

namespace {
int predictBlocks(int64_t n)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess || cus <= 0)
            cus = 256;
    }
    const int64_t tiles = (n + kPredPoints - 1) / kPredPoints;
    const int64_t want = (tiles + kPredWaves - 1) / kPredWaves;
    return (int)(want < cus ? want : cus);
}
} // namespace

hipError_t forward(const Params& p, const float* pts, int64_t n, float* out, float capMargin, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    PredSource src{};
    src.pts = pts;
    hipLaunchKernelGGL(kPredict, dim3(predictBlocks(n)), dim3(64 * kPredWaves), 0, s, p, src, n, out, capMargin, 1);
    return hipGetLastError();
}

int64_t gridCapBricks(const int32_t dims[3], int32_t zBegin, int32_t zEnd)
{
    if (dims[0] <= 0 || dims[1] <= 0 || zEnd <= zBegin) return 0;
    return (int64_t)((dims[0] + 3) / 4) * ((dims[1] + 3) / 4) * ((zEnd - zBegin + 3) / 4);
}

hipError_t gridCaps(const Params& p, const int32_t dims[3], const float xf[12], int32_t zBegin, int32_t zEnd,
                    float capMargin, float* radius, float* capSq, hipStream_t s)
{
    const int64_t bricks = gridCapBricks(dims, zBegin, zEnd);
    if (bricks <= 0) return hipSuccess;
    if (!(capMargin >= 0.f)) return hipErrorInvalidValue;
    PredSource src{};
    src.pts = nullptr;
    src.nbx = (dims[0] + 3) / 4;
    src.nby = (dims[1] + 3) / 4;
    src.zBegin = zBegin;
    for (int i = 0; i < 12; ++i) src.xf[i] = xf[i];
    hipLaunchKernelGGL(kPredict, dim3(predictBlocks(bricks)), dim3(64 * kPredWaves), 0, s, p, src, bricks, radius,
                       capMargin, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t count = (int64_t)dims[0] * dims[1] * (zEnd - zBegin);
    Xf3 A;
    for (int i = 0; i < 9; ++i) A.m[i] = xf[i];
    hipLaunchKernelGGL(kBrickCaps, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dims[0], dims[1], count,
                       src.nbx, src.nby, A, radius, capSq);
    return hipGetLastError();
}

hipError_t gradient(const Params& p, const float* pts, const float* labels, int64_t n, const TrainBuffers& tb,
                    float* grad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (int)((n + kTileRows - 1) / kTileRows);
    hipLaunchKernelGGL(kTrainTile, dim3(blocks), dim3(256), 0, s, p, pts, labels, n, 1.f / (float)n, tb.partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kReduce, dim3((kParamsPadded + 255) / 256), dim3(256), 0, s, tb.partial, blocks, grad);
    return hipGetLastError();
}

hipError_t trainStep(const Params& p, const float* pts, const float* labels, int64_t n, const TrainBuffers& tb,
                     float* mom, float* vel, int step, float lr, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (int)((n + kTileRows - 1) / kTileRows);
    hipLaunchKernelGGL(kTrainTile, dim3(blocks), dim3(256), 0, s, p, pts, labels, n, 1.f / (float)n, tb.partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, step)));
    hipLaunchKernelGGL(kAdamPartials, dim3((kParams + 255) / 256), dim3(256), 0, s, p, tb.partial, blocks, mom, vel,
                       lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t refreshCopies(const Params& p, hipStream_t s)
{
    hipLaunchKernelGGL(kRefreshAll, dim3((kParams + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

} // namespace mlp
} // namespace mi
