// mi_mlp.hpp -- the learned distance bound's network (kernels in mi_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mi {
namespace mlp {

constexpr int kHidden = 128;     // hidden width
constexpr int kLayers = 4;       // hidden layers (Linear + SiLU)
constexpr int kBands = 16;       // Fourier bands
constexpr int kFeatures = 3 + 6 * kBands;   // 99 input features (zero-padded to kHidden)
constexpr int kTileRows = 64;    // rows per workgroup

// parameter vector layout (FP32): W_l [out][in] (layer 0 uses in < kFeatures), b_l, w5, b5
constexpr int kOffW = 0;
constexpr int kOffB = kLayers * kHidden * kHidden;
constexpr int kOffW5 = kOffB + kLayers * kHidden;
constexpr int kOffB5 = kOffW5 + kHidden;
constexpr int kParams = kOffB5 + 1;
constexpr int kParamsPadded = (kParams + 255) / 256 * 256;

struct Params {
    const float* theta;      // FP32 master parameters
    float* thetaMut;
    const uint16_t* Wb;      // BF16 W_l, [l][out][in]
    uint16_t* WbMut;
    const uint16_t* WbT;     // BF16 W_l^T, [l][in][out]
    uint16_t* WbTMut;
    const uint16_t* w5b;     // BF16 output weights
    uint16_t* w5bMut;
    float center[3];
    float invScale, scale;
    float piFreq[kBands];    // pi * 2^(k/3)
    float slack;             // calibrated allowance added to every capped radius (0 until calibrated)
};

struct TrainBuffers {
    float* partial;    // [ceil(n / kTileRows)][kParamsPadded] per-tile gradients
};

// predicted distance (capMargin < 0) or capped squared search radius (pred (1 + capMargin) + slack)^2 per
// point, written at the point's index
hipError_t forward(const Params& p, const float* pts, int64_t n, float* out, float capMargin, hipStream_t s);
// Squared caps of the voxels of the z-slab [zBegin, zEnd) of a grid (x fastest): one prediction per 4x4x4
// brick at its centre c, r = pred(c) (1 + capMargin) + slack, and for voxel x the cap (r + |x - c|)^2 -- an
// upper bound of x's distance whenever r bounds c's (distance is 1-Lipschitz).  radius: scratch of one
// float per brick.
hipError_t gridCaps(const Params& p, const int32_t dims[3], const float xf[12], int32_t zBegin, int32_t zEnd,
                    float capMargin, float* radius, float* capSq, hipStream_t s);
// bricks of the slab (scratch size of gridCaps' radius, in floats)
int64_t gridCapBricks(const int32_t dims[3], int32_t zBegin, int32_t zEnd);
// gradient of 1/(2n) sum (pred/scale - label/scale)^2 over the n points into grad[kParamsPadded]
hipError_t gradient(const Params& p, const float* pts, const float* labels, int64_t n, const TrainBuffers& tb,
                    float* grad, hipStream_t s);
// one training step (forward, backward, Adam on the summed per-tile gradients) over n points
hipError_t trainStep(const Params& p, const float* pts, const float* labels, int64_t n, const TrainBuffers& tb,
                     float* mom, float* vel, int step, float lr, hipStream_t s);
hipError_t refreshCopies(const Params& p, hipStream_t s);

} // namespace mlp
} // namespace mi
