// The learned eddy-viscosity closure: network shape, parameter layout and
// kernels (fv3d_mlp.hip).
#pragma once

#include "fv3d_device.h"

namespace fv3d {

// 6 -> 64 -> 64 -> 64 -> 1; FP32 parameters in one array:
// W0[64][6] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64] b3
constexpr int kMlpIn = 6, kMlpH = 64, kMlpXS = 8;   // kMlpXS: LDS row stride of the inputs
// the fallback count of k_mlp_mut: kTripSlots counters kTripStride words
// apart (one cache line each), workgroup w adding to slot w mod kTripSlots, so
// that the atomics of a launch do not queue on one address; the reader sums them
constexpr int kTripSlots = 64, kTripStride = 16;
constexpr int kMlpW0 = 0, kMlpB0 = kMlpW0 + kMlpH * kMlpIn, kMlpW1 = kMlpB0 + kMlpH, kMlpB1 = kMlpW1 + kMlpH * kMlpH,
              kMlpW2 = kMlpB1 + kMlpH, kMlpB2 = kMlpW2 + kMlpH * kMlpH, kMlpW3 = kMlpB2 + kMlpH, kMlpB3 = kMlpW3 + kMlpH,
              kMlpParams = kMlpB3 + 1;
static_assert(kMlpParams == 8833, "6-64-64-64-1");

__global__ void k_mlp_features(const BlockDev *blocks, const long *cstart, int nblock, Gas g, const long *fstart,
                               float *feat);
__global__ void k_mlp_targets(const BlockDev *blocks, const long *cstart, int nblock, Gas g, const long *fstart,
                              float *out);
__global__ void k_mlp_forward(const float *P, const float *x, long n, float *y, float *zout);
// bounds: min and max of each feature, then the largest output (2 kMlpIn + 1)
__global__ void k_mlp_cell_features(const BlockDev *blocks, const long *cstart, int nblock, long ncell, Gas g,
                                    const float *bounds, float *xc);
__global__ void k_mlp_closure(const float *P, long ncell, float *xc);
__global__ void k_mlp_mut(const BlockDev *blocks, const long *cstart, int nblock, Gas g, const float *bounds,
                          const float *xc, unsigned long long *tripped, double relax);
__global__ void k_mlp_train_chunk(const float *P, const float *x, const float *tgt, long n, float inv_n, float *partial,
                                  float *loss);
__global__ void k_mlp_reduce(const float *partial, const float *loss, int nchunk, float *grad, float *loss_sum);
__global__ void k_mlp_adam(float *P, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                           float c1, float c2);

}  // namespace fv3d
