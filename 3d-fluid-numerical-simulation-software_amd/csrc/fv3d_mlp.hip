// Learned eddy-viscosity closure (no reference counterpart): a 6 -> 64 -> 64
// -> 64 -> 1 perceptron from local non-dimensional flow features to
// log(1 + mue_t / mue_lam), BF16 weights and activations into the matrix
// cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, tanh hidden layers,
// softplus output.  Inference runs per cell inside the residual evaluation in
// place of the Spalart-Allmaras transport; training runs on the GPU with the
// same forward code and a hand-written backward pass (per-chunk partial
// gradients, a fixed-order reduction, Adam).
//
// Matrix-core lane layout of v_mfma_f32_16x16x32_bf16 (one wave, D = A B + C,
// A 16 x 32, B 32 x 16): lane l supplies A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16], j = 0..7, and receives D[4 (l / 16) + i][l % 16],
// i = 0..3.
#include "fv3d_kernels.h"
#include "fv3d_mlp.h"

#include <math.h>

namespace fv3d {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kHS = kMlpH + 4;   // LDS row stride of an activation tile (floats)

__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline float softplus(float z) { return z > 20.0f ? z : log1pf(expf(z)); }
__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// One wave: rows r0 .. r0+15 of the activation tile in (row stride kHS,
// columns 0..63) times the weight matrix W (row n = output, 64 inputs, FP32
// in global memory, quantised to BF16 here), plus bias, tanh, into out.
__device__ inline void dense64(const float *in, int r0, const float *W, const float *b, float *out, int lane)
{
    const int row = lane & 15, kb = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const int n = nt * 16 + row;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8 a, w;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 32 * ks + 8 * kb + j;
                a[j] = (__bf16)in[(r0 + row) * kHS + k];
                w[j] = (__bf16)W[n * kMlpH + k];
            }
            acc = mfma16(a, w, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(r0 + 4 * kb + i) * kHS + n] = tanhf(acc[i] + b[n]);
    }
}

// The forward pass of one wave's 16 samples (rows r0..r0+15): x (row stride
// kMlpXS, features 0..5) -> h0, h1, h2 (tanh activations, FP32 in LDS) -> z
// (pre-softplus output) for lanes 0..15.
__device__ inline float mlp_forward16(const float *P, const float *x, float *h0, float *h1, float *h2, int r0,
                                      int lane)
{
    const int row = lane & 15, kb = lane >> 4;
    // layer 0: K = 6 features in the first 8-wide slice, zero elsewhere
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const int n = nt * 16 + row;
        bf16x8 a, w;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool in = kb == 0 && j < kMlpIn;
            a[j] = (__bf16)(in ? x[(r0 + row) * kMlpXS + j] : 0.0f);
            w[j] = (__bf16)(in ? P[kMlpW0 + n * kMlpIn + j] : 0.0f);
        }
        acc = mfma16(a, w, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h0[(r0 + 4 * kb + i) * kHS + n] = tanhf(acc[i] + P[kMlpB0 + n]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    dense64(h0, r0, P + kMlpW1, P + kMlpB1, h1, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    dense64(h1, r0, P + kMlpW2, P + kMlpB2, h2, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    float z = P[kMlpB3];
    if (lane < 16)
        for (int k = 0; k < kMlpH; ++k) z += h2[(r0 + lane) * kHS + k] * P[kMlpW3 + k];
    return z;
}

// ------------------------------------------------- register-resident inference
// Inference keeps the network in registers: each wave holds the BF16 weight
// fragments of layers 0-2 and runs the transposed products h^T = W x^T (the
// weights are the A operand, 16 samples the B operand), so that the output
// fragment of one layer is, lane for lane, the B operand of the next.  Lane l
// receives units 16 nt + 4 (l / 16) + i (nt, i = 0..3) of sample l % 16; the
// next layer's K index is ordered to match: slice ks, position 8 kb + j holds
// unit 16 (2 ks + j / 4) + 4 kb + j % 4, and the weights are loaded in that
// order.  The hidden activations never leave the registers.
struct MlpRegs {
    bf16x8 w0[4], w1[8], w2[8];   // [nt] and [2 nt + ks]
};

__device__ inline int k_unit(int ks, int kb, int j) { return 16 * (2 * ks + (j >> 2)) + 4 * kb + (j & 3); }

__device__ inline void mlp_load(const float *__restrict__ P, int lane, MlpRegs &R)
{
    const int row = lane & 15, kb = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            R.w0[nt][j] = (__bf16)((kb == 0 && j < kMlpIn) ? P[kMlpW0 + (16 * nt + row) * kMlpIn + j] : 0.0f);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int u = k_unit(ks, kb, j);
                R.w1[2 * nt + ks][j] = (__bf16)P[kMlpW1 + (16 * nt + row) * kMlpH + u];
                R.w2[2 * nt + ks][j] = (__bf16)P[kMlpW2 + (16 * nt + row) * kMlpH + u];
            }
    }
}

// tanh(x) = 1 - 2 / (1 + e^(2x)): saturates to +-1 without overflow trouble,
// absolute error near 1e-7, far below the BF16 rounding of the activations.
__device__ inline float tanh_fast(float x) { return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * x)); }

__device__ inline void dense64_t(const bf16x8 *W, const volatile float *b, float h[4][4], int kb)
{
    bf16x8 in[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) in[ks][j] = (__bf16)h[2 * ks + (j >> 2)][j & 3];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = *(const volatile f32x4 *)(b + 16 * nt + 4 * kb);
        acc = mfma16(W[2 * nt], in[0], acc);
        acc = mfma16(W[2 * nt + 1], in[1], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h[nt][i] = tanh_fast(acc[i]);
    }
}

// The small parameters of the network in LDS (biases of layers 0-2, the
// output weights and bias), staged once per workgroup: read there through a
// volatile pointer at each use, so the compiler keeps them out of the
// registers that hold the weight fragments.
constexpr int kSmallB0 = 0, kSmallB1 = 64, kSmallB2 = 128, kSmallW3 = 192, kSmallB3 = 256, kSmall = 260;

__device__ inline void stage_small(const float *__restrict__ P, float *sp, int t)
{
    for (int e = t; e < kSmall; e += blockDim.x) {
        float v = 0.0f;
        if (e < kSmallB3) v = P[(e < kSmallB1 ? kMlpB0 : e < kSmallB2 ? kMlpB1 - 64 : e < kSmallW3 ? kMlpB2 - 128
                                 : kMlpW3 - 192) + e];
        else if (e == kSmallB3) v = P[kMlpB3];
        sp[e] = v;
    }
}

// The pre-softplus output z of 16 samples; xb is the layer-0 B operand (lanes
// 0..15: the six features of sample l in positions 0..5, zeros elsewhere).
// Every lane returns z of sample l % 16.
__device__ inline float mlp_infer16(const MlpRegs &R, const float *sp, bf16x8 xb, int lane)
{
    const int kb = lane >> 4;
    const volatile float *vs = sp;
    float h[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = *(const volatile f32x4 *)(vs + kSmallB0 + 16 * nt + 4 * kb);
        acc = mfma16(R.w0[nt], xb, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h[nt][i] = tanh_fast(acc[i]);
    }
    dense64_t(R.w1, vs + kSmallB1, h, kb);
    dense64_t(R.w2, vs + kSmallB2, h, kb);
    float z = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const f32x4 w = *(const volatile f32x4 *)(vs + kSmallW3 + 16 * nt + 4 * kb);
#pragma unroll
        for (int i = 0; i < 4; ++i) z += w[i] * h[nt][i];
    }
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    return z + vs[kSmallB3];
}

// Layer-0 operand of samples 16 q .. 16 q + 15 of a wave's 64-row tile sx.
__device__ inline bf16x8 input_operand(const float *sx, int q, int lane)
{
    bf16x8 xb;
    const int row = lane & 15, kb = lane >> 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) xb[j] = (__bf16)((kb == 0 && j < kMlpIn) ? sx[(16 * q + row) * kMlpXS + j] : 0.0f);
    return xb;
}

// ------------------------------------------------------------ features
// The six inputs of a cell, from the cell-centred gradients of the solver's
// own stencil (central differences in index space times the cell
// coordinate-transformation coefficients, as CFluidProps::Curl forms them):
//   0 log(1 + d |U| / nu)            wall-distance Reynolds number
//   1 log(1 + S d^2 / nu)            strain-rate magnitude
//   2 log(1 + Omega d^2 / nu)        vorticity magnitude
//   3 (Omega^2 - S^2) / (Omega^2 + S^2)   Q criterion, normalised
//   4 log(1 + |U| Vol^(1/3) / nu)    cell Reynolds number
//   5 tanh(d U.grad p / (rho |U|^3)) pressure-gradient parameter
// with nu = mue_lam / rho (Sutherland, non-dimensional), S = sqrt(2 S_ij S_ij),
// Omega = sqrt(2 W_ij W_ij), d the wall distance.  The differences of the
// state are taken in FP64 (they cancel), the rest in FP32: the features feed
// a BF16 network.
__device__ inline void cell_features(const BlockDev &B, const Gas &g, int i, int j, int k, float f[kMlpIn])
{
    const long box = B.box, x = cidx(B, i, j, k);
    const long st[3] = {B.si, B.sj, 1};
    float d[3][4];   // d/dkesi, d/dyita, d/dzita of u v w p
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 4; ++q)
            d[a][q] = (float)(0.5 * (B.W[(1 + q) * box + x + st[a]] - B.W[(1 + q) * box + x - st[a]]));
    float gm[9];
    for (int m = 0; m < 9; ++m) gm[m] = (float)B.cg[m * box + x];
    float gr[4][3];   // d(u v w p)/d(x y z)
    for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) gr[q][c] = d[0][q] * gm[c] + d[1][q] * gm[3 + c] + d[2][q] * gm[6 + c];
    float s2 = 0.0f, w2 = 0.0f;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const float sij = 0.5f * (gr[a][b] + gr[b][a]), wij = 0.5f * (gr[a][b] - gr[b][a]);
            s2 += sij * sij;
            w2 += wij * wij;
        }
    const float S = sqrtf(2.0f * s2), Om = sqrtf(2.0f * w2);
    const float rho = (float)B.W[x], u = (float)B.W[box + x], v = (float)B.W[2 * box + x], w = (float)B.W[3 * box + x];
    const float U = sqrtf(u * u + v * v + w * w);
    const float T = (float)B.W[5 * box + x];
    // Sutherland's law (include/FluidProps.h:107-112) in FP32
    const float mue = 1.0f / (float)g.re_ref * T * sqrtf(T) * (1.0f + 110.4f / 288.15f) / (T + 110.4f / 288.15f);
    const float nu = mue / rho;
    const float dw = (float)B.dwall[x];
    const float ugp = u * gr[3][0] + v * gr[3][1] + w * gr[3][2];
    f[0] = log1pf(dw * U / nu);
    f[1] = log1pf(S * dw * dw / nu);
    f[2] = log1pf(Om * dw * dw / nu);
    f[3] = (Om * Om - S * S) / (Om * Om + S * S + 1e-30f);
    f[4] = log1pf(U * cbrtf((float)B.vol[x]) / nu);
    f[5] = tanhf(dw * ugp / (rho * U * U * U + 1e-30f));
}

// Features of every interior cell of every block into feat (6 floats per
// cell, cells in cell-list order without padding: block n at fstart[n]).
__global__ void __launch_bounds__(256) k_mlp_features(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                      int nblock, Gas g, const long *__restrict__ fstart,
                                                      float *__restrict__ feat)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    float f[kMlpIn];
    cell_features(B, g, i, j, k, f);
    const long r = fstart[n] + ((long)(i - 1) * B.nj + (j - 1)) * B.nk + (k - 1);
    for (int q = 0; q < kMlpIn; ++q) feat[r * kMlpIn + q] = f[q];
}

// The training target of every interior cell: log(1 + mue_t / mue_lam) of
// the current eddy viscosity (a Spalart-Allmaras run's), same order as
// k_mlp_features.
__global__ void __launch_bounds__(256) k_mlp_targets(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                     int nblock, Gas g, const long *__restrict__ fstart,
                                                     float *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long x = cidx(B, i, j, k);
    const double mue_lam = sutherland(g, B.W[5 * B.box + x]);
    out[fstart[n] + ((long)(i - 1) * B.nj + (j - 1)) * B.nk + (k - 1)] = (float)log1p(B.mut[x] / mue_lam);
}

// ----------------------------------------------------------- inference
// Network output of n samples (features x, 6 per sample) into y (softplus
// output) and z (pre-activation); 256 samples per workgroup, 64 per wave, the
// network in registers (mlp_infer16), as the closure evaluates it.
__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ P, const float *__restrict__ x, long n,
                                                     float *__restrict__ y, float *__restrict__ zout)
{
    __shared__ float sx[4][64 * kMlpXS];
    __shared__ __attribute__((aligned(16))) float sp[kSmall];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const long s = (long)blockIdx.x * 256 + t;
#pragma unroll
    for (int q = 0; q < kMlpXS; ++q) sx[wave][lane * kMlpXS + q] = (q < kMlpIn && s < n) ? x[s * kMlpIn + q] : 0.0f;
    stage_small(P, sp, t);
    MlpRegs R;
    mlp_load(P, lane, R);
    __syncthreads();
    float zmine = 0.0f;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const float z = mlp_infer16(R, sp, input_operand(sx[wave], q, lane), lane);
        if (q == (lane >> 4)) zmine = z;
    }
    if (s < n) {
        y[s] = softplus(zmine);
        if (zout) zout[s] = zmine;
    }
}

// The closure in the residual evaluation, two passes over the interior cells.
// First the features and the input half of the gate, one thread per cell, into
// xc (8 floats per cell of the padded cell numbering: features 0..5, then 1
// where every feature lies in the training range -- bounds: min, max per
// feature, widened by 10 % of the span -- else 0).
__global__ void __launch_bounds__(256) k_mlp_cell_features(const BlockDev *__restrict__ blocks,
                                                           const long *__restrict__ cstart, int nblock, long ncell, Gas g,
                                                           const float *__restrict__ bounds, float *__restrict__ xc)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    float f[kMlpIn] = {0, 0, 0, 0, 0, 0};
    bool ok = false;
    for_cell_item(blocks, cstart, nblock, t, [&](const BlockDev &B, int, int i, int j, int k) {
        cell_features(B, g, i, j, k, f);
        ok = true;
        for (int q = 0; q < kMlpIn; ++q) {
            const float lo = bounds[2 * q], hi = bounds[2 * q + 1], pad = 0.1f * (hi - lo);
            ok = ok && isfinite(f[q]) && f[q] >= lo - pad && f[q] <= hi + pad;
        }
    });
    float4 *o = (float4 *)(xc + 8 * t);
    o[0] = make_float4(f[0], f[1], f[2], f[3]);
    o[1] = make_float4(f[4], f[5], ok ? 1.0f : 0.0f, 0.0f);
}

// mue_t of cell (i, j, k) of block B from the network output yhat, or, for a
// cell outside the training range (ok false) or whose output is not finite or
// above the largest trained value, the Spalart-Allmaras value of its frozen
// nu~ (fv1 rho nu~ / Re, clipped as SA_Model clips it); true when the
// fallback was taken.
__device__ static bool mut_cell(const BlockDev &B, int i, int j, int k, const Gas &g, const float *__restrict__ bounds,
                                float yhat, bool ok, double relax)
{
    const long box = B.box, x = cidx(B, i, j, k);
    // both loads before any arithmetic, so a wave waits on memory once for them
    const double temp = B.W[5 * box + x], mut_old = B.mut[x];
    const double mue_lam = sutherland(g, temp);
    double mt;
    bool trip = false;
    if (ok && isfinite(yhat) && yhat <= bounds[2 * kMlpIn]) {
        mt = mue_lam * expm1((double)yhat);
    } else {
        trip = true;
        // Spalart-Allmaras fallback from the frozen nu~ (source/SA_Model.cpp:47-87)
        const double rho = B.W[x], ev = B.ev[x] < 0.0 ? 0.0 : B.ev[x];
        const double ev_lam = g.re_ref * mue_lam / rho, X = ev / ev_lam;
        const double Cv1 = F32C(7.1);
        const double fv1 = pow(X, F32C(3.0)) / (pow(X, F32C(3.0)) + Cv1 * Cv1 * Cv1);
        mt = fv1 * rho * ev / g.re_ref;
        if (mt < 0.0) mt = 0.0;
    }
    if (mue_lam > 0 && mt > 200.0 * mue_lam) mt = 200.0 * mue_lam;
    // under-relaxation of the algebraic closure: the fixed point is unchanged
    if (relax != 1.0) mt = mut_old + relax * (mt - mut_old);
    B.mut[x] = mt;
    return trip;
}

// Then the network: 256-cell chunks per workgroup in a grid-stride loop (the
// weight fragments are loaded once per wave for several chunks), each wave
// runs the network on 64 of a chunk's cells, 16 at a time, and leaves the
// softplus output in slot 7 of the cell's xc row.
__global__ void __launch_bounds__(256) k_mlp_closure(const float *__restrict__ P, long ncell, float *__restrict__ xc)
{
    __shared__ float sx[4][64 * kMlpXS];   // per wave: the features of its 64 cells
    __shared__ __attribute__((aligned(16))) float sp[kSmall];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    stage_small(P, sp, t);
    MlpRegs R;
    mlp_load(P, lane, R);
    __syncthreads();
    // every wave of the workgroup runs the same chunks: the matrix cores need
    // all 64 lanes, so cells past the end ride along as zeros
    for (long c0 = (long)blockIdx.x * 256; c0 < ncell; c0 += (long)gridDim.x * 256) {
        const long c = c0 + t;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
        if (c < ncell) {
            a = ((const float4 *)(xc + 8 * c))[0];
            b = ((const float4 *)(xc + 8 * c))[1];
        }
        const bool ok = b.z != 0.0f;
        float *r = sx[wave] + lane * kMlpXS;
        r[0] = ok ? a.x : 0.0f, r[1] = ok ? a.y : 0.0f, r[2] = ok ? a.z : 0.0f, r[3] = ok ? a.w : 0.0f;
        r[4] = ok ? b.x : 0.0f, r[5] = ok ? b.y : 0.0f, r[6] = 0.0f, r[7] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        float zmine = 0.0f;
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const float z = mlp_infer16(R, sp, input_operand(sx[wave], q, lane), lane);
            if (q == (lane >> 4)) zmine = z;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (c < ncell) xc[8 * c + 7] = softplus(zmine);
    }
}

// mue_t of every interior cell from the network outputs in xc; tripped counts
// the fallbacks.
__global__ void __launch_bounds__(256) k_mlp_mut(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                 int nblock, Gas g, const float *__restrict__ bounds,
                                                 const float *__restrict__ xc,
                                                 unsigned long long *__restrict__ tripped, double relax)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // the gate flag and network output, loaded unconditionally (an index past
    // the list reads item 0, unused) before the block search, which they do
    // not depend on, so the wave waits on memory once for them and the block
    const float2 v = ((const float2 *)(xc + 8 * (c < cstart[nblock] ? c : 0)))[3];
    bool trip = false;
    for_cell_item(blocks, cstart, nblock, c, [&](const BlockDev &B, int, int i, int j, int k) {
        trip = mut_cell(B, i, j, k, g, bounds, v.y, v.x != 0.0f, relax);
    });
    // one atomic per wave for the fallback count (every lane reaches this
    // point: the grid is whole 256-thread workgroups and no lane returns),
    // into the workgroup's slot
    const unsigned long long m = __ballot(trip);
    if ((threadIdx.x & 63) == 0 && m)
        atomicAdd(tripped + (blockIdx.x % kTripSlots) * kTripStride, (unsigned long long)__popcll(m));
}

// ------------------------------------------------------------ training
// Loss = mean over the n samples of (softplus(z) - t)^2.  One workgroup per
// chunk of 64 samples: forward (activations kept in LDS), then the backward
// pass; the chunk's gradient goes to partial[chunk * kMlpParams + p] and its
// sum of squared errors to loss[chunk].  Matrix products of the backward pass
// run on the matrix cores with BF16 operands (deltas and activations
// quantised as the forward pass quantises them), FP32 accumulation.
__global__ void __launch_bounds__(256) k_mlp_train_chunk(const float *__restrict__ P, const float *__restrict__ x,
                                                         const float *__restrict__ tgt, long n, float inv_n,
                                                         float *__restrict__ partial, float *__restrict__ loss)
{
    __shared__ float sx[64 * kMlpXS];
    __shared__ float h[3][64 * kHS];   // activations, overwritten by the deltas
    __shared__ float dz[64];
    __shared__ float red[64];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, row = lane & 15, kb = lane >> 4;
    const long s0 = (long)blockIdx.x * 64;
    float *G = partial + (long)blockIdx.x * kMlpParams;
    for (int e = t; e < 64 * kMlpXS; e += 256) {
        const int r = e / kMlpXS, q = e % kMlpXS;
        sx[e] = (q < kMlpIn && s0 + r < n) ? x[(s0 + r) * kMlpIn + q] : 0.0f;
    }
    __syncthreads();
    const float z = mlp_forward16(P, sx, h[0], h[1], h[2], wave * 16, lane);
    if (lane < 16) {
        const int r = wave * 16 + lane;
        const bool in = s0 + r < n;
        const float e = in ? softplus(z) - tgt[s0 + r] : 0.0f;
        dz[r] = in ? 2.0f * inv_n * e * sigmoid(z) : 0.0f;
        red[r] = e * e;
    }
    __syncthreads();
    if (t == 0) {
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += red[r];
        loss[blockIdx.x] = s;
    }
    // output layer: dW3, db3, then delta2 = dz W3 (1 - h2^2) in place
    if (t < kMlpH) {
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += dz[r] * h[2][r * kHS + t];
        G[kMlpW3 + t] = s;
    } else if (t == kMlpH) {
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += dz[r];
        G[kMlpB3] = s;
    }
    __syncthreads();
    for (int e = t; e < 64 * kMlpH; e += 256) {
        const int r = e / kMlpH, c = e % kMlpH;
        const float a = h[2][r * kHS + c];
        h[2][r * kHS + c] = dz[r] * P[kMlpW3 + c] * (1.0f - a * a);
    }
    __syncthreads();
    // hidden layers L = 2, 1: delta in h[L], activations of the layer below in h[L-1]
    for (int L = 2; L >= 1; --L) {
        const float *Wl = P + (L == 2 ? kMlpW2 : kMlpW1);
        const int oW = L == 2 ? kMlpW2 : kMlpW1, oB = L == 2 ? kMlpB2 : kMlpB1;
        float *dl = h[L], *al = h[L - 1];
        // dW[n][k] = sum_s delta[s][n] a[s][k]: output tiles (mt, nt), four per wave
        for (int tile = wave * 4; tile < wave * 4 + 4; ++tile) {
            const int mt = tile >> 2, nt = tile & 3;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < 2; ++ks) {
                bf16x8 a, b;
                for (int j = 0; j < 8; ++j) {
                    const int s = 32 * ks + 8 * kb + j;
                    a[j] = (__bf16)dl[s * kHS + mt * 16 + row];
                    b[j] = (__bf16)al[s * kHS + nt * 16 + row];
                }
                acc = mfma16(a, b, acc);
            }
            for (int i = 0; i < 4; ++i) G[oW + (mt * 16 + 4 * kb + i) * kMlpH + nt * 16 + row] = acc[i];
        }
        if (t < kMlpH) {
            float s = 0.0f;
            for (int r = 0; r < 64; ++r) s += dl[r * kHS + t];
            G[oB + t] = s;
        }
        __syncthreads();
        // delta below[s][k] = (sum_n delta[s][n] W[n][k]) (1 - a[s][k]^2), in place of a
        for (int tile = wave * 4; tile < wave * 4 + 4; ++tile) {
            const int mt = tile >> 2, nt = tile & 3;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < 2; ++ks) {
                bf16x8 a, b;
                for (int j = 0; j < 8; ++j) {
                    const int nn = 32 * ks + 8 * kb + j;
                    a[j] = (__bf16)dl[(mt * 16 + row) * kHS + nn];
                    b[j] = (__bf16)Wl[nn * kMlpH + nt * 16 + row];
                }
                acc = mfma16(a, b, acc);
            }
            for (int i = 0; i < 4; ++i) {
                const int s = mt * 16 + 4 * kb + i, c = nt * 16 + row;
                const float av = al[s * kHS + c];
                al[s * kHS + c] = acc[i] * (1.0f - av * av);
            }
        }
        __syncthreads();
    }
    // input layer: dW0[n][f] = sum_s delta0[s][n] x[s][f], db0
    for (int o = t; o < kMlpH * kMlpIn; o += 256) {
        const int nn = o / kMlpIn, f = o % kMlpIn;
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += h[0][r * kHS + nn] * (float)(__bf16)sx[r * kMlpXS + f];
        G[kMlpW0 + o] = s;
    }
    if (t < kMlpH) {
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s += h[0][r * kHS + t];
        G[kMlpB0 + t] = s;
    }
}

// Fixed-order sum of the chunk gradients (and losses) of one step.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, const float *__restrict__ loss,
                                                    int nchunk, float *__restrict__ grad, float *__restrict__ loss_sum)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float s = 0.0f;
        for (int c = 0; c < nchunk; ++c) s += partial[(long)c * kMlpParams + p];
        grad[p] = s;
    } else if (p == kMlpParams) {
        float s = 0.0f;
        for (int c = 0; c < nchunk; ++c) s += loss[c];
        *loss_sum = s;
    }
}

// Adam (Kingma and Ba 2015) on the FP32 master weights.
__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ grad, float lr, float b1, float b2, float eps,
                                                  float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr, vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace fv3d
