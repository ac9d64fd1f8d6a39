// Two-grid cycle transfers (source/Solver.cpp:61-103, source/Prolong_Interpolate.cpp,
// source/ComputeAllBlockVars.cpp:269-309 and 374-439).
//
// The coarse level is a second set of blocks with nodes (2i-1, 2j-1, 2k-1) of
// the fine ones; coarse cell (I,J,K) covers fine cells 2I-1..2I along each
// axis.  Restriction and prolongation are one thread per cell and bound by HBM
// reads of the finer level; every coarse value a prolongation thread reads is
// shared with its neighbours along k and comes from L2.
#include "fv3d_kernels.h"

namespace fv3d {

// W from Q for one interior cell (CFluidProps::PrimitiveVars,
// source/PrimitiveVars.cpp:7-29) and the limiter's bad-cell test
// (source/Limit_FlowVars.cpp:26-28)
__device__ static inline void prim_vars(const BlockDev &B, long x, const double q[5], const Gas &g, int *bad)
{
    const long box = B.box;
    const double Gamma = g.gamma, Rgas = g.rgas;
    double d = q[0], u = q[1] / q[0], v = q[2] / q[0], w = q[3] / q[0];
    double V2 = u * u + v * v + w * w;
    double p = (Gamma - F32C(1.0)) * (q[4] - q[0] * V2 / F32C(2.0));
    double T = p / d / Rgas;
    B.W[x] = d; B.W[box + x] = u; B.W[2 * box + x] = v; B.W[3 * box + x] = w; B.W[4 * box + x] = p; B.W[5 * box + x] = T;
    const double dmin = F32C(1.0E-6), dmax = F32C(1000.0), vmax = F32C(1000.0);
    if (d < dmin || d > dmax || p < dmin || p > dmax || u > vmax || v > vmax || w > vmax) atomicOr(bad, 1);
}

// CFluidProps::PrimitiveVars over every interior cell (source/PrimitiveVars.cpp:
// 7-29), with the limiter's bad-cell flags.
__global__ void __launch_bounds__(256) k_prim_all(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                  int nblock, Gas g, int *__restrict__ bad_flag)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long x = cidx(B, i, j, k);
    double q[5];
    for (int c = 0; c < 5; ++c) q[c] = B.Q[c * B.box + x];
    prim_vars(B, x, q, g, bad_flag + n);
}

// CSolver::Prolong (fine to coarse, source/Prolong_Interpolate.cpp:7-50): the
// volume-weighted state cv_2h = sum(Q V) / V_H, which becomes the coarse
// state, and Res_2h = sum(R) of the eight fine cells, in the reference's order
// (111, 112, 121, 122, 211, 212, 221, 222).  Fused: the coarse PrimitiveVars
// and the forcing of the coarse LU-SGS (ComputeForceFunction /
// AddForceFunction, source/ComputeAllBlockVars.cpp:269-309):
// R = R + (Res_2h - R), R the coarse residual the previous forced sweep used.
__global__ void __launch_bounds__(256) k_mg_restrict(const BlockDev *__restrict__ fine, const BlockDev *__restrict__ coarse,
                                                     const long *__restrict__ cstart, int nblock, Gas g,
                                                     int *__restrict__ bad_flag)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(coarse, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &F = fine[n], &B = coarse[n];
    double sq[5], sr[5];
    bool first = true;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                const long y = cidx(F, 2 * i - 1 + a, 2 * j - 1 + b, 2 * k - 1 + c);
                const double v = F.vol[y];
                for (int q = 0; q < 5; ++q) {
                    const double qv = F.Q[q * F.box + y] * v, rr = F.R[q * F.box + y];
                    sq[q] = first ? qv : sq[q] + qv;
                    sr[q] = first ? rr : sr[q] + rr;
                }
                first = false;
            }
    const long x = cidx(B, i, j, k), box = B.box;
    const double V2 = B.vol[x];
    double q[5];
    for (int c = 0; c < 5; ++c) {
        q[c] = sq[c] / V2;
        B.Q[c * box + x] = q[c];
        B.cv2h[c * box + x] = q[c];
        B.r2h[c * box + x] = sr[c];
        const double rc = B.R[c * box + x];
        const double qf = sr[c] - rc;
        B.R[c * box + x] = rc + qf;
    }
    prim_vars(B, x, q, g, bad_flag + n);
}

// CMesh::CornerTreatment (source/ComputeAllBlockVars.cpp:374-439) on the coarse
// state: the twelve edges of the ghost box from their two face-ghost
// neighbours, then the eight corners from three edges and three face ghosts.
// One workgroup per block.  (The same treatment of the correction array only
// ever averages zeros: its ghost cells are never written.)
__global__ void __launch_bounds__(256) k_mg_corner(const BlockDev *__restrict__ coarse)
{
    const BlockDev &B = coarse[blockIdx.x];
    const int IM = B.IM, JM = B.JM, KM = B.KM;
    const long box = B.box;
    const int ne = 4 * (IM - 1) + 4 * (JM - 1) + 4 * (KM - 1);
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        long x, p, q;
        int s = e;
        if (s < 4 * (IM - 1)) {
            const int i = s / 4 + 1, w = s % 4;
            if (w == 0) { x = cidx(B, i, 0, 0); p = cidx(B, i, 1, 0); q = cidx(B, i, 0, 1); }
            else if (w == 1) { x = cidx(B, i, JM, 0); p = cidx(B, i, JM - 1, 0); q = cidx(B, i, JM, 1); }
            else if (w == 2) { x = cidx(B, i, 0, KM); p = cidx(B, i, 0, KM - 1); q = cidx(B, i, 1, KM); }
            else { x = cidx(B, i, JM, KM); p = cidx(B, i, JM - 1, KM); q = cidx(B, i, JM, KM - 1); }
        } else if ((s -= 4 * (IM - 1)) < 4 * (JM - 1)) {
            const int j = s / 4 + 1, w = s % 4;
            if (w == 0) { x = cidx(B, 0, j, 0); p = cidx(B, 1, j, 0); q = cidx(B, 0, j, 1); }
            else if (w == 1) { x = cidx(B, 0, j, KM); p = cidx(B, 1, j, KM); q = cidx(B, 0, j, KM - 1); }
            else if (w == 2) { x = cidx(B, IM, j, KM); p = cidx(B, IM - 1, j, KM); q = cidx(B, IM, j, KM - 1); }
            else { x = cidx(B, IM, j, 0); p = cidx(B, IM - 1, j, 0); q = cidx(B, IM, j, 1); }
        } else {
            s -= 4 * (JM - 1);
            const int k = s / 4 + 1, w = s % 4;
            if (w == 0) { x = cidx(B, 0, 0, k); p = cidx(B, 1, 0, k); q = cidx(B, 0, 1, k); }
            else if (w == 1) { x = cidx(B, IM, 0, k); p = cidx(B, IM, 1, k); q = cidx(B, IM - 1, 0, k); }
            else if (w == 2) { x = cidx(B, 0, JM, k); p = cidx(B, 1, JM, k); q = cidx(B, 0, JM - 1, k); }
            else { x = cidx(B, IM, JM, k); p = cidx(B, IM - 1, JM, k); q = cidx(B, IM, JM - 1, k); }
        }
        for (int c = 0; c < 5; ++c) B.Q[c * box + x] = F32C(0.5) * (B.Q[c * box + p] + B.Q[c * box + q]);
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        // corner (ci,cj,ck) and its inward neighbours along each axis
        const int w = threadIdx.x;
        const int ci = (w & 1) ? IM : 0, cj = (w & 2) ? JM : 0, ck = (w & 4) ? KM : 0;
        const int i1 = ci ? IM - 1 : 1, j1 = cj ? JM - 1 : 1, k1 = ck ? KM - 1 : 1;
        const long x = cidx(B, ci, cj, ck);
        const long e1 = cidx(B, i1, cj, ck), e2 = cidx(B, ci, j1, ck), e3 = cidx(B, ci, cj, k1);
        const long f1 = cidx(B, i1, j1, ck), f2 = cidx(B, i1, cj, k1), f3 = cidx(B, ci, j1, k1);
        for (int c = 0; c < 5; ++c) {
            const double *Q = B.Q + c * box;
            const double s2 = ((Q[e1] + Q[e2]) + Q[e3]) * F32C(2.0);
            const double s1 = (Q[f1] + Q[f2]) + Q[f3];
            B.Q[c * box + x] = (s2 - s1) / F32C(3.0);
        }
    }
}

// CSolver::Interpolate (coarse to fine, source/Prolong_Interpolate.cpp:52-156):
// weights 27, 9, 3, 1 (/64) on the nearest coarse cell, the three sharing a
// face with it on the fine cell's side, the three sharing an edge and the one
// sharing a corner; STATE sets the fine state from the coarse one (end of the
// coarse phase), else the coarse correction Q_H - cv_2h is added (zero in the
// coarse ghost cells, which the reference never writes).  A workgroup takes a
// tile of kMgTI x kMgTJ x kMgTK coarse cells: it stages their values and a
// one-cell halo in LDS (each coarse value read once), then forms the eight
// fine cells under every coarse cell of the tile.  Fused: the fine
// PrimitiveVars and the bad-cell flags for the limiter.
template <bool STATE>
__global__ void __launch_bounds__(256) k_mg_prolong(const BlockDev *__restrict__ fine, const BlockDev *__restrict__ coarse,
                                                    const int4 *__restrict__ tiles, Gas g, int *__restrict__ bad_flag)
{
    constexpr int HI = kMgTI + 2, HJ = kMgTJ + 2, HK = kMgTK + 2, NH = HI * HJ * HK;
    __shared__ double sd[5][NH];
    const int4 tl = tiles[blockIdx.x];
    const int n = tl.x, I0 = tl.y, J0 = tl.z, K0 = tl.w;
    const BlockDev &F = fine[n], &B = coarse[n];
    const long cbox = B.box;
    for (int e = threadIdx.x; e < NH; e += blockDim.x) {
        const int c3 = e % HK, b3 = (e / HK) % HJ, a3 = e / (HK * HJ);
        const int I = I0 - 1 + a3, J = J0 - 1 + b3, K = K0 - 1 + c3;
        const bool inbox = I <= B.ni + 1 && J <= B.nj + 1 && K <= B.nk + 1;
        const bool interior = I >= 1 && I <= B.ni && J >= 1 && J <= B.nj && K >= 1 && K <= B.nk;
        const long y = inbox ? cidx(B, I, J, K) : 0;
        for (int c = 0; c < 5; ++c) {
            double v = 0.0;
            if (STATE) {
                if (inbox) v = B.Q[c * cbox + y];
            } else if (interior) {
                v = B.Q[c * cbox + y] - B.cv2h[c * cbox + y];
            }
            sd[c][e] = v;
        }
    }
    __syncthreads();
    const double a1 = (double)(27.0f / 64.0f), a2 = (double)(9.0f / 64.0f), a3 = (double)(3.0f / 64.0f),
                 a4 = (double)(1.0f / 64.0f);
    constexpr int FI = 2 * kMgTI, FJ = 2 * kMgTJ, FK = 2 * kMgTK;
    for (int e = threadIdx.x; e < FI * FJ * FK; e += blockDim.x) {
        const int fk = e % FK, fj = (e / FK) % FJ, fi = e / (FK * FJ);
        const int i = 2 * I0 - 1 + fi, j = 2 * J0 - 1 + fj, k = 2 * K0 - 1 + fk;
        if (i > F.ni || j > F.nj || k > F.nk) continue;
        // nearest coarse cell and the neighbour towards the fine cell, in halo coordinates
        const int i1 = (i + 1) / 2 - (I0 - 1), j1 = (j + 1) / 2 - (J0 - 1), k1 = (k + 1) / 2 - (K0 - 1);
        const int i2 = (i & 1) ? i1 - 1 : i1 + 1, j2 = (j & 1) ? j1 - 1 : j1 + 1, k2 = (k & 1) ? k1 - 1 : k1 + 1;
        auto at = [&](int p, int q, int r) { return (p * HJ + q) * HK + r; };
        const int h111 = at(i1, j1, k1), h112 = at(i1, j1, k2), h121 = at(i1, j2, k1), h211 = at(i2, j1, k1);
        const int h221 = at(i2, j2, k1), h122 = at(i1, j2, k2), h212 = at(i2, j1, k2), h222 = at(i2, j2, k2);
        const long x = cidx(F, i, j, k);
        double qn[5];
        for (int c = 0; c < 5; ++c) {
            const double *S = sd[c];
            const double t1 = S[h111] * a1;
            const double t2 = ((S[h112] + S[h121]) + S[h211]) * a2;
            const double t3 = ((S[h221] + S[h122]) + S[h212]) * a3;
            const double t4 = S[h222] * a4;
            const double v = ((t1 + t2) + t3) + t4;
            qn[c] = STATE ? v : F.Q[c * F.box + x] + v;
            F.Q[c * F.box + x] = qn[c];
        }
        prim_vars(F, x, qn, g, bad_flag + n);
    }
}

// N-level V-cycle, full approximation storage on the levels below the first
// coarse one (the reference has two levels at most, source/UserInput.cpp:
// 1057-1058, and its coarse level takes the restricted fine residual as its
// right-hand side, source/ComputeAllBlockVars.cpp:269-309).  On level l, after
// restriction (cv_2h = Q_l, Res_2h = the restricted total residual of level
// l-1) and a residual evaluation R(Q_l):
//   FORCE: P = Res_2h - R(Q_l), R = Res_2h  (the right-hand side of the sweep
//          is the restricted residual, as on the reference's coarse level, but
//          with lambda and dt of the current state);
//   ADD:   R = R(Q_l after the sweep) + P  (the total residual restricted to
//          level l+1).
// At a converged fine state every forcing cancels its residual and every
// correction Q_l - cv_2h vanishes, so the fixed point is the fine level's.
template <bool FORCE>
__global__ void __launch_bounds__(256) k_fas(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                             int nblock)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long x = cidx(B, i, j, k), box = B.box;
    for (int c = 0; c < 5; ++c) {
        const long y = c * box + x;
        if (FORCE) {
            B.fasP[y] = B.r2h[y] - B.R[y];
            B.R[y] = B.r2h[y];
        } else {
            B.R[y] = B.R[y] + B.fasP[y];
        }
    }
}

template __global__ void k_fas<true>(const BlockDev *, const long *, int);
template __global__ void k_fas<false>(const BlockDev *, const long *, int);

template __global__ void k_mg_prolong<true>(const BlockDev *, const BlockDev *, const int4 *, Gas, int *);
template __global__ void k_mg_prolong<false>(const BlockDev *, const BlockDev *, const int4 *, Gas, int *);

}  // namespace fv3d
