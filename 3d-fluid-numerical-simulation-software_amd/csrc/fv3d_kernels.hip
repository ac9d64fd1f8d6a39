// HIP kernels of the LU-SGS iteration for gfx950.
//
// One launch per stage covers every block of the mesh through a work list, so
// multi-block meshes cost no more launches than a single block.  Reductions
// (residual norm, section integrals) are two-pass with a fixed summation order
// and therefore deterministic run to run.
#include "fv3d_kernels.h"

#include <math.h>

namespace fv3d {

// ===================================================================== BC
// One thread per patch face.  Inlet (source/BndCondInlet.cpp:7-127), general
// and centrifugal outlet (source/BndCondOutlet.cpp:116-305), no-slip wall
// (source/BndCondWall.cpp:20-284).  Writes ghost cells of Q/W and the two
// Dummy layers of the patch side.
__device__ static void set_dummies(const BlockDev &B, int s, int a, int b, const double cv[5], const double pv[6])
{
    for (int l = 1; l <= kND; ++l) {
        for (int c = 0; c < 5; ++c) *dummy_ptr(B, s, l, c, a, b) = cv[c];
        for (int c = 0; c < 6; ++c) *dummy_ptr(B, s, l, 5 + c, a, b) = pv[c];
    }
}

// The pitch-angle sign of each periodic patch's block, updated in table order
// (source/BndCondInteriorPeriodic.cpp:54-59); one thread, so patches that share
// a block see each other's flips in order.
__device__ static void periodic_flip(PeriodicDev *__restrict__ per, int nper, int *__restrict__ pitch_sign)
{
    for (int p = 0; p < nper; ++p) {
        if (per[p].flip) pitch_sign[per[p].block] = -pitch_sign[per[p].block];
        per[p].sign = pitch_sign[per[p].block];
    }
}

// Thread 0 also updates the periodic signs (nper > 0) for the exchange that
// follows: no ghost fill here reads them, so this saves a launch.
__global__ void k_bc(const BlockDev *__restrict__ blocks, const PatchDev *__restrict__ patches, int npatch,
                     const long *__restrict__ pstart, Gas g, const ScalarsDev *__restrict__ sc,
                     PeriodicDev *__restrict__ per, int nper, int *__restrict__ pitch_sign)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && nper > 0) periodic_flip(per, nper, pitch_sign);
    if (t >= pstart[npatch]) return;
    const int q = find_range(pstart, npatch, t);
    const PatchDev P = patches[q];
    const BlockDev &B = blocks[P.block];
    long r = t - pstart[q];
    int a = P.a0 + (int)(r / P.nb), b = P.b0 + (int)(r % P.nb);   // tangential cell indices
    const int d = P.dir, hi = P.hi;
    const int N = d == 0 ? B.IM : d == 1 ? B.JM : B.KM;
    // interior cell next to the face and the ghost cell behind it
    int ii, jj, kk, gi, gj, gk;
    int pi = hi ? N - 1 : 1, pg = hi ? N : 0;
    if (d == 0) { ii = pi; gi = pg; jj = gj = a; kk = gk = b; }
    else if (d == 1) { jj = pi; gj = pg; ii = gi = a; kk = gk = b; }
    else { kk = pi; gk = pg; ii = gi = a; jj = gj = b; }
    const long xi = cidx(B, ii, jj, kk), xg = cidx(B, gi, gj, gk);
    const long box = B.box;
    const int s = 2 * d + hi;
    const double Gamma = g.gamma, Rgas = g.rgas;
    double cv[5], pv[6];
    if (P.kind == 0) {
        // inlet: K faces only
        if (d != 2) return;
        const int kf = hi ? N : 1;
        const long xf = cidx(B, ii, jj, kf);
        double Sx = B.fS[2][xf], Sy = B.fS[2][box + xf], Sz = B.fS[2][2 * box + xf];
        double n_m = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
        double uvel = B.W[1 * box + xi], vvel = B.W[2 * box + xi], wvel = B.W[3 * box + xi];
        double vdn = uvel * Sx / n_m + vvel * Sy / n_m + wvel * Sz / n_m;
        if (!hi) vdn = F32C(-1.0) * vdn;
        const double *A = g.inlet_dir;
        double Vn_m = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        double c_d = sqrt(Gamma * Rgas * B.W[5 * box + xi]);
        double V2_d = uvel * uvel + vvel * vvel + wvel * wvel;
        double costh = F32C(-1.0) * vdn / sqrt(V2_d);
        double c_02 = c_d * c_d + (Gamma - F32C(1.0)) * V2_d / F32C(2.0);
        double R_minus = vdn - F32C(2.0) * c_d / (Gamma - F32C(1.0));
        double meduA = (-R_minus * (Gamma - F32C(1.0))) / ((Gamma - F32C(1.0)) * costh * costh + F32C(2.0));
        double meduB = sqrt(c_02 * ((Gamma - F32C(1.0)) * costh * costh + F32C(2.0)) / (Gamma - F32C(1.0)) /
                                (R_minus * R_minus) - (Gamma - F32C(1.0)) / F32C(2.0));
        double c_b = meduA * (F32C(1.0) + costh * meduB);
        double Ttin = g.tt_in, Ptin = g.pt_in;
        double TempBC = Ttin * (c_b * c_b / c_02);
        if (TempBC > Ttin) TempBC = F32C(0.99) * Ttin;
        double PressBC = Ptin * pow(TempBC / Ttin, Gamma / (Gamma - F32C(1.0)));
        double DensBC = PressBC / Rgas / TempBC;
        double VBC = sqrt(F32C(2.0) * Gamma / (Gamma - 1) * Rgas * (Ttin - TempBC));
        double u = VBC * A[0] / Vn_m, v = VBC * A[1] / Vn_m, w = VBC * A[2] / Vn_m;
        double V2 = u * u + v * v + w * w;
        double ener = PressBC / DensBC / (Gamma - F32C(1.0)) + V2 / F32C(2.0);
        pv[0] = DensBC; pv[1] = u; pv[2] = v; pv[3] = w; pv[4] = PressBC; pv[5] = TempBC;
        cv[0] = DensBC; cv[1] = DensBC * u; cv[2] = DensBC * v; cv[3] = DensBC * w; cv[4] = DensBC * ener;
        if (B.ev) B.ev[xg] = g.ev_ratio * F32C(1.0);   // source/BndCondInlet.cpp:104-105
    } else if (P.kind == 1) {
        // general / centrifugal outlet, K faces only (the radial-equilibrium
        // outlet is k_bc_axial's)
        if (d != 2 || P.axial) return;
        const int kf = hi ? N : 1;
        const long xf = cidx(B, ii, jj, kf);
        double s_ = B.fs[2][xf];
        double nx = B.fS[2][xf] / s_, ny = B.fS[2][box + xf] / s_, nz = B.fS[2][2 * box + xf] / s_;
        double d0 = B.W[xi], u0 = B.W[box + xi], v0 = B.W[2 * box + xi], w0 = B.W[3 * box + xi];
        double p0 = B.W[4 * box + xi], T0 = B.W[5 * box + xi];
        double Pout = sc[P.inst].p_out;
        double dens0 = d0;
        double c0 = sqrt(Gamma * Rgas * T0);
        double Vn = u0 * nx + v0 * ny + w0 * nz;
        double Ma_n = Vn / c0;
        if (P.centrifugal || Ma_n <= 1.0) {
            pv[4] = Pout;
            pv[0] = d0 + (Pout - p0) / c0 / c0;
            pv[1] = u0 + nx * (p0 - Pout) / dens0 / c0;
            pv[2] = v0 + ny * (p0 - Pout) / dens0 / c0;
            pv[3] = w0 + nz * (p0 - Pout) / dens0 / c0;
        } else {
            pv[4] = p0; pv[0] = d0; pv[1] = u0; pv[2] = v0; pv[3] = w0;
        }
        pv[5] = pv[4] / pv[0] / Rgas;
        double V2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
        double ener = pv[4] / pv[0] / (Gamma - F32C(1.0)) + V2 / F32C(2.0);
        cv[0] = pv[0]; cv[1] = pv[0] * pv[1]; cv[2] = pv[0] * pv[2]; cv[3] = pv[0] * pv[3]; cv[4] = pv[0] * ener;
        if (B.ev) B.ev[xg] = B.ev[xi];   // source/BndCondOutlet.cpp:180-181, 282-283
    } else {
        // no-slip wall: mirror the relative velocity about the wall velocity
        const int f = hi ? N : 1;
        const long xf = d == 0 ? cidx(B, f, jj, kk) : d == 1 ? cidx(B, ii, f, kk) : cidx(B, ii, jj, f);
        double x = B.fc[d][xf], y = B.fc[d][box + xf];
        double rot = P.rot;
        double Vwall_x = -rot * y, Vwall_y = rot * x;
        double Vin_x = B.W[box + xi] - Vwall_x, Vin_y = B.W[2 * box + xi] - Vwall_y, Vin_z = B.W[3 * box + xi] - F32C(0.0);
        double Vout_x = F32C(0.0) - Vin_x, Vout_y = F32C(0.0) - Vin_y, Vout_z = F32C(0.0) - Vin_z;
        pv[0] = B.W[xi];
        pv[1] = Vout_x + Vwall_x;
        pv[2] = Vout_y + Vwall_y;
        pv[3] = Vout_z + F32C(0.0);
        pv[4] = B.W[4 * box + xi];
        pv[5] = B.W[5 * box + xi];
        cv[0] = B.Q[xi];
        cv[1] = pv[1] * cv[0];
        cv[2] = pv[2] * cv[0];
        cv[3] = pv[3] * cv[0];
        cv[4] = B.Q[4 * box + xi];
        for (int c = 0; c < 5; ++c) B.Q[c * box + xg] = cv[c];
        for (int c = 0; c < 6; ++c) B.W[c * box + xg] = pv[c];
        if (B.ev) B.ev[xg] = F32C(0.0);   // source/BndCondWall.cpp:104-107
        double cd[5], pd[6];
        for (int c = 0; c < 5; ++c) cd[c] = (cv[c] + B.Q[c * box + xi]) * 0.5;
        for (int c = 0; c < 6; ++c) pd[c] = 0.5 * (pv[c] + B.W[c * box + xi]);
        set_dummies(B, s, a, b, cd, pd);
        return;
    }
    for (int c = 0; c < 5; ++c) B.Q[c * box + xg] = cv[c];
    for (int c = 0; c < 6; ++c) B.W[c * box + xg] = pv[c];
    set_dummies(B, s, a, b, cv, pv);
}

// CBndConds::BndCondOutlet_Axial (source/BndCondOutlet.cpp:7-114): static
// pressure in radial equilibrium, dp/dr = rho c_u^2 / r, fixed to Pout at
// i = (int)(0.5 IM) and integrated outwards along i on every j line; c_u is the
// area-weighted circumferential velocity of the last cell layer per i.  As in
// the reference it covers the whole K+ face of the block, reads the density of
// the ghost layer left by the previous application, and integrates towards
// the tip with the radii of the faces at KM-1.  One workgroup per patch: the
// per-i averages (serial over j, the reference's order), then one thread per
// j line, then the ghost cells and Dummy layers.
__global__ void __launch_bounds__(256) k_bc_axial(const BlockDev *__restrict__ blocks,
                                                  const PatchDev *__restrict__ patches, const int *__restrict__ list,
                                                  Gas g, const ScalarsDev *__restrict__ sc)
{
    __shared__ double cu[kAxialMaxI];
    const PatchDev P = patches[list[blockIdx.x]];
    if (P.dir != 2 || !P.hi) return;   // K+ faces only (source/BndCondOutlet.cpp:33-41)
    const BlockDev &B = blocks[P.block];
    const int IM = B.IM, JM = B.JM, KM = B.KM, k = KM - 1;
    const long box = B.box;
    const double *fc = B.fc[2], *fs = B.fs[2];
    const double Gamma = g.gamma, Rgas = g.rgas;
    for (int i = 1 + (int)threadIdx.x; i <= IM - 1; i += blockDim.x) {
        double c_u = 0.0, S_temp = 0.0;
        for (int j = 1; j <= JM - 1; ++j) {
            const long xc = cidx(B, i, j, k), xf = cidx(B, i, j, KM);
            const double th = atan(fc[box + xf] / fc[xf]);
            const double cu_temp = F32C(-1.0) * B.W[box + xc] * sin(th) + B.W[2 * box + xc] * cos(th);
            c_u = c_u + cu_temp * fs[xf];
            S_temp = S_temp + fs[xf];
        }
        cu[i] = c_u / S_temp;
    }
    __syncthreads();
    const int iref = (int)(0.5 * IM);
    const double Pout = sc[P.inst].p_out;
    auto rad = [&](int i, int j, int kk) {
        const long xf = cidx(B, i, j, kk);
        return sqrt(fc[xf] * fc[xf] + fc[box + xf] * fc[box + xf]);
    };
    double *p = B.W + 4 * box;
    for (int j = 1 + (int)threadIdx.x; j <= JM - 1; j += blockDim.x) {
        p[cidx(B, iref, j, KM)] = Pout;
        for (int i = iref - 1; i >= 1; --i) {
            const double rn = rad(i, j, KM), rn_1 = rad(i + 1, j, KM);
            const double dp_dr = B.Q[cidx(B, i, j, KM)] * cu[i] * cu[i] / rn;
            p[cidx(B, i, j, KM)] = p[cidx(B, i + 1, j, KM)] - dp_dr * (rn_1 - rn);
        }
        for (int i = iref + 1; i <= IM - 1; ++i) {
            const double rn_1 = rad(i, j, k), rn = rad(i - 1, j, k);
            const double dp_dr = B.Q[cidx(B, i, j, KM)] * cu[i] * cu[i] / rn;
            p[cidx(B, i, j, KM)] = p[cidx(B, i - 1, j, KM)] + dp_dr * (rn_1 - rn);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (IM - 1) * (JM - 1); e += blockDim.x) {
        const int i = 1 + e / (JM - 1), j = 1 + e % (JM - 1);
        const long xg = cidx(B, i, j, KM), xi = cidx(B, i, j, k);
        double cv[5], pv[6];
        for (int c = 0; c < 4; ++c) cv[c] = B.Q[c * box + xi];
        pv[0] = cv[0];
        pv[1] = B.W[box + xi]; pv[2] = B.W[2 * box + xi]; pv[3] = B.W[3 * box + xi];
        pv[4] = p[xg];
        pv[5] = pv[4] / Rgas / cv[0];
        const double V2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
        cv[4] = pv[4] / (Gamma - F32C(1.0)) + pv[0] * V2 / F32C(2.0);
        for (int c = 0; c < 5; ++c) B.Q[c * box + xg] = cv[c];
        for (int c = 0; c < 6; ++c) B.W[c * box + xg] = pv[c];
        if (B.ev) B.ev[xg] = B.ev[xi];   // source/BndCondOutlet.cpp:102-103
        set_dummies(B, 5, i, j, cv, pv);
    }
}

// Global time step of CTimeDiscr::TimeStep (source/TimeStep.cpp:64-86): the
// minimum over each block's cells, starting from 1E10 (the caller fills dtmin
// with its bits), then every cell of the block takes it and the LU-SGS
// diagonal is formed again from it.  Time steps are positive, so the bits
// order as the values.
__global__ void __launch_bounds__(256) k_dt_block_min(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                      int nblock, unsigned long long *__restrict__ dtmin)
{
    constexpr int kSlots = 16;
    __shared__ unsigned long long smin[kSlots];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long t0 = (long)blockIdx.x * blockDim.x;
    const int n0 = t0 < cstart[nblock] ? find_range(cstart, nblock, t0) : nblock - 1;
    if (threadIdx.x < kSlots) smin[threadIdx.x] = ~0ull;
    __syncthreads();
    int n, i, j, k;
    if (cell_item(blocks, cstart, nblock, t, n, i, j, k)) {
        const BlockDev &B = blocks[n];
        const unsigned long long v = (unsigned long long)__double_as_longlong(B.dt[cidx(B, i, j, k)]);
        if (n - n0 < kSlots) atomicMin(&smin[n - n0], v);
        else atomicMin(dtmin + n, v);
    }
    __syncthreads();
    if (threadIdx.x < kSlots && n0 + (int)threadIdx.x < nblock && smin[threadIdx.x] != ~0ull)
        atomicMin(dtmin + n0 + threadIdx.x, smin[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_dt_global(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                   int nblock, Gas g, const unsigned long long *__restrict__ dtmin)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long box = B.box, x = cidx(B, i, j, k);
    const double dt = __longlong_as_double((long long)dtmin[n]);
    B.dt[x] = dt;
    double lvd[3] = {0.0, 0.0, 0.0};
    if (g.viscous)
        for (int d = 0; d < 3; ++d) lvd[d] = B.lamv[d * box + x];
    B.alpha[x] = B.vol[x] / dt + g.relax * (B.lamc[x] + B.lamc[box + x] + B.lamc[2 * box + x]) +
                 F32C(2.0) * (lvd[0] + lvd[1] + lvd[2]);
    if (B.alpha2)
        B.alpha2[x] = B.vol[x] / dt + g.relax * (B.lamc[x] + B.lamc[box + x] + B.lamc[2 * box + x]) +
                      F32C(3.0) * (lvd[0] + lvd[1] + lvd[2]);
}

// ================================================== interior / periodic
__global__ void k_periodic_flip(PeriodicDev *__restrict__ per, int nper, int *__restrict__ pitch_sign)
{
    if (threadIdx.x != 0) return;
    periodic_flip(per, nper, pitch_sign);
}

// One transfer of source/ShadowOutIn.cpp: Trans rotates the momentum and the
// velocity about z by dr when dr != 0 (lines 188-199).
__global__ void k_exchange(const BlockDev *__restrict__ blocks, const ExchDev *__restrict__ el, long n,
                           const PeriodicDev *__restrict__ per)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExchDev e = el[t];
    const BlockDev &A = blocks[e.src_block], &B = blocks[e.dst_block];
    double q[5], w[6];
    for (int c = 0; c < 5; ++c) q[c] = A.Q[c * A.box + e.src];
    for (int c = 0; c < 6; ++c) w[c] = A.W[c * A.box + e.src];
    if (e.rot) {
        const PeriodicDev &P = per[e.per];
        const bool pos = (e.rot == 1) == (P.sign > 0);   // dr = +delta or -delta
        const double cr = pos ? P.cs[0] : P.cs[2], sr = pos ? P.cs[1] : P.cs[3];
        const double qx = q[1], qy = q[2], ux = w[1], uy = w[2];
        q[1] = qx * cr - qy * sr;
        q[2] = qy * cr + qx * sr;
        w[1] = ux * cr - uy * sr;
        w[2] = uy * cr + ux * sr;
    }
    double *D = B.D[e.side] + (long)(e.layer - 1) * 11 * B.Dt[e.side] + e.slot;
    for (int c = 0; c < 5; ++c) D[c * B.Dt[e.side]] = q[c];
    for (int c = 0; c < 6; ++c) D[(5 + c) * B.Dt[e.side]] = w[c];
    if (e.ghost >= 0) {
        for (int c = 0; c < 5; ++c) B.Q[c * B.box + e.ghost] = q[c];
        for (int c = 0; c < 6; ++c) B.W[c * B.box + e.ghost] = w[c];
        if (B.ev) B.ev[e.ghost] = A.ev[e.src];   // nu~ of the ghost layer only (source/ShadowOutIn.cpp:339-345)
    }
}

// ======================================================= spectral + dt
// source/Spectral_curl_grad.cpp:9-42 and source/TimeStep.cpp:7-62, with the
// viscous spectral radii for Navier-Stokes (TimeStep.cpp:38-49).
__global__ void k_spectral_dt(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart, int nblock, Gas g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for_cell_item(blocks, cstart, nblock, t, [&](const BlockDev &B, int, int i, int j, int k) {
        const long box = B.box, x = cidx(B, i, j, k);
        const double Gamma = g.gamma, Rgas = g.rgas;
        double uvelr = B.W[box + x] + B.w_rot * B.cc[box + x];
        double vvelr = B.W[2 * box + x] - B.w_rot * B.cc[x];
        double wvelr = B.W[3 * box + x];
        double csoun = sqrt(Gamma * Rgas * B.W[5 * box + x]);
        double lam[3], eq[3];
        const long xn[3] = {x + B.si, x + B.sj, x + 1};
        for (int d = 0; d < 3; ++d) {
            const double *S = B.fS[d];
            double Sx = (S[x] + S[xn[d]]) * F32C(0.5);
            double Sy = (S[box + x] + S[box + xn[d]]) * F32C(0.5);
            double Sz = (S[2 * box + x] + S[2 * box + xn[d]]) * F32C(0.5);
            eq[d] = Sx * Sx + Sy * Sy + Sz * Sz;
            double CS = csoun * sqrt(eq[d]);
            lam[d] = fabs(uvelr * Sx + vvelr * Sy + wvelr * Sz) + CS;
            B.lamc[d * box + x] = lam[d];
        }
        int flag = *B.flag_low_dt;
        double dt_factor = flag == 0 ? F32C(1.0) : (double)(1.0f / (1.0f + (float)flag));
        double lc = lam[0] + lam[1] + lam[2];
        double lv = 0.0, lvd[3] = {0.0, 0.0, 0.0};
        if (g.viscous) {
            const double mue_lam = sutherland(g, B.W[5 * box + x]), mue_tur = B.mut ? B.mut[x] : 0.0, dens = B.W[x];
            const double a = (double)(4.0f / 3.0f) / dens, b = Gamma / dens;
            const double m = a > b ? a : b;
            for (int d = 0; d < 3; ++d) {
                lvd[d] = m * (mue_lam / g.pr_lam + mue_tur / g.pr_tur) * eq[d] / B.vol[x];
                B.lamv[d * box + x] = lvd[d];
            }
            lv = lvd[0] + lvd[1] + lvd[2];
        }
        double dt = dt_factor * (*B.cfl) * B.vol[x] / (lc + g.c_timestep * lv);
        B.dt[x] = dt;
        // diagonal of the LU-SGS operator (source/LU_SGS_Marching.cpp:76-77)
        B.alpha[x] = B.vol[x] / dt + g.relax * (lam[0] + lam[1] + lam[2]) + F32C(2.0) * (lvd[0] + lvd[1] + lvd[2]);
        // and of its nu~ equation (source/LU_SGS_Marching.cpp:81-82)
        if (B.alpha2)
            B.alpha2[x] = B.vol[x] / dt + g.relax * (lam[0] + lam[1] + lam[2]) + F32C(3.0) * (lvd[0] + lvd[1] + lvd[2]);
    });
}

// Decrement of the time-step damping counter at the end of CTimeDiscr::TimeStep
// (source/TimeStep.cpp:89-92), after every cell has read it.
__global__ void k_dt_flag_decrement(const BlockDev *__restrict__ blocks, int nblock)
{
    for (int n = threadIdx.x; n < nblock; n += blockDim.x) {
        int *f = blocks[n].flag_low_dt;
        if (*f > 0) *f = *f - 1;
    }
}

// =================================================== convective residual
// Face flux of CSpaceDiscr::Upwind_Convective (source/Upwind_Convective.cpp:5-360)
// with the WallFlux override (source/WallFlux.cpp:70-175).
__device__ static void face_flux(const BlockDev &B, const Gas &g, int d, int i, int j, int k, double out[5])
{
    const long box = B.box;
    const int N = d == 0 ? B.IM : d == 1 ? B.JM : B.KM;
    const int p = d == 0 ? i : d == 1 ? j : k;
    const long xf = cidx(B, i, j, k);
    if (p == 1 || p == N) {
        const int s = 2 * d + (p == N);
        const int ta = d == 0 ? j : i, tb = d == 2 ? j : k;
        const long tix = (long)ta * B.Db[s] + tb;
        if (B.wmask[s] != nullptr && B.wmask[s][tix]) {
            const long xb = xf - (d == 0 ? B.si : d == 1 ? B.sj : 1);
            double wa[5], wb[5];
            for (int c = 0; c < 5; ++c) { wa[c] = B.W[c * box + xf]; wb[c] = B.W[c * box + xb]; }
            wall_face_core(g.gamma, wa, wb, B.fS[d][xf], B.fS[d][box + xf], B.fS[d][2 * box + xf], B.fc[d][xf],
                           B.fc[d][box + xf], B.wrot[s][tix], out);
            return;
        }
    }
    // four-cell stencil along d; cells outside [1, N-1] come from the Dummy
    // layers.  Characteristic reconstruction (any variable code other than
    // 100 and 200, source/UserInput.cpp:117-130) takes the conservative
    // stencil and the primitive variables of F and B.
    const bool prim = g.recon_var == 200, chr = g.recon_var != 100 && g.recon_var != 200;
    double st[4][5], wFB[2][6];
    const long dstride = d == 0 ? B.si : d == 1 ? B.sj : 1;
    const double *src = prim ? B.W : B.Q;
    for (int o = 0; o < 4; ++o) {
        int pc = p - 2 + o;
        if (pc >= 1 && pc <= N - 1) {
            long xc = xf + (long)(o - 2) * dstride;
            for (int c = 0; c < 5; ++c) st[o][c] = src[c * box + xc];
            if (chr && (o == 1 || o == 2))
                for (int c = 0; c < 6; ++c) wFB[o - 1][c] = B.W[c * box + xc];
        } else {
            int s = 2 * d + (pc >= N ? 1 : 0);
            int layer = pc < 1 ? 1 - pc : pc - (N - 1);
            int ta = d == 0 ? j : i, tb = d == 2 ? j : k;
            int c0 = prim ? 5 : 0;
            for (int c = 0; c < 5; ++c) st[o][c] = *dummy_ptr(B, s, layer, c0 + c, ta, tb);
            if (chr && (o == 1 || o == 2))
                for (int c = 0; c < 6; ++c) wFB[o - 1][c] = *dummy_ptr(B, s, layer, 5 + c, ta, tb);
        }
    }
    const long xd = d == 0 ? cidx(B, i, j + 1, k + 1) : d == 1 ? cidx(B, i + 1, j, k + 1) : cidx(B, i + 1, j + 1, k);
    // a non-rotating block skips the face centroid (its grid velocity is zero)
    double fcx = 0.0, fcy = 0.0;
    if (B.w_rot != 0.0) { fcx = B.fc[d][xf]; fcy = B.fc[d][box + xf]; }
    face_core(g, st, wFB[0], wFB[1], B.fS[d][xf], B.fS[d][box + xf], B.fS[d][2 * box + xf], B.fs[d][xf], fcx, fcy,
              B.w_rot, B.node[xd] - B.node[xf], B.node[box + xd] - B.node[box + xf],
              B.node[2 * box + xd] - B.node[2 * box + xf], out);
}

// Tile of TI x TJ x TK cells per workgroup; every face of the tile is computed
// once into LDS, then each cell sums its six faces and subtracts the rotation
// source (source/ComputeAllBlockVars.cpp:85-118, 237-254).  Optionally stores
// the face fluxes for inspection.
// Three waves per SIMD: the register budget (168) holds the face code with a
// few spills, and the third wave hides more latency than the spills cost.
template <int TI, int TJ, int TK>
__global__ void __launch_bounds__(TI *TJ *TK) __attribute__((amdgpu_waves_per_eu(3))) k_convective(const BlockDev *__restrict__ blocks, const int4 *__restrict__ tiles,
                                                           Gas g, int store_faces, double *fI0, double *fJ0, double *fK0)
{
    __shared__ double sI[5][TI + 1][TJ][TK];
    __shared__ double sJ[5][TI][TJ + 1][TK];
    __shared__ double sK[5][TI][TJ][TK + 1];
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;   // padding of the XCD-ordered tile list
    const BlockDev &B = blocks[tl.x];
    const int t = threadIdx.x;
    const int tk = t % TK, tj = (t / TK) % TJ, ti = t / (TK * TJ);
    const int i = tl.y + ti + 1, j = tl.z + tj + 1, k = tl.w + tk + 1;
    const bool valid = i <= B.ni && j <= B.nj && k <= B.nk;
    double f[5];
    if (valid) {
        // the cell's three low faces, then the high faces on the tile's upper
        // edges, all through one copy of the face code: the loop stays rolled
        // because six inlined copies of it (every scheme and reconstruction)
        // would not fit the instruction cache
#pragma nounroll
        for (int e = 0; e < 6; ++e) {
            const int d = e < 3 ? e : e - 3, hi = e < 3 ? 0 : 1;
            if (hi && !(d == 0 ? (ti == TI - 1 || i == B.ni) : d == 1 ? (tj == TJ - 1 || j == B.nj)
                                                                    : (tk == TK - 1 || k == B.nk)))
                continue;
            face_flux(B, g, d, i + (d == 0 ? hi : 0), j + (d == 1 ? hi : 0), k + (d == 2 ? hi : 0), f);
            if (d == 0)
                for (int c = 0; c < 5; ++c) sI[c][ti + hi][tj][tk] = f[c];
            else if (d == 1)
                for (int c = 0; c < 5; ++c) sJ[c][ti][tj + hi][tk] = f[c];
            else
                for (int c = 0; c < 5; ++c) sK[c][ti][tj][tk + hi] = f[c];
        }
    }
    __syncthreads();
    if (!valid) return;
    const long box = B.box, x = cidx(B, i, j, k);
    const double w_rot = B.w_rot;
    const double vol = B.vol[x];
    const double dens = B.W[x], uvel = B.W[box + x], vvel = B.W[2 * box + x];
    double Qc[5] = {0.0, dens * (w_rot * vvel) * vol, F32C(-1.0) * dens * (w_rot * uvel) * vol, 0.0, 0.0};
    for (int c = 0; c < 5; ++c) {
        double s = sI[c][ti + 1][tj][tk] - sI[c][ti][tj][tk];
        s = s + sJ[c][ti][tj + 1][tk];
        s = s - sJ[c][ti][tj][tk];
        s = s + sK[c][ti][tj][tk + 1];
        s = s - sK[c][ti][tj][tk];
        s = s - Qc[c];
        // Res = Fc - Fv - Fd with Fv = Fd = 0 for the inviscid upwind path
        B.R[c * box + x] = (s - 0.0) - 0.0;
    }
    if (store_faces) {
        for (int c = 0; c < 5; ++c) {
            fI0[c * box + x] = sI[c][ti][tj][tk];
            fJ0[c * box + x] = sJ[c][ti][tj][tk];
            fK0[c * box + x] = sK[c][ti][tj][tk];
            if (ti == TI - 1 || i == B.ni) fI0[c * box + x + B.si] = sI[c][ti + 1][tj][tk];
            if (tj == TJ - 1 || j == B.nj) fJ0[c * box + x + B.sj] = sJ[c][ti][tj + 1][tk];
            if (tk == TK - 1 || k == B.nk) fK0[c * box + x + 1] = sK[c][ti][tj][tk + 1];
        }
    }
}

#define FV3D_CONV(S, A, B, C) \
    template __global__ void k_convective<A, B, C>(const BlockDev *, const int4 *, Gas, int, double *, double *, double *);
FV3D_FOR_TILES(FV3D_CONV)
#undef FV3D_CONV

// Face-parallel form of the same residual: every face of every block once, one
// thread per face (faces of a block and direction i-slowest, k fastest), into
// B.fcv; k_conv_sum then forms each cell's residual from its six faces and the
// rotation source in k_convective's order, so the two forms agree bit for bit.
__global__ void __launch_bounds__(256) k_conv_faces(const BlockDev *__restrict__ blocks, const long *__restrict__ fstart,
                                                    int nlist, Gas g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fstart[nlist]) return;
    const int q = find_range(fstart, nlist, t);
    const int d = q % 3;
    const BlockDev &B = blocks[q / 3];
    const long r = t - fstart[q];
    const int nj = B.nj + (d == 1 ? 1 : 0), nk = B.nk + (d == 2 ? 1 : 0);
    const int i = (int)(r / ((long)nj * nk)) + 1, j = (int)(r / nk % nj) + 1, k = (int)(r % nk) + 1;
    double f[5];
    face_flux(B, g, d, i, j, k, f);
    const long x = cidx(B, i, j, k);
    for (int c = 0; c < 5; ++c) B.fcv[(long)(d * 5 + c) * B.box + x] = f[c];
}

__global__ void __launch_bounds__(256) k_conv_sum(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                  int nblock)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long box = B.box, x = cidx(B, i, j, k);
    const double w_rot = B.w_rot;
    const double vol = B.vol[x];
    const double dens = B.W[x], uvel = B.W[box + x], vvel = B.W[2 * box + x];
    double Qc[5] = {0.0, dens * (w_rot * vvel) * vol, F32C(-1.0) * dens * (w_rot * uvel) * vol, 0.0, 0.0};
    for (int c = 0; c < 5; ++c) {
        const double *FI = B.fcv + (long)c * box, *FJ = B.fcv + (long)(5 + c) * box, *FK = B.fcv + (long)(10 + c) * box;
        double s = FI[x + B.si] - FI[x];
        s = s + FJ[x + B.sj];
        s = s - FJ[x];
        s = s + FK[x + 1];
        s = s - FK[x];
        s = s - Qc[c];
        B.R[c * box + x] = (s - 0.0) - 0.0;
    }
}

// ================================================================ LU-SGS
// Flux difference of CTimeDiscr::delt_Fn_up / delt_Fn_down
// (source/LU_SGS_Marching.cpp:293-419) at neighbour cell nb whose faces along
// d are nb (lower) and nb + stride (upper).
__device__ static void delt_fn(const BlockDev &B, double Gamma, int d, long nb, double df[5])
{
    const long box = B.box;
    const long st = d == 0 ? B.si : d == 1 ? B.sj : 1;
    const long fa = nb + st, fb = nb;   // upper face first, as the reference sums them
    const double *S = B.fS[d], *s = B.fs[d];
    double nx = (S[fa] / s[fa] + S[fb] / s[fb]) / F32C(2.0);
    double ny = (S[box + fa] / s[fa] + S[box + fb] / s[fb]) / F32C(2.0);
    double nz = (S[2 * box + fa] / s[fa] + S[2 * box + fb] / s[fb]) / F32C(2.0);
    double uvel_e = F32C(-1.0) * B.cc[box + nb] * B.w_rot;
    double vvel_e = B.cc[nb] * B.w_rot;
    double o[5], nw[5];
    for (int c = 0; c < 5; ++c) { o[c] = B.Q[c * box + nb]; nw[c] = o[c] + B.dQ[c * box + nb]; }
    double dens = nw[0], uvel = nw[1] / nw[0], vvel = nw[2] / nw[0], wvel = nw[3] / nw[0], E = nw[4] / nw[0];
    double press_new = (Gamma - F32C(1.0)) * dens * (E - F32C(0.5) * (uvel * uvel + vvel * vvel + wvel * wvel));
    double press_old = B.W[4 * box + nb];
    double Vn_old = o[1] / o[0] * nx + o[2] / o[0] * ny + o[3] / o[0] * nz;
    double Vn_new = uvel * nx + vvel * ny + wvel * nz;
    double Vn_e = uvel_e * nx + vvel_e * ny + F32C(0.0) * nz;
    df[0] = nw[0] * (Vn_new - Vn_e) - o[0] * (Vn_old - Vn_e);
    df[1] = (nw[1] * (Vn_new - Vn_e) + press_new * nx) - (o[1] * (Vn_old - Vn_e) + press_old * nx);
    df[2] = (nw[2] * (Vn_new - Vn_e) + press_new * ny) - (o[2] * (Vn_old - Vn_e) + press_old * ny);
    df[3] = (nw[3] * (Vn_new - Vn_e) + press_new * nz) - (o[3] * (Vn_old - Vn_e) + press_old * nz);
    df[4] = ((nw[4] + press_new) * (Vn_new - Vn_e) + press_new * Vn_e) - ((o[4] + press_old) * (Vn_old - Vn_e) + press_old * Vn_e);
}

// One hyperplane i+j+k = plane of every block.  Thread (n, i, j) handles cell
// (i, j, plane-i-j) of block n.  Cells of one plane depend only on cells of the
// previous plane, so the result equals the reference's serial order exactly.
template <bool FORWARD>
__global__ void k_lusgs_plane(const BlockDev *__restrict__ blocks, const long *__restrict__ pstart, int nblock, int step,
                              double relax, Gas g)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pstart[nblock]) return;
    const int n = find_range(pstart, nblock, t);
    const BlockDev &B = blocks[n];
    const long r = t - pstart[n];
    const int i = (int)(r / B.nj) + 1, j = (int)(r % B.nj) + 1;
    const int last = B.IM + B.JM + B.KM - 3;
    const int plane = FORWARD ? 3 + step : last - step;
    if (plane < 3 || plane > last) return;
    const int k = plane - i - j;
    if (k < 1 || k > B.nk) return;
    const long box = B.box, x = cidx(B, i, j, k);
    const double Gamma = g.gamma;
    double alpha = B.alpha[x];   // Vol/dt + relax*sum(lamc) + 2*sum(lamv), from k_spectral_dt
    double A[3][5], V[3][5];
    const long st[3] = {B.si, B.sj, 1};
    const int pos[3] = {i, j, k};
    const int lim[3] = {B.ni, B.nj, B.nk};
    for (int d = 0; d < 3; ++d) {
        bool ok = FORWARD ? pos[d] != 1 : pos[d] != lim[d];
        if (!ok) {
            for (int c = 0; c < 5; ++c) { A[d][c] = 0.0; V[d][c] = 0.0; }
            continue;
        }
        const long nb = FORWARD ? x - st[d] : x + st[d];
        double df[5];
        delt_fn(B, Gamma, d, nb, df);
        const double *s = B.fs[d];
        double sd = F32C(0.5) * (s[nb] + s[nb + st[d]]);
        double rl = relax * B.lamc[d * box + nb];
        // lamv * dQ (lamv = 0 on the inviscid path)
        double lv = B.lamv ? B.lamv[d * box + nb] : 0.0;
        for (int c = 0; c < 5; ++c) {
            double dq = B.dQ[c * box + nb];
            A[d][c] = FORWARD ? (df[c] * sd + dq * rl) * F32C(0.5) : (df[c] * sd - dq * rl) * F32C(0.5);
            V[d][c] = dq * lv;
        }
    }
    for (int c = 0; c < 5; ++c) {
        double s;
        if (FORWARD) {
            s = B.R[c * box + x] * F32C(-1.0) + A[0][c];
            s = s + A[1][c]; s = s + A[2][c];
        } else {
            s = B.dQ[c * box + x] * alpha - A[0][c];
            s = s - A[1][c]; s = s - A[2][c];
        }
        s = s + V[0][c]; s = s + V[1][c]; s = s + V[2][c];
        B.dQ[c * box + x] = s / alpha;
    }
}

template __global__ void k_lusgs_plane<true>(const BlockDev *, const long *, int, int, double, Gas);
template __global__ void k_lusgs_plane<false>(const BlockDev *, const long *, int, int, double, Gas);

// the row-layout sweep in the arithmetic type T
template <class T>
using gcT = const T __attribute__((address_space(1))) *;
template <class T>
using gT = T __attribute__((address_space(1))) *;
#define TC(x) ((T)(x##f))

namespace {
// row-layout sweep: per step, the neighbour's static data with the terms of
// the flux difference that do not depend on dQ, the upstream dQ, and (wave 0)
// the cell's right-hand side and diagonal
template <class T>
struct RowPf {
    T q[5], m[5], of[5], vn_old, cl;   // state, metrics, old-state flux terms
    T p, l, lv;
    T x[kSweepD];                      // upstream tile's dQ (and dnu~)
    T r[5], a;
    T ev, rev, a2;                     // Spalart-Allmaras: neighbour's nu~; own nu~ right-hand side and diagonal
};
template <int D>
struct Dir {
    static constexpr int value = D;
};
// A direction wave issues the static loads of a step (the neighbour's state and
// metrics, its own right-hand side) kStaticAhead steps before the step and
// forms their early terms one step before it, so a load has kStaticAhead - 1
// steps to arrive; it cycles through kRing row buffers.
#ifndef FV3D_STATIC_AHEAD
#define FV3D_STATIC_AHEAD 3
#endif
constexpr int kStaticAhead = FV3D_STATIC_AHEAD;
constexpr int kRing = kStaticAhead + 1;
// The upstream tile's dQ for a step is read kCrossAhead steps before the step
// (the word may still be empty then; the step re-reads it until it is not).
// Reading it earlier than one step ahead finds it empty more often than it
// hides latency: measured on the 2.54M-cell vane passage, two steps ahead
// gained 3-7 % on the fine FP32 sweeps and lost as much on FP64 and coarse ones.
#ifndef FV3D_CROSS_AHEAD
#define FV3D_CROSS_AHEAD 1
#endif
constexpr int kCrossAhead = FV3D_CROSS_AHEAD;
static_assert(kCrossAhead >= 1 && kCrossAhead <= kStaticAhead, "the cross read lands in a free row buffer");
constexpr int kRowS0 = -kStaticAhead;   // first (idle) step of the row-layout sweep loops
// 1/x for the FP32 operator: the hardware reciprocal (within 1 ulp), one
// instruction where a correctly rounded division takes about ten; the FP32
// operator only preconditions the FP64 residual, whose fixed point it keeps
__device__ inline float rcp_op(float x) { return __builtin_amdgcn_rcpf(x); }
// whether a mailbox word still holds the empty pattern
__device__ inline bool mbox_empty(double x) { return __builtin_bit_cast(unsigned long long, x) == kMboxEmpty; }
__device__ inline bool mbox_empty(float x) { return __builtin_bit_cast(unsigned int, x) == kMboxEmpty32; }
// the mailbox set of a sweep direction, in words of T (both sets share one
// allocation whose second half starts at mbox[1])
template <class T>
__device__ inline T *mbox_base(const BlockDev &B, bool fwd)
{
    return (T *)B.mbox[0] + (fwd ? 0 : (B.mbox[1] - B.mbox[0]));
}
}  // namespace

// ---------------------------------------------------------------------------
// Tile-skewed sweep layout (BlockDev::srow / sdq).
//
// Lane (tj,tk) of row r of tile (J,K) is cell (r+1-tj-tk, J*TJ+tj+1, K*TK+tk+1).
// k_sweep_fill writes whole rows, one wave per row, giving lanes outside the
// block a harmless unit state; it runs at set-up (metrics and those lanes).
// Each iteration k_sweep_pack copies the state of the cells in, and after the
// sweeps k_sweep_unpack copies dQ back to the cell box.
template <class T>
__global__ void __launch_bounds__(256) k_sweep_fill(const BlockDev *__restrict__ blocks, int block, long nrows, int what)
{
    constexpr int TJ = kKbaTJ, TK = kKbaTK;
    const BlockDev &B = blocks[block];
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    const int lane = threadIdx.x & 63, tj = lane / TK, tk = lane % TK;
    const int slot = (int)(row / B.nr), r = (int)(row % B.nr);
    const int J = slot / B.ntk, K = slot % B.ntk;
    const int i = r + 1 - tj - tk, j = J * TJ + tj + 1, k = K * TK + tk + 1;
    const bool valid = i >= 1 && i <= B.ni && j <= B.nj && k <= B.nk;
    const long box = B.box, x = valid ? cidx(B, i, j, k) : 0;
    T *dst = (T *)B.srow + row * (kSweepQ * 64) + lane;
    if (what & kGatherState) {
        for (int c = 0; c < 5; ++c) dst[(kSQ_Q + c) * 64] = valid ? B.Q[c * box + x] : (c == 0 ? 1.0 : 0.0);
        dst[kSQ_P * 64] = valid ? B.W[4 * box + x] : 1.0;
        for (int d = 0; d < 3; ++d) dst[(kSQ_LAM + d) * 64] = valid ? B.lamc[d * box + x] : 0.0;
        for (int c = 0; c < 5; ++c) dst[(kSQ_R + c) * 64] = valid ? B.R[c * box + x] : 0.0;
        dst[kSQ_A * 64] = valid ? B.alpha[x] : 1.0;
        for (int d = 0; d < 3; ++d) dst[(kSQ_LV + d) * 64] = (valid && B.lamv) ? B.lamv[d * box + x] : 0.0;
        dst[kSQ_EV * 64] = (valid && B.ev) ? B.ev[x] : 0.0;
        dst[kSQ_REV * 64] = (valid && B.res_ev) ? B.res_ev[x] : 0.0;
        dst[kSQ_A2 * 64] = (valid && B.alpha2) ? B.alpha2[x] : 1.0;
    }
    if (what & kGatherMetrics)
        for (int q = 0; q < 15; ++q) dst[(kSQ_SWG + q) * 64] = valid ? B.swg[q * box + x] : 0.0;
}

namespace {
// Work item of workgroup b in a grid of n.  The hardware hands workgroup b to
// XCD b mod 8, so XCD x runs b = x, x + 8, ...; XCD x takes the x-th run of
// consecutive work items, in the order it runs them.  Neighbouring items share
// cache lines at their edges (a 16-cell run of k starts at any offset within a
// 128 B line), which are then fetched into, and written back from, one L2.
__device__ inline long xcd_item(long b, long n)
{
    const long x = b % 8;
    long start = 0;
    for (long y = 0; y < x; ++y) start += (n - y + 7) / 8;   // items of the XCDs before x
    return start + b / 8;
}
// (tile, chunk) of pack/unpack work item wg of a block: k tiles fastest, then
// chunks of i, then j tiles, so consecutive items are neighbours along k (they
// share the cell-box lines at their k edges) and items ntk apart neighbours
// along i (they share the partly written rows at their chunk edges)
__device__ inline void pack_item(const BlockDev &B, long wg, int nchunk, int &slot, int &i0)
{
    const int K = (int)(wg % B.ntk);
    const long rest = wg / B.ntk;
    i0 = (int)(rest % nchunk) * kPackChunk + 1;
    slot = (int)(rest / nchunk) * B.ntk + K;
}
}  // namespace

// One workgroup per (tile, chunk of kPackChunk i-planes, group of quantities):
// the quantities are split into qsplit groups so that a small grid still
// spreads over the chip.  The chunk's cells are read plane by plane along k
// (coalesced) into LDS, then written out row by row: the rows a chunk touches
// are partly written (the lanes whose cell lies in the chunk), each as one
// contiguous run.
template <class T>
__global__ void __launch_bounds__(256) k_sweep_pack(const BlockDev *__restrict__ blocks, const long *__restrict__ gstart,
                                                    int nblock, int nq, int qsplit)
{
    constexpr int TJ = kKbaTJ, TK = kKbaTK;
    __shared__ double buf[kPackChunk][64];
    // the block of this workgroup: gstart holds the running (tile, chunk) counts
    const long item = xcd_item(blockIdx.x, gridDim.x);
    const int qg = (int)(item % qsplit);
    const int block = find_range(gstart, nblock, item / qsplit);
    const BlockDev &B = blocks[block];
    const long wg = item / qsplit - gstart[block];
    const int nchunk = (B.ni + kPackChunk - 1) / kPackChunk;
    int slot, i0;
    pack_item(B, wg, nchunk, slot, i0);
    const int J = slot / B.ntk, K = slot % B.ntk;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tj = lane / TK, tk = lane % TK, j = J * TJ + tj + 1, k = K * TK + tk + 1;
    const bool col = j <= B.nj && k <= B.nk;
    const long box = B.box;
    const int NQ = nq;   // 15; 18 with the lamv rows of Navier-Stokes; 21 with the nu~ rows of Spalart-Allmaras
    // the cell-box array and row quantity of each packed quantity
    for (int qi = qg * NQ / qsplit; qi < (qg + 1) * NQ / qsplit; ++qi) {
        const double *src;
        int rq;
        if (qi < 5) { src = B.Q + qi * box; rq = kSQ_Q + qi; }
        else if (qi == 5) { src = B.W + 4 * box; rq = kSQ_P; }
        else if (qi < 9) { src = B.lamc + (qi - 6) * box; rq = kSQ_LAM + qi - 6; }
        else if (qi < 14) { src = B.R + (qi - 9) * box; rq = kSQ_R + qi - 9; }
        else if (qi == 14) { src = B.alpha; rq = kSQ_A; }
        else if (qi < 18) { src = B.lamv + (qi - 15) * box; rq = kSQ_LV + qi - 15; }
        else if (qi == 18) { src = B.ev; rq = kSQ_EV; }
        else if (qi == 19) { src = B.res_ev; rq = kSQ_REV; }
        else { src = B.alpha2; rq = kSQ_A2; }
        for (int ci = w; ci < kPackChunk; ci += 4) {
            const int i = i0 + ci;
            if (col && i <= B.ni) buf[ci][lane] = src[cidx(B, i, j, k)];
        }
        __syncthreads();
        // rows r = i + tj + tk - 1 for i in the chunk
        for (int rr = w; rr < kPackChunk + TJ + TK - 2; rr += 4) {
            const int r = i0 - 1 + rr, i = r + 1 - tj - tk;
            if (col && i >= i0 && i < i0 + kPackChunk && i <= B.ni)
                ((T *)B.srow)[((long)slot * B.nr + r) * (kSweepQ * 64) + rq * 64 + lane] = (T)buf[i - i0][lane];
        }
        __syncthreads();
    }
}

// The reverse of k_sweep_pack for dQ (and dnu~), components split into
// csplit groups the same way.
template <class T>
__global__ void __launch_bounds__(256) k_sweep_unpack(const BlockDev *__restrict__ blocks, const long *__restrict__ gstart,
                                                      int nblock, int nc, int csplit)
{
    constexpr int TJ = kKbaTJ, TK = kKbaTK;
    __shared__ double buf[kPackChunk][64];
    // the block of this workgroup: gstart holds the running (tile, chunk) counts
    const long item = xcd_item(blockIdx.x, gridDim.x);
    const int cg = (int)(item % csplit);
    const int block = find_range(gstart, nblock, item / csplit);
    const BlockDev &B = blocks[block];
    const long wg = item / csplit - gstart[block];
    const int nchunk = (B.ni + kPackChunk - 1) / kPackChunk;
    int slot, i0;
    pack_item(B, wg, nchunk, slot, i0);
    const int J = slot / B.ntk, K = slot % B.ntk;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tj = lane / TK, tk = lane % TK, j = J * TJ + tj + 1, k = K * TK + tk + 1;
    const bool col = j <= B.nj && k <= B.nk;
    const long box = B.box;
    const int NC = nc;   // 6: dnu~ too (the nu~ equation in the sweep)
    for (int c = cg * NC / csplit; c < (cg + 1) * NC / csplit; ++c) {
        for (int rr = w; rr < kPackChunk + TJ + TK - 2; rr += 4) {
            const int r = i0 - 1 + rr, i = r + 1 - tj - tk;
            if (col && i >= i0 && i < i0 + kPackChunk && i <= B.ni)
                buf[i - i0][lane] = (double)((const T *)B.sdq)[((long)slot * B.nr + r) * (kSweepD * 64) + c * 64 + lane];
        }
        __syncthreads();
        double *dst = c < 5 ? B.dQ + c * box : B.dev;
        for (int ci = w; ci < kPackChunk; ci += 4) {
            const int i = i0 + ci;
            if (col && i <= B.ni) dst[cidx(B, i, j, k)] = buf[ci][lane];
        }
        __syncthreads();
    }
}

template __global__ void k_sweep_fill<double>(const BlockDev *, int, long, int);
template __global__ void k_sweep_fill<float>(const BlockDev *, int, long, int);
template __global__ void k_sweep_pack<double>(const BlockDev *, const long *, int, int, int);
template __global__ void k_sweep_pack<float>(const BlockDev *, const long *, int, int, int);
template __global__ void k_sweep_unpack<double>(const BlockDev *, const long *, int, int, int);
template __global__ void k_sweep_unpack<float>(const BlockDev *, const long *, int, int, int);

// The persistent wavefront form of the sweeps, on the tile-skewed layout.  The
// (j,k) cross-section is cut into 4 x 16 tiles, one workgroup of four waves
// per tile; tiles are handed out by a ticket counter in topological order, so
// a tile waits only on tiles already running.  Waves 0, 1, 2 each own one
// neighbour direction (i, j, k): a step ahead they load the neighbour's
// state and metrics, then form the flux difference and off-diagonal terms of
// the step's 64 cells in LDS; wave 3 combines them into dQ, keeps dQ in LDS for
// the next step and stores it.  Every load is issued unconditionally with
// clamped indices, so the compiler's memory-counter waits are exact.  The
// per-hyperplane kernel k_lusgs_plane above is the reference's own schedule
// and gives the same bits.  A step touches one row of its tile: the three
// direction waves read their neighbours' state and metrics from the previous
// row (a lane shift) or, on the tile's upstream edge, from the upstream tile's
// row, so every load is a contiguous run of one row.  dQ goes to the tile's
// dQ rows.
//
// Across a tile edge dQ travels through the upstream tile's mailbox: its
// combine wave stores the dQ of its edge lanes there as soon as it has them,
// and the consumer reads the words a step ahead and, before using them, waits
// (bounded) while any word it needs still holds the empty pattern.  Tiles
// start in ticket order, so every upstream tile is already running.
//
// With SA the nu~ equation of Spalart-Allmaras rides along as a sixth
// component (source/LU_SGS_Marching.cpp:88-157, 178-257): the direction waves
// form (nu~ + dnu~) Vn_new - nu~ Vn_old of their neighbour from the same Vn
// as the flow terms, and the combine wave divides by alpha2.  The reference
// updates both per cell from the same neighbour increments, so this is its
// order.
template <bool FWD, int TJ, int TK, bool VISC, bool SA, class T>
__global__ void __launch_bounds__(256) k_lusgs_sl(const BlockDev *__restrict__ blocks, const KbaTile *__restrict__ tiles,
                                                  int ntiles, int *__restrict__ ticket, int *__restrict__ err,
                                                  double relax, Gas g)
{
    static_assert(TJ * TK == 64, "64 cells per tile");
    static_assert(!SA || VISC, "Spalart-Allmaras runs with the viscous terms");
    // The FP32 operator forms the flux difference without cancellation; the
    // FP64 one follows the reference's expression order.
    constexpr bool kDeltaForm = sizeof(T) == 4;
    constexpr int NC = SA ? 6 : 5;   // increments carried per cell
    __shared__ T sdq[2][NC][64];
    __shared__ T sA[3][NC][64];
    __shared__ T sV[3][NC][64];
    __shared__ T sR[SA ? 8 : 6][64];   // R, alpha; with SA the nu~ right-hand side and alpha2
    __shared__ int s_q;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int tj = lane / TK, tk = lane % TK;
    const T Gamma = (T)g.gamma;
    constexpr int RQ = kSweepQ * 64, RD = kSweepD * 64;   // words per row
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(ticket, 1);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_q);
        __syncthreads();
        if (q >= ntiles) return;
        const KbaTile tl = tiles[q];
        const BlockDev &B = blocks[tl.block];
        const int ni = B.ni, nj = B.nj, nk = B.nk, nr = B.nr, ntk = B.ntk;
        const int slot = (tl.j0 / TJ) * ntk + tl.k0 / TK;
        const gcT<T> rows = (gcT<T>)((const T *)B.srow + (long)slot * nr * RQ);
        const gT<T> drows = (gT<T>)((T *)B.sdq + (long)slot * nr * RD);
        const int lj = FWD ? tj : TJ - 1 - tj, lk = FWD ? tk : TK - 1 - tk;
        const int j = tl.j0 + tj + 1, k = tl.k0 + tk + 1;
        const bool col = j <= nj && k <= nk;
        const int nsteps = nr;
        // steps run: [kRowS0, send), a whole number of kRing-step unrollings
        const int send = (nsteps - kRowS0 + kRing - 1) / kRing * kRing + kRowS0;
        auto clampr = [&](int r) { return min(max(r, 0), nr - 1); };
        // own row of a step, and whether this lane's cell is in the block
        auto row_of = [&](int step) { return FWD ? step : nr - 1 - step; };
        auto cell_at = [&](int step, int &i) -> bool {
            const int l2 = step - lj - lk;
            i = FWD ? l2 + 1 : ni - l2;
            return col && l2 >= 0 && l2 < ni;
        };
        auto bar = []() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        };
        auto dir_loop = [&](auto dc) {
            constexpr int d = decltype(dc)::value;
            constexpr int TE = d == 1 ? TJ : TK;   // tile extent along d
            const int dep = d == 1 ? tl.dep_j : d == 2 ? tl.dep_k : -1;
            const bool edge = d == 1 ? tj == (FWD ? 0 : TJ - 1) : d == 2 ? tk == (FWD ? 0 : TK - 1) : false;
            // upstream tile's rows (this tile's own when there is none: the
            // values are then never used)
            const int uslot = dep < 0 ? slot : d == 1 ? slot + (FWD ? -ntk : ntk) : slot + (FWD ? -1 : 1);
            const gcT<T> urows = (gcT<T>)((const T *)B.srow + (long)uslot * nr * RQ);
            const gcT<T> udrows = (gcT<T>)((const T *)B.sdq + (long)uslot * nr * RD);
            // the upstream lane facing this lane across the tile edge
            const int ulane = d == 1 ? (FWD ? (TJ - 1) * TK : 0) + tk : tj * TK + (FWD ? TK - 1 : 0);
            // in-tile neighbour lane (lanes on the edge read the upstream row instead)
            const int nlane = (d == 0 || edge) ? lane : d == 1 ? lane + (FWD ? -TK : TK) : lane + (FWD ? -1 : 1);
            const bool from_up = d != 0 && edge && dep >= 0;
            // the upstream tile's mailbox words this lane reads (the other lanes
            // read the same words as the edge lanes)
            const gT<T> umbox = (gT<T>)(mbox_base<T>(B, FWD) + (long)uslot * nr * kMboxRow);
            const int mword = d == 1 ? tk : kMboxJ + tj;
            const int mstride = d == 1 ? TK : TJ;
            // upstream dQ of a step, read one step ahead without waiting
            auto issue_cross = [&](RowPf<T> &bn, int step) {
                if (d == 0) return;
                const int urow = clampr(FWD ? row_of(step) + TE - 1 : row_of(step) - (TE - 1));
                for (int c = 0; c < NC; ++c)
                    bn.x[c] = __hip_atomic_load(umbox + urow * kMboxRow + mword + c * mstride, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            };
            // before using the upstream dQ of step s: re-read while a word the
            // step needs is still empty (bounded; a run-out sets the error word)
            auto settle_cross = [&](RowPf<T> &b, int s, bool need) {
                if (d == 0) return;
                auto empty = [&](const T *x) {
                    bool e = false;
                    for (int c = 0; c < NC; ++c) e = e || (mbox_empty(x[c]));
                    return e;
                };
                if (!__builtin_amdgcn_ballot_w64(need && empty(b.x))) return;
                const int urow = clampr(FWD ? row_of(s) + TE - 1 : row_of(s) - (TE - 1));
                long spins = 0;
                while (__builtin_amdgcn_ballot_w64(need && empty(b.x))) {
                    __builtin_amdgcn_s_sleep(1);
                    for (int c = 0; c < NC; ++c)
                        b.x[c] = __hip_atomic_load(umbox + urow * kMboxRow + mword + c * mstride, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    if (++spins > (1L << 22)) { atomicExch(err, 1); break; }
                }
            };
            // the neighbour's state and metrics of a step (static during the
            // sweep, so no wait), and for wave 0 the cell's right-hand side
            auto issue_static = [&](RowPf<T> &bn, int step) {
                const int rt = row_of(step);
                const int urow = clampr(FWD ? rt + TE - 1 : rt - (TE - 1));
                const int nrow = clampr(FWD ? rt - 1 : rt + 1);
                const gcT<T> src = (from_up ? urows + urow * RQ + ulane : rows + nrow * RQ + nlane);
                for (int c = 0; c < 5; ++c) bn.q[c] = src[(kSQ_Q + c) * 64];
                bn.p = src[kSQ_P * 64];
                bn.l = src[(kSQ_LAM + d) * 64];
                if (VISC) bn.lv = src[(kSQ_LV + d) * 64];
                for (int c = 0; c < 5; ++c) bn.m[c] = src[(kSQ_SWG + d * 5 + c) * 64];
                if constexpr (SA) bn.ev = src[kSQ_EV * 64];
                if (d == 0) {
                    const int ro = clampr(rt);
                    for (int c = 0; c < 5; ++c)
                        bn.r[c] = FWD ? rows[ro * RQ + (kSQ_R + c) * 64 + lane] : ((gcT<T>)drows)[ro * RD + c * 64 + lane];
                    bn.a = rows[ro * RQ + kSQ_A * 64 + lane];
                    if constexpr (SA) {
                        // forward: Res_ev; backward: the forward dnu~
                        bn.rev = FWD ? rows[ro * RQ + kSQ_REV * 64 + lane] : ((gcT<T>)drows)[ro * RD + 5 * 64 + lane];
                        bn.a2 = rows[ro * RQ + kSQ_A2 * 64 + lane];
                    }
                }
            };
            // the terms of the flux difference that use only the old state
            // (source/LU_SGS_Marching.cpp:293-419; same operations, evaluated early)
            auto precompute = [&](RowPf<T> &b) {
                const T *o = b.q;
                const T nx = b.m[0], ny = b.m[1], nz = b.m[2], Vn_e = b.m[4];
                const T press_old = b.p;
                if constexpr (kDeltaForm) {
                    // old velocity and total enthalpy per volume for the
                    // cancellation-free difference below (one reciprocal of
                    // the density: this operator only preconditions)
                    const T r0 = rcp_op(o[0]);
                    b.of[0] = o[1] * r0;
                    b.of[1] = o[2] * r0;
                    b.of[2] = o[3] * r0;
                    b.of[3] = o[4] + press_old;
                    b.vn_old = b.of[0] * nx + b.of[1] * ny + b.of[2] * nz;
                } else {
                    const T Vn_old = o[1] / o[0] * nx + o[2] / o[0] * ny + o[3] / o[0] * nz;
                    b.vn_old = Vn_old;
                    b.of[0] = o[0] * (Vn_old - Vn_e);
                    b.of[1] = o[1] * (Vn_old - Vn_e) + press_old * nx;
                    b.of[2] = o[2] * (Vn_old - Vn_e) + press_old * ny;
                    b.of[3] = o[3] * (Vn_old - Vn_e) + press_old * nz;
                    b.of[4] = (o[4] + press_old) * (Vn_old - Vn_e) + press_old * Vn_e;
                }
                b.cl = (T)relax * b.l;
            };
            // step s: the dQ-dependent part of the flux difference (critical
            // path), then, while the combine wave works, the loads and early
            // terms of the next steps
            auto work = [&](RowPf<T> &b, RowPf<T> &b1, RowPf<T> &bc, RowPf<T> &bn, int s) {
                int i;
                const bool valid = cell_at(s, i);
                const int p = d == 0 ? i : d == 1 ? j : k, lim = d == 0 ? ni : d == 1 ? nj : nk;
                const bool use = valid && (FWD ? p != 1 : p != lim);
                settle_cross(b, s, use && from_up);
                T dqn[NC];
                for (int c = 0; c < NC; ++c) {
                    const T in_tile = sdq[(s - 1) & 1][c][nlane];
                    dqn[c] = (d != 0 && edge) ? b.x[c] : in_tile;
                }
                const T nx = b.m[0], ny = b.m[1], nz = b.m[2], sd = b.m[3], Vn_e = b.m[4];
                T df[5];
                T vn_new = T(0), dvn = T(0);   // the neighbour's new normal velocity and its change
                if constexpr (kDeltaForm) {
                    // F_n(Q+dQ) - F_n(Q) written as sums of products with a
                    // difference factor, so its rounding error scales with dQ:
                    //   du = (dm - u d rho)/(rho + d rho),  dVn = du.n,
                    //   dp = (gamma-1)(dE - (m.du + dm.u')/2),  w' = Vn' - Vn_e
                    //   d(rho w) = d rho w' + rho dVn,  d(m w + p n) = dm w' + m dVn + dp n,
                    //   d((E+p) w + p Vn_e) = (dE + dp) w' + (E+p) dVn + dp Vn_e
                    const T *q = b.q;
                    const T r1 = rcp_op(q[0] + dqn[0]);
                    const T du = (dqn[1] - b.of[0] * dqn[0]) * r1, dv = (dqn[2] - b.of[1] * dqn[0]) * r1,
                            dw = (dqn[3] - b.of[2] * dqn[0]) * r1;
                    const T u1 = b.of[0] + du, v1 = b.of[1] + dv, w1 = b.of[2] + dw;
                    const T dVn = du * nx + dv * ny + dw * nz;
                    const T wn = (b.vn_old + dVn) - Vn_e;
                    const T dp = (Gamma - TC(1.0)) *
                                 (dqn[4] - TC(0.5) * ((q[1] * du + q[2] * dv + q[3] * dw) +
                                                      (dqn[1] * u1 + dqn[2] * v1 + dqn[3] * w1)));
                    vn_new = b.vn_old + dVn;
                    dvn = dVn;
                    df[0] = dqn[0] * wn + q[0] * dVn;
                    df[1] = dqn[1] * wn + q[1] * dVn + dp * nx;
                    df[2] = dqn[2] * wn + q[2] * dVn + dp * ny;
                    df[3] = dqn[3] * wn + q[3] * dVn + dp * nz;
                    df[4] = (dqn[4] + dp) * wn + b.of[3] * dVn + dp * Vn_e;
                } else {
                    T nw[5];
                    for (int c = 0; c < 5; ++c) nw[c] = b.q[c] + dqn[c];
                    T dens = nw[0], uvel = nw[1] / nw[0], vvel = nw[2] / nw[0], wvel = nw[3] / nw[0], E = nw[4] / nw[0];
                    T press_new = (Gamma - TC(1.0)) * dens * (E - TC(0.5) * (uvel * uvel + vvel * vvel + wvel * wvel));
                    T Vn_new = uvel * nx + vvel * ny + wvel * nz;
                    vn_new = Vn_new;
                    df[0] = nw[0] * (Vn_new - Vn_e) - b.of[0];
                    df[1] = (nw[1] * (Vn_new - Vn_e) + press_new * nx) - b.of[1];
                    df[2] = (nw[2] * (Vn_new - Vn_e) + press_new * ny) - b.of[2];
                    df[3] = (nw[3] * (Vn_new - Vn_e) + press_new * nz) - b.of[3];
                    df[4] = ((nw[4] + press_new) * (Vn_new - Vn_e) + press_new * Vn_e) - b.of[4];
                }
                T A[5], V[5];
                for (int c = 0; c < 5; ++c) {
                    const T a = FWD ? (df[c] * sd + dqn[c] * b.cl) * TC(0.5) : (df[c] * sd - dqn[c] * b.cl) * TC(0.5);
                    const T v = dqn[c] * (VISC ? b.lv : T(0));   // lamv * dQ, lamv = 0 on the inviscid path
                    A[c] = use ? a : T(0);
                    V[c] = use ? v : T(0);
                }
                for (int c = 0; c < 5; ++c) { sA[d][c][lane] = A[c]; sV[d][c][lane] = V[c]; }
                if constexpr (SA) {
                    // (nu~ + dnu~) Vn_new - nu~ Vn_old (source/LU_SGS_Marching.cpp:351-353); the
                    // FP32 operator writes it as nu~ dVn + dnu~ Vn_new
                    const T dev = dqn[5];
                    T dfe;
                    if constexpr (kDeltaForm) {
                        dfe = b.ev * dvn + dev * vn_new;
                    } else {
                        (void)dvn;
                        const T ev_new = b.ev + dev;
                        dfe = ev_new * vn_new - b.ev * b.vn_old;
                    }
                    const T a = FWD ? (dfe * sd + b.cl * dev) * TC(0.5) : (dfe * sd - b.cl * dev) * TC(0.5);
                    sA[d][5][lane] = use ? a : T(0);
                    sV[d][5][lane] = use ? b.lv * dev : T(0);
                }
                if (d == 0) {
                    for (int c = 0; c < 5; ++c) sR[c][lane] = b.r[c];
                    sR[5][lane] = b.a;
                    if constexpr (SA) {
                        sR[6][lane] = b.rev;
                        sR[7][lane] = b.a2;
                    }
                }
                bar();   // terms of step s are in LDS
                issue_cross(bc, s + kCrossAhead);
                issue_static(bn, s + kStaticAhead);
                precompute(b1);
                bar();   // dQ of step s is in LDS
            };
            // b[u] serves step s0+u: static data issued kStaticAhead steps
            // ahead (into the buffer the previous step has finished with), the
            // upstream dQ kCrossAhead steps ahead.  The loop starts kStaticAhead idle
            // steps early (every lane invalid) so that no load precedes it and
            // every step runs the same code; nothing is in flight on entry.
            RowPf<T> b[kRing] = {};
            __builtin_amdgcn_s_waitcnt(0);
            for (int s0 = kRowS0; s0 < send; s0 += kRing) {
#pragma unroll
                for (int u = 0; u < kRing; ++u)
                    work(b[u], b[(u + 1) % kRing], b[(u + kCrossAhead) % kRing], b[(u + kStaticAhead) % kRing], s0 + u);
            }
        };
        if (wave == 0) {
            dir_loop(Dir<0>());
        } else if (wave == 1) {
            dir_loop(Dir<1>());
        } else if (wave == 2) {
            dir_loop(Dir<2>());
        } else {
            // ---- combine wave (source/LU_SGS_Marching.cpp:155, 255)
            const gT<T> mbox = (gT<T>)(mbox_base<T>(B, FWD) + (long)slot * nr * kMboxRow);
            // downstream edges of this tile: the j edge (lanes tj = jl, word tk)
            // and the k edge (lanes tk = kl, word tj)
            const bool jedge = tj == (FWD ? TJ - 1 : 0), kedge = tk == (FWD ? TK - 1 : 0);
            for (int s = kRowS0; s < send; ++s) {
                int i;
                const bool valid = cell_at(s, i);
                const int ro = clampr(row_of(s));
                T outv[NC];
                bar();
                if (valid) {
                    const T alpha = sR[5][lane];
                    T ra = T(0);
                    if constexpr (kDeltaForm) ra = rcp_op(alpha);
                    for (int c = 0; c < 5; ++c) {
                        T v;
                        if (FWD) {
                            v = sR[c][lane] * TC(-1.0) + sA[0][c][lane];
                            v = v + sA[1][c][lane]; v = v + sA[2][c][lane];
                        } else {
                            v = sR[c][lane] * alpha - sA[0][c][lane];
                            v = v - sA[1][c][lane]; v = v - sA[2][c][lane];
                        }
                        v = v + sV[0][c][lane]; v = v + sV[1][c][lane]; v = v + sV[2][c][lane];
                        if constexpr (kDeltaForm) outv[c] = v * ra;
                        else outv[c] = v / alpha;
                        sdq[s & 1][c][lane] = outv[c];
                    }
                    if constexpr (SA) {
                        // source/LU_SGS_Marching.cpp:156-157, 256-257
                        const T alpha2 = sR[7][lane];
                        T v;
                        if (FWD) {
                            v = sR[6][lane] * TC(-1.0) + sA[0][5][lane];
                            v = v + sA[1][5][lane]; v = v + sA[2][5][lane];
                        } else {
                            v = sR[6][lane] * alpha2 - sA[0][5][lane];
                            v = v - sA[1][5][lane]; v = v - sA[2][5][lane];
                        }
                        v = v + sV[0][5][lane]; v = v + sV[1][5][lane]; v = v + sV[2][5][lane];
                        if constexpr (kDeltaForm) outv[5] = v * rcp_op(alpha2);
                        else outv[5] = v / alpha2;
                        sdq[s & 1][5][lane] = outv[5];
                    }
                }
                bar();
                if (valid) {
                    // hand the edge values to the downstream tiles, then keep dQ
                    if (jedge)
                        for (int c = 0; c < NC; ++c)
                            __hip_atomic_store(mbox + ro * kMboxRow + tk + c * TK, outv[c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    if (kedge)
                        for (int c = 0; c < NC; ++c)
                            __hip_atomic_store(mbox + ro * kMboxRow + kMboxJ + tj + c * TJ, outv[c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    for (int c = 0; c < NC; ++c) drows[ro * RD + c * 64 + lane] = outv[c];
                }
            }
        }
    }
}

#define FV3D_SL(F, V, S, T) \
    template __global__ void k_lusgs_sl<F, kKbaTJ, kKbaTK, V, S, T>(const BlockDev *, const KbaTile *, int, int *, int *, double, Gas);
FV3D_SL(true, false, false, double)
FV3D_SL(false, false, false, double)
FV3D_SL(true, true, false, double)
FV3D_SL(false, true, false, double)
FV3D_SL(true, true, true, double)
FV3D_SL(false, true, true, double)
FV3D_SL(true, false, false, float)
FV3D_SL(false, false, false, float)
FV3D_SL(true, true, false, float)
FV3D_SL(false, true, false, float)
FV3D_SL(true, true, true, float)
FV3D_SL(false, true, true, float)
#undef FV3D_SL

__global__ void k_fill_u64(unsigned long long *p, long n, unsigned long long v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_i32(int *p, int n, int v)
{
    for (int e = threadIdx.x; e < n; e += blockDim.x) p[e] = v;
}

__global__ void k_fill_int(int *p, long n, int v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// ===================================================== update + norms
// Q += dQ (source/LU_SGS_Marching.cpp:262-287), Res_Calculate terms
// (source/ComputeAllBlockVars.cpp:340-357), PrimitiveVars
// (source/PrimitiveVars.cpp:7-29) and the Limit_Vars trigger test
// (source/Limit_FlowVars.cpp:28-29).  One partial sum per workgroup.  With
// Spalart-Allmaras (sa set) also nu~ += dnu~ (source/LU_SGS_Marching.cpp:271-272)
// and, when sa_hi is given, the flag of a block holding a value the limiter
// replaces by its neighbours' mean (above 1000).
__global__ void __launch_bounds__(256) k_update(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart, int nblock,
                                                Gas g, double *__restrict__ partial, int *__restrict__ bad_flag,
                                                int *__restrict__ sa_hi, int sa)
{
    __shared__ double red[5][256];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double rm[5] = {0, 0, 0, 0, 0};
    for_cell_item(blocks, cstart, nblock, t, [&](const BlockDev &B, int n, int i, int j, int k) {
        const long box = B.box, x = cidx(B, i, j, k);
        double q[5];
        const double vol = B.vol[x];
        for (int c = 0; c < 5; ++c) {
            double o = B.Q[c * box + x];
            q[c] = o + B.dQ[c * box + x];
            rm[c] = fabs((o - q[c]) / o) * vol;
            B.Q[c * box + x] = q[c];
        }
        const double Gamma = g.gamma, Rgas = g.rgas;
        double d = q[0], u = q[1] / q[0], v = q[2] / q[0], w = q[3] / q[0];
        double V2 = u * u + v * v + w * w;
        double p = (Gamma - F32C(1.0)) * (q[4] - q[0] * V2 / F32C(2.0));
        double T = p / d / Rgas;
        B.W[x] = d; B.W[box + x] = u; B.W[2 * box + x] = v; B.W[3 * box + x] = w; B.W[4 * box + x] = p; B.W[5 * box + x] = T;
        const double dmin = F32C(1.0E-6), dmax = F32C(1000.0), vmax = F32C(1000.0);
        if (d < dmin || d > dmax || p < dmin || p > dmax || u > vmax || v > vmax || w > vmax) atomicOr(bad_flag + n, 1);
        if (sa) {
            const double e = B.ev[x] + B.dev[x];
            B.ev[x] = e;
            if (sa_hi && e > F32C(1000.0)) atomicOr(sa_hi + n, 1);
        }
    });
    for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = rm[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = red[c][threadIdx.x] + red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 5; ++c) partial[(long)c * gridDim.x + blockIdx.x] = red[c][0];
}

// Limit_Vars (source/Limit_FlowVars.cpp:8-119) in the reference's serial order:
// cells are repaired one after the other, so later repairs see earlier ones.
// Runs as one thread and returns at once unless k_update flagged a bad cell.
__device__ static bool bad_cell(const BlockDev &B, long x)
{
    const long box = B.box;
    const double dmin = F32C(1.0E-6), dmax = F32C(1000.0), vmax = F32C(1000.0);
    double d = B.W[x], p = B.W[4 * box + x];
    return d < dmin || d > dmax || p < dmin || p > dmax || B.W[box + x] > vmax || B.W[2 * box + x] > vmax ||
           B.W[3 * box + x] > vmax;
}

// Res_Calculate runs after Limit_Vars in the reference
// (source/LU_SGS_Marching.cpp:287-290, source/Solver.cpp:92-93), but k_update
// formed the norm partials from the state before the repair.  With partial
// given, each repaired cell's |(q_old - q)/q_old| Vol terms are moved from the
// pre-repair value to the repaired one in its 256-cell group's partial
// (q_old = the pre-repair value minus dQ).  Without it (the limiter after a
// prolongation, which no norm follows) nothing is patched.
__global__ void k_limiter_serial(const BlockDev *__restrict__ blocks, int nblock, Gas g, int *__restrict__ bad_flag,
                                 const long *__restrict__ cstart, double *__restrict__ partial, int npart)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double Gamma = g.gamma, Rgas = g.rgas;
    const double dmin = F32C(1.0E-6), dmax = F32C(1000.0), vmax = F32C(1000.0);
    for (int n = 0; n < nblock; ++n) {
        if (!bad_flag[n]) continue;
        bad_flag[n] = 0;
        const BlockDev &B = blocks[n];
        const long box = B.box;
        int nneg = 0;
        for (int i = 1; i <= B.ni; ++i)
            for (int j = 1; j <= B.nj; ++j)
                for (int k = 1; k <= B.nk; ++k) {
                    const long x = cidx(B, i, j, k);
                    if (!bad_cell(B, x)) continue;
                    nneg++;
                    int kn = 0, i1 = 0, j1 = 0, k1 = 0;
                    double d = 0.0, u = 0.0, v = 0.0, w = 0.0, p = 0.0;
                    for (int ia = -1; ia <= 1; ++ia)
                        for (int ja = -1; ja <= 1; ++ja)
                            for (int ka = -1; ka <= 1; ++ka) {
                                k1 = k + ka; j1 = j + ja; i1 = i + ia;
                                if (i1 > 0 && i1 < B.IM && j1 > 0 && j1 < B.JM && k1 > 0 && k1 < B.KM) {
                                    const long y = cidx(B, i1, j1, k1);
                                    if (!bad_cell(B, y)) {
                                        d = d + B.W[y]; u = u + B.W[box + y]; v = v + B.W[2 * box + y];
                                        w = w + B.W[3 * box + y]; p = p + B.W[4 * box + y];
                                        kn = kn + 1;
                                    }
                                }
                            }
                    if (kn == 0) {
                        const long y = cidx(B, i1, j1, k1);
                        d = B.W[y]; u = B.W[box + y]; v = B.W[2 * box + y]; w = B.W[3 * box + y]; p = B.W[4 * box + y];
                        if (d < dmin) d = dmin;
                        if (d > dmax) d = dmax;
                        if (p < dmin) p = dmin;
                        if (p > dmax) p = dmax;
                        if (fabs(u) > vmax) u = u < 0 ? -fabs(vmax) : fabs(vmax);
                        if (fabs(v) > vmax) v = v < 0 ? -fabs(vmax) : fabs(vmax);
                        if (fabs(w) > vmax) w = w < 0 ? -fabs(vmax) : fabs(vmax);
                    } else {
                        double sn1 = (double)(1.0f / (float)kn);
                        d = d * sn1; u = u * sn1; v = v * sn1; w = w * sn1; p = p * sn1;
                    }
                    B.W[x] = d; B.W[box + x] = u; B.W[2 * box + x] = v; B.W[3 * box + x] = w; B.W[4 * box + x] = p;
                    B.W[5 * box + x] = p / d / Rgas;
                    double qn[5];
                    qn[0] = d; qn[1] = d * u; qn[2] = d * v; qn[3] = d * w;
                    qn[4] = p / (Gamma - F32C(1.0)) + d * (u * u + v * v + w * w) / F32C(2.0);
                    if (partial) {
                        const long t = cstart[n] + ((long)(i - 1) * B.nj + (j - 1)) * B.nk + (k - 1);
                        const double vol = B.vol[x];
                        for (int c = 0; c < 5; ++c) {
                            const double qp = B.Q[c * box + x], o = qp - B.dQ[c * box + x];
                            partial[(long)c * npart + t / 256] += fabs((o - qn[c]) / o) * vol - fabs((o - qp) / o) * vol;
                        }
                    }
                    for (int c = 0; c < 5; ++c) B.Q[c * box + x] = qn[c];
                }
        if (nneg > 0) *B.flag_low_dt = 10;
    }
}

// ============================================= section integrals, norms
// Mass flow, mass-averaged total pressure and total temperature over one K
// plane (source/Ouput.cpp:9-95).  One workgroup per section, fixed-order tree
// reduction; sums[sec*5 + {mass, ptm, ttm}].
// grid: 2 x kSecParts workgroups; workgroup p of a section sums the cells
// p*256 + t, (p + kSecParts)*256 + t, ... into sums[(sec*kSecParts + p)*3 + c]
__global__ void __launch_bounds__(256) k_sections(const BlockDev *__restrict__ blocks, const int2 *__restrict__ secs, Gas g,
                                                  double *__restrict__ sums)
{
    __shared__ double red[3][256];
    const int sec = blockIdx.x / kSecParts, part = blockIdx.x % kSecParts;
    const int2 sc = secs[sec];
    const BlockDev &B = blocks[sc.x];
    const int K = sc.y;
    const long box = B.box;
    const double Gamma = g.gamma, Rgas = g.rgas;
    double m = 0.0, pt = 0.0, tt = 0.0;
    for (int f = part * blockDim.x + threadIdx.x; f < B.ni * B.nj; f += kSecParts * blockDim.x) {
        int i = f / B.nj + 1, j = f % B.nj + 1;
        long a = cidx(B, i, j, K), z = cidx(B, i, j, K - 1);
        double dens = F32C(0.5) * (B.W[a] + B.W[z]), u = F32C(0.5) * (B.W[box + a] + B.W[box + z]);
        double v = F32C(0.5) * (B.W[2 * box + a] + B.W[2 * box + z]), w = F32C(0.5) * (B.W[3 * box + a] + B.W[3 * box + z]);
        const double *S = B.fS[2];
        double mc = dens * (u * S[a] + v * S[box + a] + w * S[2 * box + a]);
        double temp = F32C(0.5) * (B.W[5 * box + a] + B.W[5 * box + z]);
        double Ma = sqrt(u * u + v * v + w * w) / sqrt(Gamma * Rgas * temp);
        double Pt_cell = B.W[4 * box + a] * pow(F32C(1.0) + (Gamma - F32C(1.0)) / F32C(2.0) * Ma * Ma, Gamma / (Gamma - F32C(1.0)));
        m += mc;
        pt += pow(Pt_cell, (Gamma - F32C(1.0)) / Gamma) * mc;
        tt += B.W[5 * box + a] * (1 + (Gamma - F32C(1.0)) / F32C(2.0) * Ma * Ma) * mc;
    }
    red[0][threadIdx.x] = m; red[1][threadIdx.x] = pt; red[2][threadIdx.x] = tt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) sums[blockIdx.x * 3 + c] = red[c][0];
}

// Residual norm (fixed-order sum of the per-workgroup partials), performance
// parameters (source/Ouput.cpp:97-107,282-300), CFL ramp for the next iteration
// (source/Solver.cpp:49-60) and the history row.  One workgroup.
__global__ void __launch_bounds__(256) k_finalize(const double *__restrict__ partial, int npart, const int2 *__restrict__ parts,
                                                  const double *__restrict__ secsum, ScalarsDev *lv, ScalarsDev *ramp, Gas g,
                                                  double *__restrict__ hist_row)
{
    const int inst = blockIdx.x;
    ScalarsDev *sc = lv + inst;
    ramp += inst;
    secsum += (long)inst * 2 * kSecParts * 3;
    if (hist_row) hist_row += 7L * inst;
    const int p0 = parts[inst].x, np = parts[inst].y;
    // component 5: Res_ev of Spalart-Allmaras (k_sa_res)
    const int nc = g.turbulence == 200 ? 6 : 5;
    __shared__ double red[6][256];
    double r[6] = {0, 0, 0, 0, 0, 0};
    // each thread adds partials p = t, t + 256, ... in that order; eight of
    // them are loaded before the first is added, so a thread waits on memory
    // once per eight rather than once per partial (d_partial holds six
    // components, so the sixth is always in bounds)
    int p = threadIdx.x;
    for (; p + 7 * 256 < np; p += 8 * 256) {
        double v[8][6];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < 6; ++c) v[u][c] = partial[(long)c * npart + p0 + p + u * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                if (c < nc) r[c] += v[u][c];
    }
    for (; p < np; p += 256)
        for (int c = 0; c < nc; ++c) r[c] += partial[(long)c * npart + p0 + p];
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = r[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int c = 0; c < 5; ++c) sc->res[c] = red[c][0] / sc->total_vol;
    sc->res_ev = red[5][0] / sc->total_vol;
    const double Gamma = g.gamma, Rgas = g.rgas;
    // section partials in a fixed order: section 0 inlet, 1 outlet; mass, ptm, ttm
    double sums[6] = {0, 0, 0, 0, 0, 0};
    for (int sec = 0; sec < 2; ++sec)
        for (int p = 0; p < kSecParts; ++p)
            for (int c = 0; c < 3; ++c) sums[sec * 3 + c] += secsum[(sec * kSecParts + p) * 3 + c];
    // sums: section 0 inlet, section 1 outlet: mass, ptm, ttm
    double mi = sums[0] * sc->dens_ref * sc->vel_ref * sc->nblade_in;
    double mo = sums[3] * sc->dens_ref * sc->vel_ref * sc->nblade_out;
    double pti = pow(sums[1] / sums[0], Gamma / (Gamma - F32C(1.0))) * sc->dens_ref * sc->vel_ref * sc->vel_ref;
    double pto = pow(sums[4] / sums[3], Gamma / (Gamma - F32C(1.0))) * sc->dens_ref * sc->vel_ref * sc->vel_ref;
    double tti = sums[2] / sums[0] * sc->temp_ref, tto = sums[5] / sums[3] * sc->temp_ref;
    double PR = pto / pti;
    double L_adk = Gamma / (Gamma - F32C(1.0)) * Rgas * tti * (pow(PR, (Gamma - F32C(1.0)) / Gamma) - F32C(1.0));
    double L_k = Gamma / (Gamma - F32C(1.0)) * Rgas * tti * (tto / tti - F32C(1.0));
    ramp->iter += 1;
    if (hist_row) {
        hist_row[0] = ramp->iter; hist_row[1] = sc->res[0]; hist_row[2] = mi; hist_row[3] = mo;
        hist_row[4] = PR; hist_row[5] = L_adk / L_k; hist_row[6] = ramp->cfl;
    }
    sc->mass_in = mi; sc->mass_out = mo; sc->pr = PR; sc->eff = L_adk / L_k;
    // CFL ramp applied for the next iteration (source/Solver.cpp:49-59), on the
    // density residual of the level that ran
    if (sc->res[0] <= ramp->cfl_res) {
        ramp->cfl = ramp->cfl * F32C(1.2);
        ramp->cfl_res = ramp->cfl_res / F32C(10.0);
    }
    if (ramp->cfl > F32C(10.0)) ramp->cfl = F32C(10.0);
}

// ============================================================ mixing plane
// CMesh::BndCondMixingPlane (source/BndCondMixingPlane.cpp:287-522), one
// workgroup per radial line i of a plane.  The threads form the per-face
// terms of each side's pitch-wise sums into LDS; one thread adds them in j
// order (the reference's order) and forms the mixed-out state of each side:
// the flux-averaged mass, momentum, energy (and rho nu~) fluxes, pressure from
// the quadratic of lines 370 / 432, then velocity, density, temperature and
// the speed of sound.  The threads then write the characteristic ghost
// states: block 1's K-high ghosts (and Dummy layers) from the downstream
// average plus its outgoing waves, block 2's K-low ghosts from the upstream
// average plus its outgoing wave; nu~ ghosts take the upstream average.
struct MixAvg {
    double d, u, v, w, p, c, ev;
};

__device__ static void mix_terms(const BlockDev &B, int i, int j, int kf, int kc, int sa, double *t)
{
    const long xf = cidx(B, i, j, kf), xc = cidx(B, i, j, kc), box = B.box;
    const double s = B.fs[2][xf];
    const double nx = B.fS[2][xf] / s, ny = B.fS[2][box + xf] / s, nz = B.fS[2][2 * box + xf] / s;
    const double u = B.W[box + xc], v = B.W[2 * box + xc], w = B.W[3 * box + xc], p = B.W[4 * box + xc];
    const double Vn = u * nx + v * ny + w * nz;
    t[0] = s;
    t[1] = nx * s;
    t[2] = ny * s;
    t[3] = nz * s;
    t[4] = B.Q[xc] * Vn * s;
    t[5] = (B.Q[box + xc] * Vn + nx * p) * s;
    t[6] = (B.Q[2 * box + xc] * Vn + ny * p) * s;
    t[7] = (B.Q[3 * box + xc] * Vn + nz * p) * s;
    t[8] = (B.Q[4 * box + xc] + p) * Vn * s;
    t[9] = sa ? B.Q[xc] * B.ev[xc] * Vn * s : 0.0;
}

__device__ static MixAvg mix_reduce(const double (*t)[10], int nj, const Gas &g, int sa)
{
    const double Gamma = g.gamma, Rgas = g.rgas;
    double s_sum = 0.0, nsx = 0.0, nsy = 0.0, nsz = 0.0, I[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nj; ++j) {
        nsx = nsx + t[j][1];
        nsy = nsy + t[j][2];
        nsz = nsz + t[j][3];
        s_sum = s_sum + t[j][0];
        for (int m = 1; m <= 6; ++m) I[m] = I[m] + t[j][3 + m];
    }
    MixAvg a;
    for (int m = 1; m <= 5; ++m) I[m] = I[m] / s_sum;
    a.ev = 0.0;
    if (sa) {
        I[6] = I[6] / s_sum;
        a.ev = I[6] / I[1];
    }
    const double nx = nsx / s_sum, ny = nsy / s_sum, nz = nsz / s_sum;
    const double tA = I[2] * nx + I[3] * ny + I[4] * nz;
    const double tB = F32C(2.0) * I[1] * I[5] - (I[2] * I[2] + I[3] * I[3] + I[4] * I[4]);
    const double tC = nx * nx + ny * ny + nz * nz;
    a.p = (tA / tC + sqrt(tA * tA / tC / tC - (Gamma * Gamma - F32C(1.0)) * tB / tC)) / (Gamma + F32C(1.0));
    a.u = (I[2] - a.p * nx) / I[1];
    a.v = (I[3] - a.p * ny) / I[1];
    a.w = (I[4] - a.p * nz) / I[1];
    a.d = I[1] / (a.u * nx + a.v * ny + a.w * nz);
    const double T = a.p / Rgas / a.d;
    a.c = sqrt(Gamma * Rgas * T);
    return a;
}

__device__ static void mix_ghost(const BlockDev &B, int s, int i, int j, int kg, double d, double u, double v,
                                 double w, double p, double ev, const Gas &g, int sa)
{
    const double Gamma = g.gamma, Rgas = g.rgas;
    const double pv[6] = {d, u, v, w, p, p / Rgas / d};
    const double V2 = u * u + v * v + w * w;
    const double cv[5] = {d, d * u, d * v, d * w, p / (Gamma - F32C(1.0)) + d * V2 / F32C(2.0)};
    const long x = cidx(B, i, j, kg), box = B.box;
    for (int c = 0; c < 5; ++c) B.Q[c * box + x] = cv[c];
    for (int c = 0; c < 6; ++c) B.W[c * box + x] = pv[c];
    if (sa) B.ev[x] = ev;
    set_dummies(B, s, i, j, cv, pv);
}

__global__ void __launch_bounds__(256) k_mixing_plane(const BlockDev *__restrict__ blocks,
                                                      const MixDev *__restrict__ mix, int nmix, Gas g, int sa)
{
    __shared__ double t[kMixMaxJ][10];
    __shared__ MixAvg avg[2];
    int q = 0;
    while (q + 1 < nmix && mix[q + 1].line0 <= (int)blockIdx.x) ++q;
    const MixDev M = mix[q];
    const int i = M.i0 + ((int)blockIdx.x - M.line0);
    const BlockDev &U = blocks[M.b1], &D = blocks[M.b2];
    const int KM = U.KM;
    for (int side = 0; side < 2; ++side) {
        const int nj = side ? M.nj2 : M.nj1, j0 = side ? M.j2 : M.j1;
        for (int r = threadIdx.x; r < nj; r += blockDim.x)
            mix_terms(side ? D : U, i, j0 + r, side ? 1 : KM, side ? 1 : KM - 1, sa, t[r]);
        __syncthreads();
        if (threadIdx.x == 0) avg[side] = mix_reduce(t, nj, g, sa);
        __syncthreads();
    }
    const MixAvg u = avg[0], d = avg[1];
    for (int r = threadIdx.x; r < M.nj1; r += blockDim.x) {   // lines 460-491
        const int j = M.j1 + r;
        const long x = cidx(U, i, j, KM - 1), box = U.box;
        const double pd = U.W[x], pu = U.W[box + x], pvv = U.W[2 * box + x], pw = U.W[3 * box + x],
                     pp = U.W[4 * box + x];
        const double C1 = -u.c * u.c * (pd - u.d) + (pp - u.p);
        const double C2 = u.d * u.c * (pw - u.w) + (pp - u.p);
        const double C5 = 0.0;
        mix_ghost(U, 5, i, j, KM, d.d + (-C1 + (C2 + C5) / F32C(2.0)) / (d.c * d.c), pu, pvv,
                  d.w + (C2 - C5) / (F32C(2.0) * d.d * d.c), d.p + (C2 + C5) / F32C(2.0), u.ev, g, sa);
    }
    for (int r = threadIdx.x; r < M.nj2; r += blockDim.x) {   // lines 494-522
        const int j = M.j2 + r;
        const long x = cidx(D, i, j, 1), box = D.box;
        const double pw = D.W[3 * box + x], pp = D.W[4 * box + x];
        const double C1 = 0.0, C2 = 0.0, C3 = 0.0, C4 = 0.0;
        const double C5 = (F32C(0.0) - d.d * d.c * (pw - d.w) + (pp - d.p));
        mix_ghost(D, 4, i, j, 0, u.d + (-C1 + (C2 + C5) / F32C(2.0)) / (u.c * u.c), u.u + C4 / (u.d * u.c),
                  u.v + C3 / (u.d * u.c), u.w - C5 / (F32C(2.0) * u.d * u.c), u.p + (C2 + C5) / F32C(2.0), u.ev, g,
                  sa);
    }
}

}  // namespace fv3d
