// Navier-Stokes terms of the LU-SGS iteration for gfx950: the
// coordinate-transformation coefficients (set-up), the face gradients and
// viscous face fluxes summed into the residual, the vorticity magnitude and
// the Baldwin-Lomax eddy viscosity.
//
// Arithmetic follows the reference's expression order (the library is built
// with FP contraction off), so results match the CPU oracle to the last bits
// except where libm pow rounds differently.
#include "fv3d_kernels.h"
#include "fv3d_tile.h"

#include <math.h>

namespace fv3d {

// ------------------------------------------------------- metric coefficients
// The nine coefficients from the coordinate derivatives along kesi (e1), yita
// (e2) and zita (e3) (source/GeometryVariables.cpp:191-210, repeated at
// 252-267, 310-326, 367-387).  gradz_kesi repeats the gradx_kesi expression in
// the reference; that is kept.  Coefficient m goes to out[m * stride].
__device__ static void metric9(const double e1[3], const double e2[3], const double e3[3], double *out, long stride)
{
    const double lim_zero = F32C(1E-20);
    double partA = (e1[0] * (e2[1] * e3[2] - e3[1] * e2[2]));
    double partB = (e2[0] * (e1[1] * e3[2] - e3[1] * e1[2]));
    double partC = (e3[0] * (e1[1] * e2[2] - e2[1] * e1[2]));
    double J1 = (partA - partB + partC);
    if (fabs(J1) <= lim_zero) J1 = lim_zero;
    const double J = F32C(1.0) / J1;
    out[0 * stride] = J * (e2[1] * e3[2] - e3[1] * e2[2]);
    out[1 * stride] = J * (e2[2] * e3[0] - e3[2] * e2[0]);
    out[2 * stride] = J * (e2[1] * e3[2] - e3[1] * e2[2]);
    out[3 * stride] = J * (e3[1] * e1[2] - e1[1] * e3[2]);
    out[4 * stride] = J * (e3[2] * e1[0] - e1[2] * e3[0]);
    out[5 * stride] = J * (e3[0] * e1[1] - e1[0] * e3[1]);
    out[6 * stride] = J * (e1[1] * e2[2] - e2[1] * e1[2]);
    out[7 * stride] = J * (e1[2] * e2[0] - e2[2] * e1[0]);
    out[8 * stride] = J * (e1[0] * e2[1] - e2[0] * e1[1]);
}

// One thread per (kind, box index): kind 0-2 the faces of direction kind,
// kind 3 the cells (source/GeometryVariables.cpp:108-391).
__global__ void k_metrics(const BlockDev *__restrict__ blocks, int block, double *fg0, double *fg1, double *fg2,
                          double *cg)
{
    const BlockDev &B = blocks[block];
    const long box = B.box;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * box) return;
    const int kind = (int)(t / box);
    const long x = t % box;
    const int i = (int)(x / B.si), j = (int)((x % B.si) / B.sj), k = (int)(x % B.sj);
    const int IM = B.IM, JM = B.JM, KM = B.KM;
    const int hi[4][3] = {{IM, JM - 1, KM - 1}, {IM - 1, JM, KM - 1}, {IM - 1, JM - 1, KM}, {IM - 1, JM - 1, KM - 1}};
    if (i < 1 || j < 1 || k < 1 || i > hi[kind][0] || j > hi[kind][1] || k > hi[kind][2]) return;
    const double *N = B.node;
    auto node = [&](int a, int b, int c, int q) { return N[q * box + cidx(B, a, b, c)]; };
    // 0.5*(n1 + n2) - 0.5*(n3 + n4), component q
    auto hp = [&](int a1, int b1, int c1, int a2, int b2, int c2, int a3, int b3, int c3, int a4, int b4, int c4, int q) {
        return F32C(0.5) * (node(a1, b1, c1, q) + node(a2, b2, c2, q)) -
               F32C(0.5) * (node(a3, b3, c3, q) + node(a4, b4, c4, q));
    };
    // one-sided at the first and last index, central between, of a centroid array
    auto d3 = [&](const double *P, long xm, long x0, long xp, bool first, bool last, int q) {
        if (first) return P[q * box + xp] - P[q * box + x0];
        if (last) return P[q * box + x0] - P[q * box + xm];
        return F32C(0.5) * (P[q * box + xp] - P[q * box + xm]);
    };
    double e1[3], e2[3], e3[3];
    if (kind == 3) {
        const double *C = B.cc;
        for (int q = 0; q < 3; ++q) {
            e1[q] = d3(C, x - B.si, x, x + B.si, i == 1, i == IM - 1, q);
            e2[q] = d3(C, x - B.sj, x, x + B.sj, j == 1, j == JM - 1, q);
            e3[q] = d3(C, x - 1, x, x + 1, k == 1, k == KM - 1, q);
        }
        metric9(e1, e2, e3, cg + x, box);
    } else if (kind == 0) {
        const double *F = B.fc[0];
        for (int q = 0; q < 3; ++q) {
            e1[q] = d3(F, x - B.si, x, i < IM ? x + B.si : x, i == 1, i == IM, q);
            e2[q] = hp(i, j + 1, k, i, j + 1, k + 1, i, j, k, i, j, k + 1, q);
            // the reference averages Node[i][j][k] with itself here (line 247)
            e3[q] = hp(i, j, k + 1, i, j + 1, k + 1, i, j, k, i, j, k, q);
        }
        metric9(e1, e2, e3, fg0 + x, box);
    } else if (kind == 1) {
        const double *F = B.fc[1];
        for (int q = 0; q < 3; ++q) {
            e1[q] = hp(i + 1, j, k, i + 1, j, k + 1, i, j, k, i, j, k + 1, q);
            e2[q] = d3(F, x - B.sj, x, j < JM ? x + B.sj : x, j == 1, j == JM, q);
            e3[q] = hp(i, j, k + 1, i + 1, j, k + 1, i, j, k, i + 1, j, k, q);
        }
        metric9(e1, e2, e3, fg1 + x, box);
    } else {
        const double *F = B.fc[2];
        for (int q = 0; q < 3; ++q) {
            e1[q] = hp(i + 1, j + 1, k, i + 1, j, k, i, j, k, i, j + 1, k, q);
            e2[q] = hp(i + 1, j + 1, k, i, j + 1, k, i + 1, j, k, i, j, k, q);
            e3[q] = d3(F, x - 1, x, k < KM ? x + 1 : x, k == 1, k == KM, q);
        }
        metric9(e1, e2, e3, fg2 + x, box);
    }
}

// Viscous flux of one face (source/ViscousFlux.cpp:34-108); without a
// turbulence model mue_tur = cond_tur = 0 (lines 68-72).
template <int D, class TL>
__device__ static void face_fv(const double *sw, int l, int i, int j, int k, const BlockDev &B, const Gas &g,
                               double out[5], double *gstore)
{
    constexpr int NH = TL::NH;
    const int st = D == 0 ? TL::HJ * TL::HK : D == 1 ? TL::HK : 1;
    double dv[3][4];
    face_dv<D, TL>(sw, l, i, j, k, B, dv);
    const long box = B.box, x = cidx(B, i, j, k);
    const double *G = B.fg[D] + x;
    double gm[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) gm[m] = G[m * box];
    double gr[3][4];   // d/dx, d/dy, d/dz of u v w T
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) gr[c][q] = dv[0][q] * gm[c] + dv[1][q] * gm[3 + c] + dv[2][q] * gm[6 + c];
    if (gstore)
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < 4; ++q) gstore[(c * 4 + q) * box] = gr[c][q];
    const double Gamma = g.gamma, Rgas = g.rgas;
    const double uf = F32C(0.5) * (sw[0 * NH + l] + sw[0 * NH + l - st]);
    const double vf = F32C(0.5) * (sw[1 * NH + l] + sw[1 * NH + l - st]);
    const double wf = F32C(0.5) * (sw[2 * NH + l] + sw[2 * NH + l - st]);
    const double tf = F32C(0.5) * (sw[3 * NH + l] + sw[3 * NH + l - st]);
    const double mue_lam = sutherland(g, tf);
    const double cond_lam = mue_lam * Gamma * Rgas / (Gamma - F32C(1.0)) / g.pr_lam;
    double mue_tur = 0.0, cond_tur = 0.0;
    if (g.turbulence) {
        // face average of the eddy viscosity (source/ViscousFlux.cpp:59-62)
        mue_tur = (B.mut[x - (D == 0 ? B.si : D == 1 ? B.sj : 1)] + B.mut[x]) / F32C(2.0);
        cond_tur = mue_tur * Gamma * Rgas / (Gamma - F32C(1.0)) / g.pr_tur;
    }
    const double mue = mue_lam + mue_tur, cond = cond_lam + cond_tur;
    const double c43 = (double)(4.0f / 3.0f), c23 = (double)(2.0f / 3.0f);
    const double *gx = gr[0], *gy = gr[1], *gz = gr[2];
    const double txx = mue * (c43 * gx[0] - c23 * (gy[1] + gz[2]));
    const double txy = mue * (gy[0] + gx[1]);
    const double txz = mue * (gz[0] + gx[2]);
    const double tyy = mue * (c43 * gy[1] - c23 * (gx[0] + gz[2]));
    const double tyz = mue * (gz[1] + gy[2]);
    const double tzz = mue * (c43 * gz[2] - c23 * (gx[0] + gy[1]));
    const double F5x = uf * txx + vf * txy + wf * txz + cond * gx[3];
    const double F5y = uf * txy + vf * tyy + wf * tyz + cond * gy[3];
    const double F5z = uf * txz + vf * tyz + wf * tzz + cond * gz[3];
    const double *S = B.fS[D];
    const double Sx = S[x], Sy = S[box + x], Sz = S[2 * box + x];
    out[0] = 0.0 * Sx + 0.0 * Sy + 0.0 * Sz;
    out[1] = txx * Sx + txy * Sy + txz * Sz;
    out[2] = txy * Sx + tyy * Sy + tyz * Sz;
    out[3] = txz * Sx + tyz * Sy + tzz * Sz;
    out[4] = F5x * Sx + F5y * Sy + F5z * Sz;
}

template <int TI, int TJ, int TK>
__global__ void __launch_bounds__(TI *TJ *TK) k_viscous(const BlockDev *__restrict__ blocks, const int4 *__restrict__ tiles,
                                                        Gas g, int store, ViscDebug dbg)
{
    using TL = VTile<TI, TJ, TK>;
    constexpr int NT = TI * TJ * TK, NH = TL::NH;
    __shared__ double sw[4 * NH];
    // momentum and energy components of the face fluxes; the continuity
    // component of the viscous flux is zero on every face (F1 = 0,
    // source/ViscousFlux.cpp:150-160), so it takes no LDS and three
    // workgroups fit on a CU instead of two
    __shared__ double sf[4 * TL::NF];
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;   // padding of the XCD-ordered tile list
    const BlockDev &B = blocks[tl.x];
    const long box = B.box;
    const int t = threadIdx.x;
    // u v w T of the tile's cells and halo, clamped into the cell box
    for (int n = t; n < NH; n += NT) {
        const int hk = n % TL::HK, hj = (n / TL::HK) % TL::HJ, hi = n / (TL::HK * TL::HJ);
        const int ci = min(tl.y + hi, B.IM), cj = min(tl.z + hj, B.JM), ck = min(tl.w + hk, B.KM);
        const long x = cidx(B, ci, cj, ck);
        sw[0 * NH + n] = B.W[1 * box + x];
        sw[1 * NH + n] = B.W[2 * box + x];
        sw[2 * NH + n] = B.W[3 * box + x];
        sw[3 * NH + n] = B.W[5 * box + x];
    }
    __syncthreads();
    // every face of the tile once: I faces (TI+1)xTJxTK, then J, then K
    for (int f = t; f < TL::NF; f += NT) {
        int D, a, b, c;
        if (f < TL::NFI) {
            D = 0; c = f % TK; b = (f / TK) % TJ; a = f / (TK * TJ);
        } else if (f < TL::NFI + TL::NFJ) {
            const int r = f - TL::NFI;
            D = 1; c = r % TK; b = (r / TK) % (TJ + 1); a = r / (TK * (TJ + 1));
        } else {
            const int r = f - TL::NFI - TL::NFJ;
            D = 2; c = r % (TK + 1); b = (r / (TK + 1)) % TJ; a = r / ((TK + 1) * TJ);
        }
        // higher-index cell of the face
        const int i = tl.y + a + 1, j = tl.z + b + 1, k = tl.w + c + 1;
        const bool ok = i <= B.ni + (D == 0) && j <= B.nj + (D == 1) && k <= B.nk + (D == 2);
        if (!ok) continue;
        const int l = TL::h(a + 1, b + 1, c + 1);
        double o[5];
        const long x = cidx(B, i, j, k);
        double *gs = store ? dbg.grad[D] + x : nullptr;
        if (D == 0) face_fv<0, TL>(sw, l, i, j, k, B, g, o, gs);
        else if (D == 1) face_fv<1, TL>(sw, l, i, j, k, B, g, o, gs);
        else face_fv<2, TL>(sw, l, i, j, k, B, g, o, gs);
#pragma unroll
        for (int q = 1; q < 5; ++q) sf[(q - 1) * TL::NF + f] = o[q];
        if (store)
            for (int q = 0; q < 5; ++q) dbg.fv[D][q * box + x] = o[q];
    }
    __syncthreads();
    // cell sums (source/ViscousFlux.cpp:268-280) and Res = Fc - Fv - Fd with Fd = 0
    // (source/ComputeAllBlockVars.cpp:245)
    const int tk = t % TK, tj = (t / TK) % TJ, ti = t / (TK * TJ);
    const int i = tl.y + ti + 1, j = tl.z + tj + 1, k = tl.w + tk + 1;
    if (i > B.ni || j > B.nj || k > B.nk) return;
    const int fI0 = (ti * TJ + tj) * TK + tk, fI1 = ((ti + 1) * TJ + tj) * TK + tk;
    const int fJ0 = TL::NFI + (ti * (TJ + 1) + tj) * TK + tk, fJ1 = TL::NFI + (ti * (TJ + 1) + tj + 1) * TK + tk;
    const int fK0 = TL::NFI + TL::NFJ + (ti * TJ + tj) * (TK + 1) + tk, fK1 = fK0 + 1;
    const long x = cidx(B, i, j, k);
    // density: R - 0 - 0 leaves R as the convective kernel wrote it
    if (store) dbg.fvc[x] = 0.0;
#pragma unroll
    for (int q = 1; q < 5; ++q) {
        const double *F = sf + (q - 1) * TL::NF;
        double s = F[fI1] - F[fI0];
        s = s + F[fJ1];
        s = s - F[fJ0];
        s = s + F[fK1];
        s = s - F[fK0];
        if (store) dbg.fvc[q * box + x] = s;
        B.R[q * box + x] = (B.R[q * box + x] - s) - 0.0;
    }
}

#define FV3D_VISC(S, A, B, C) template __global__ void k_viscous<A, B, C>(const BlockDev *, const int4 *, Gas, int, ViscDebug);
FV3D_FOR_TILES(FV3D_VISC)
#undef FV3D_VISC

// ---------------------------------------------------------------- vorticity
// CFluidProps::Curl (source/Spectral_curl_grad.cpp:44-83), one thread per cell.
__global__ void k_curl(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart, int nblock)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for_cell_item(blocks, cstart, nblock, t, [&](const BlockDev &B, int, int i, int j, int k) {
        const long box = B.box, x = cidx(B, i, j, k);
        const long st[3] = {B.si, B.sj, 1};
        double d[3][3];   // d/dkesi, d/dyita, d/dzita of u v w
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 3; ++q) d[a][q] = F32C(0.5) * (B.W[(1 + q) * box + x + st[a]] - B.W[(1 + q) * box + x - st[a]]);
        double gm[9];
        for (int m = 0; m < 9; ++m) gm[m] = B.cg[m * box + x];
        const double v_x = d[0][1] * gm[0] + d[1][1] * gm[3] + d[2][1] * gm[6];
        const double w_x = d[0][2] * gm[0] + d[1][2] * gm[3] + d[2][2] * gm[6];
        const double u_y = d[0][0] * gm[1] + d[1][0] * gm[4] + d[2][0] * gm[7];
        const double w_y = d[0][2] * gm[1] + d[1][2] * gm[4] + d[2][2] * gm[7];
        const double u_z = d[0][0] * gm[2] + d[1][0] * gm[5] + d[2][0] * gm[8];
        const double v_z = d[0][1] * gm[2] + d[1][1] * gm[5] + d[2][1] * gm[8];
        B.omega[x] = sqrt(F32C(0.5) * ((w_y - v_z) * (w_y - v_z) + (u_z - w_x) * (u_z - w_x) + (v_x - u_y) * (v_x - u_y)));
    });
}

// ------------------------------------------------------------- Baldwin-Lomax
// CBlocks::BL_Model (source/BL_Model.cpp:8-279) with CBlocks::BL_model_1d
// (283-329).  The reference visits the walls of a block in table order: a
// cell's first visit sets mue_t, later walls keep the minimum, and each wall
// zeroes its first cell.  Every value is non-negative (mue_in and mue_out are
// products of non-negative factors), so that is the minimum over the visiting
// walls whatever the order, and the first cell's zero wins it; all walls of
// all blocks therefore run in one launch and fold their values into blmin with
// an atomic minimum on the bits (ordered as the values for non-negative
// doubles).  Cells that no wall line visits keep mue_t (k_bl_merge).  A
// workgroup serves kBlLines grid lines normal to the wall that
// are neighbours along their fastest tangential index (so a load of one cell
// position is contiguous over the lines) and splits each line into G groups;
// thread (group g, line l) holds the cells n0 = 1 + g + G c, c < kBlCpt, of its
// line in registers.  The reference's sequential maxima become reductions
// over the groups in LDS: tau_w and U_dif are plain maxima (NaN entries never
// win, as with its strict comparisons; a NaN first entry is kept), and
// (F_max, y_max) is the largest positive F with ties going to the smallest n0,
// which is the first one its loop meets.  The first cell ends at zero; the
// ghost cell gets zero and then the negated first-cell value, as WallFlux
// writes it (source/WallFlux.cpp:106-110).
template <int G>
__global__ void __launch_bounds__(kBlLines *G) k_bl_lines(const BlockDev *__restrict__ blocks,
                                                          const PatchDev *__restrict__ patches,
                                                          const int2 *__restrict__ walls, int nwalls, Gas g)
{
    constexpr int L = kBlLines, C = kBlCpt;
    __shared__ double s_a[G][L], s_b[G][L];
    __shared__ int s_n[G][L];
    __shared__ double s_line[6][L];   // Tw, Udif, Ret, Fmax, y_max, Fwake
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    // the wall and the lines of this workgroup (walls[w].y is its line count,
    // rounded up to whole groups of L lines by the host)
    long t = (long)blockIdx.x * L;
    int w = 0;
    while (w < nwalls && t >= walls[w].y) { t -= walls[w].y; ++w; }
    const PatchDev P = patches[walls[min(w, nwalls - 1)].x];
    const BlockDev &B = blocks[P.block];
    const long line = t + lane;
    const bool active = w < nwalls && line < (long)P.na * P.nb;
    const int a = P.a0 + (int)(active ? line / P.nb : 0), cc = P.b0 + (int)(active ? line % P.nb : 0);
    const int D = P.dir, hi = P.hi;
    const int N = D == 0 ? B.IM : D == 1 ? B.JM : B.KM;
    const int M = D == 0 ? B.IM : B.JM;   // line length of the reference's call (K walls need JM == KM)
    const long box = B.box;
    auto cell = [&](int n0) -> long {
        const int n = hi ? N - n0 : n0;
        return D == 0 ? cidx(B, n, a, cc) : D == 1 ? cidx(B, a, n, cc) : cidx(B, a, cc, n);
    };
    const long xw = D == 0 ? cidx(B, hi ? N : 1, a, cc) : D == 1 ? cidx(B, a, hi ? N : 1, cc) : cidx(B, a, cc, hi ? N : 1);
    const double *FC = B.fc[D];
    const double wx = FC[xw], wy = FC[box + xw], wz = FC[2 * box + xw];
    const double A_plus = F32C(26.0), Cwk = F32C(1.0), kap = F32C(0.4), Ckleb = F32C(0.3), alpha = F32C(0.0168),
                 Ccp = F32C(1.6);
    double dis[C], om[C], dens[C], damp[C];
    bool ok[C];
    double tw = -1.0, ud = 0.0, t1 = 0.0, d1 = 0.0, m1 = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int n0 = 1 + grp + G * c;
        ok[c] = active && n0 < M;
        const long x = cell(ok[c] ? n0 : 1);
        const double dx = B.cc[x] - wx, dy = B.cc[box + x] - wy, dz = B.cc[2 * box + x] - wz;
        // POW(x, 2.0) of the reference compiles to x*x
        dis[c] = sqrt(dx * dx + dy * dy + dz * dz);
        om[c] = B.omega[x];
        dens[c] = B.W[x];
        const double ml = sutherland(g, B.W[5 * box + x]);
        const double tv = fabs(om[c] * ml);
        const double u = B.W[box + x], v = B.W[2 * box + x], ww = B.W[3 * box + x];
        const double vv = sqrt(u * u + v * v + ww * ww);
        if (ok[c] && tv > tw) tw = tv;
        if (ok[c] && vv > ud) ud = vv;
        if (c == 0 && grp == 0) {
            t1 = tv;
            d1 = dens[0];
            m1 = ml;
        }
    }
    s_a[grp][lane] = tw;
    s_b[grp][lane] = ud;
    __syncthreads();
    if (grp == 0) {
        double Tw = t1, Udif = 0.0;
        for (int q = 0; q < G; ++q) {
            if (s_a[q][lane] > Tw) Tw = s_a[q][lane];
            if (s_b[q][lane] > Udif) Udif = s_b[q][lane];
        }
        s_line[0][lane] = Tw;
        s_line[1][lane] = Udif;
        s_line[2][lane] = sqrt(d1 * Tw) / m1;
    }
    __syncthreads();
    const double Ret = s_line[2][lane];
    // largest positive F, smallest n0 on ties
    double bF = 0.0, bY = 0.0;
    int bN = 1 << 30;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double y_plus = dis[c] * Ret;
        damp[c] = F32C(1.0) - exp(-y_plus / A_plus);
        const double F = dis[c] * fabs(om[c]) * damp[c];
        const int n0 = 1 + grp + G * c;
        if (ok[c] && F > 0.0 && (F > bF || (F == bF && n0 < bN))) {
            bF = F;
            bY = dis[c];
            bN = n0;
        }
    }
    s_a[grp][lane] = bF;
    s_b[grp][lane] = bY;
    s_n[grp][lane] = bN;
    __syncthreads();
    if (grp == 0) {
        for (int q = 1; q < G; ++q)
            if (s_a[q][lane] > bF || (s_a[q][lane] == bF && s_n[q][lane] < bN)) {
                bF = s_a[q][lane];
                bY = s_b[q][lane];
                bN = s_n[q][lane];
            }
        const double Udif = s_line[1][lane];
        const double wa = bY * bF, wb = Cwk * bY * Udif * Udif / bF;
        s_line[3][lane] = bF;
        s_line[4][lane] = bY;
        s_line[5][lane] = wa < wb ? wa : wb;
    }
    __syncthreads();
    const double y_max = s_line[4][lane], Fwake = s_line[5][lane];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (!ok[c]) continue;
        const int n0 = 1 + grp + G * c;
        const double bL = kap * dis[c] * damp[c];
        const double mue_in = dens[c] * bL * bL * om[c];
        const double Fkleb = F32C(1.0) / (F32C(1.0) + F32C(5.5) * pow((Ckleb * dis[c] / y_max), 6.0));
        const double mue_out = alpha * Ccp * dens[c] * Fwake * Fkleb;
        const double mt = fabs(mue_in) > fabs(mue_out) ? mue_out : mue_in;
        const long x = cell(n0);
        // the first cell ends at zero (source/BL_Model.cpp:101-118)
        const double v = n0 == 1 ? 0.0 : mt;
        atomicMin(B.blmin + x, (unsigned long long)__double_as_longlong(v));
    }
    if (active && grp == 0) {
        const long xg = D == 0 ? cidx(B, hi ? N : 0, a, cc) : D == 1 ? cidx(B, a, hi ? N : 0, cc) : cidx(B, a, cc, hi ? N : 0);
        B.mut[xg] = -0.0;
    }
}

// mue_t of every visited interior cell from the minimum over its walls; the
// minimum resets to "not visited" for the next evaluation.
__global__ void __launch_bounds__(256) k_bl_merge(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                  int nblock)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long x = cidx(B, i, j, k);
    const unsigned long long v = B.blmin[x];
    if (v != kBlUnvisited) {
        B.mut[x] = __longlong_as_double((long long)v);
        B.blmin[x] = kBlUnvisited;
    }
}

template __global__ void k_bl_lines<16>(const BlockDev *, const PatchDev *, const int2 *, int, Gas);
template __global__ void k_bl_lines<32>(const BlockDev *, const PatchDev *, const int2 *, int, Gas);
template __global__ void k_bl_lines<64>(const BlockDev *, const PatchDev *, const int2 *, int, Gas);

}  // namespace fv3d
