// Spalart-Allmaras turbulence model of the LU-SGS iteration for gfx950: the
// wall distance (set-up), the residual of the nu~ equation, the eddy
// viscosity, the nu~ part of the LU-SGS sweeps and of the variable limiter.
//
// Arithmetic follows the reference's expression order (the library is built
// with FP contraction off); results match the CPU oracle to the last bits
// except where the device pow/exp/sqrt round differently from libm.
#include "fv3d_kernels.h"
#include "fv3d_tile.h"

#include <math.h>

namespace fv3d {

// Model constants of CSpaceDiscr::SA_Model (source/SA_Model.cpp:42-44): float
// literals stored in doubles, the derived Cw1 formed in double.
struct SaConst {
    double Cb1 = F32C(0.1355), Cb2 = F32C(0.622), Cv1 = F32C(7.1), sigma = (double)(2.0f / 3.0f), kavf = F32C(0.41);
    double Cw2 = F32C(0.3), Cw3 = F32C(2.0), Ct3 = F32C(1.2), Ct4 = F32C(0.5);
    __device__ double Cw1() const { return Cb1 / kavf / kavf + (F32C(1.0) + Cb2) / sigma; }
};

__device__ static inline double sa_fv1(double X, double Cv1)
{
    return pow(X, F32C(3.0)) / (pow(X, F32C(3.0)) + Cv1 * Cv1 * Cv1);
}

// nu~ of an interior cell after the clipping of SA_Model (source/SA_Model.cpp:
// 64-87): negative values become zero, values above 200 times the laminar
// kinematic viscosity are capped.  The face fluxes and the source see this
// value, the face gradients the value before it (they are formed earlier,
// in CFluidProps::Gradient).
__device__ static inline double sa_clip(double ev, double temp, double dens, const Gas &g)
{
    if (ev < 0.0) ev = 0.0;
    const double mue_lam = sutherland(g, temp);
    const double ev_lam = g.re_ref * mue_lam / dens;
    const double Max = 200.0;
    if (ev_lam > 0 && ev > (Max * ev_lam)) ev = Max * ev_lam;
    return ev;
}

// ----------------------------------------------------------- wall distance
// CMesh::WallDistanceSA (source/WallDistanceSA.cpp:9-132): for every cell the
// smallest distance from its centroid to the centroid of any wall face, from
// 10000.  wf holds the wall-face centroids (x, y, z) of every instance's
// walls; a block takes those of its own instance (BlockDev::wf0, nwf); each
// workgroup stages 256 of them at a time in LDS (the instances of the
// workgroup's first and last cells bound the range it stages).  FP64 throughout: this is the arithmetic-bound
// set-up kernel (cells x wall faces distance evaluations).  The minimum is
// taken over the squared distances and rooted once: the correctly rounded
// square root is monotone, so sqrt(min(d^2, 10000^2)) is the reference's
// running minimum of the roots started from 10000.
__global__ void __launch_bounds__(256) k_wall_distance(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                       int nblock, const double *__restrict__ wf)
{
    __shared__ double sx[256], sy[256], sz[256];
    __shared__ int s_lo, s_hi;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n = 0, i, j, k;
    const bool valid = cell_item(blocks, cstart, nblock, t, n, i, j, k);
    double cx = 0.0, cy = 0.0, cz = 0.0;
    long x = 0;
    int f0 = 0, f1 = 0;   // this cell's wall faces
    if (valid) {
        const BlockDev &B = blocks[n];
        x = cidx(B, i, j, k);
        cx = B.cc[x]; cy = B.cc[B.box + x]; cz = B.cc[2 * B.box + x];
        f0 = B.wf0;
        f1 = B.wf0 + B.nwf;
    }
    // the union of the workgroup's face ranges (one range unless the
    // workgroup spans two instances)
    if (threadIdx.x == 0) { s_lo = 0x7fffffff; s_hi = 0; }
    __syncthreads();
    if (valid && f1 > f0) {
        atomicMin(&s_lo, f0);
        atomicMax(&s_hi, f1);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    double d2min = 10000.0 * 10000.0;
    for (int base = lo; base < hi; base += 256) {
        __syncthreads();
        const int q = base + (int)threadIdx.x;
        if (q < hi) {
            sx[threadIdx.x] = wf[3 * (long)q];
            sy[threadIdx.x] = wf[3 * (long)q + 1];
            sz[threadIdx.x] = wf[3 * (long)q + 2];
        }
        __syncthreads();
        const int m = min(256, hi - base);
        for (int e = 0; e < m; ++e) {
            const int q2 = base + e;
            const double dx = cx - sx[e], dy = cy - sy[e], dz = cz - sz[e];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (q2 >= f0 && q2 < f1) d2min = fmin(d2min, d2);
        }
    }
    if (valid) blocks[n].dwall[x] = sqrt(d2min);
}

// ------------------------------------------------------------ nu~ residual
// CSpaceDiscr::SA_Model (source/SA_Model.cpp:90-210) with the convective cell
// sum of CMesh::ConvectiveFlux (source/ComputeAllBlockVars.cpp:109-114), the
// WallFlux zero of the nu~ convection through walls (source/WallFlux.cpp:105,
// 136, 167) and Res_ev = Fc_tur - Fv_tur (source/ComputeAllBlockVars.cpp:249).
// A tile stages u v w T, nu~ (as it stands and clipped) and rho of its cells and
// a one-cell halo in LDS; every face of the tile is evaluated once: the face
// gradient of nu~ (the shared 33-case rule), the upwind convection, the
// diffusion (1 + c_b2)/sigma (nu + nu~) grad nu~ . S / Re and the plain
// gradient flux grad nu~ . S.  The reference reads its face gradients through
// aliased pointers (source/SA_Model.cpp:21-32): all three components are the
// x derivative on I faces, and the y and z derivatives trade places on K
// faces; that is kept.  The cell part adds the source.
template <int TI, int TJ, int TK>
__global__ void __launch_bounds__(TI *TJ *TK) k_sa_faces(const BlockDev *__restrict__ blocks, const int4 *__restrict__ tiles,
                                                         Gas g)
{
    using TL = VTile<TI, TJ, TK>;
    constexpr int NT = TI * TJ * TK, NH = TL::NH;
    // planes: u v w T, nu~ as it stands, rho, nu~ clipped
    __shared__ double sw[7 * NH];
    __shared__ double sf[3 * TL::NF];   // convective, viscous, plain gradient flux
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;   // padding of the XCD-ordered tile list
    const BlockDev &B = blocks[tl.x];
    const long box = B.box;
    const int t = threadIdx.x;
    const SaConst K;
    const double Re = g.re_ref;
    for (int n = t; n < NH; n += NT) {
        const int hk = n % TL::HK, hj = (n / TL::HK) % TL::HJ, hi = n / (TL::HK * TL::HJ);
        const int ci = min(tl.y + hi, B.IM), cj = min(tl.z + hj, B.JM), ck = min(tl.w + hk, B.KM);
        const long x = cidx(B, ci, cj, ck);
        const double T = B.W[5 * box + x], rho = B.W[x], ev = B.ev[x];
        sw[0 * NH + n] = B.W[1 * box + x];
        sw[1 * NH + n] = B.W[2 * box + x];
        sw[2 * NH + n] = B.W[3 * box + x];
        sw[3 * NH + n] = T;
        sw[4 * NH + n] = ev;
        sw[5 * NH + n] = rho;
        const bool interior = ci >= 1 && ci <= B.ni && cj >= 1 && cj <= B.nj && ck >= 1 && ck <= B.nk;
        sw[6 * NH + n] = interior ? sa_clip(ev, T, rho, g) : ev;
    }
    __syncthreads();
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
        const int i = tl.y + a + 1, j = tl.z + b + 1, k = tl.w + c + 1;
        const bool ok = i <= B.ni + (D == 0) && j <= B.nj + (D == 1) && k <= B.nk + (D == 2);
        if (!ok) continue;
        const int l = TL::h(a + 1, b + 1, c + 1);
        const int st = D == 0 ? TL::HJ * TL::HK : D == 1 ? TL::HK : 1;
        const int lm = l - st;
        double dv[3][1];
        if (D == 0) face_dv<0, TL, 1>(sw + 4 * NH, l, i, j, k, B, dv);
        else if (D == 1) face_dv<1, TL, 1>(sw + 4 * NH, l, i, j, k, B, dv);
        else face_dv<2, TL, 1>(sw + 4 * NH, l, i, j, k, B, dv);
        const long x = cidx(B, i, j, k);
        const double *G = B.fg[D] + x;
        double gr[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) gr[q] = dv[0][0] * G[q * box] + dv[1][0] * G[(3 + q) * box] + dv[2][0] * G[(6 + q) * box];
        const double gx = gr[0], gy = D == 0 ? gr[0] : D == 1 ? gr[1] : gr[2], gz = D == 0 ? gr[0] : D == 1 ? gr[2] : gr[1];
        const double densf = F32C(0.5) * (sw[5 * NH + l] + sw[5 * NH + lm]);
        const double tempf = F32C(0.5) * (sw[3 * NH + l] + sw[3 * NH + lm]);
        const double evf = F32C(0.5) * (sw[6 * NH + l] + sw[6 * NH + lm]);
        const double ev_lam = Re * sutherland(g, tempf) / densf;
        const double tx = (F32C(1.0) + K.Cb2) / K.sigma * (ev_lam + evf) * gx;
        const double ty = (F32C(1.0) + K.Cb2) / K.sigma * (ev_lam + evf) * gy;
        const double tz = (F32C(1.0) + K.Cb2) / K.sigma * (ev_lam + evf) * gz;
        const double *S = B.fS[D];
        const double Sx = S[x], Sy = S[box + x], Sz = S[2 * box + x], s = B.fs[D][x];
        const double nx = Sx / s, ny = Sy / s, nz = Sz / s;
        const double Vn1 = sw[lm] * nx + sw[NH + lm] * ny + sw[2 * NH + lm] * nz;
        const double Vn2 = sw[l] * nx + sw[NH + l] * ny + sw[2 * NH + l] * nz;
        double fc = (F32C(0.5) * (Vn1 + fabs(Vn1)) * sw[6 * NH + lm] + F32C(0.5) * (Vn2 - fabs(Vn2)) * sw[6 * NH + l]) * s;
        // a wall face carries no nu~ convection
        const int p = D == 0 ? i : D == 1 ? j : k, N = D == 0 ? B.IM : D == 1 ? B.JM : B.KM;
        if (p == 1 || p == N) {
            const int side = 2 * D + (p == N);
            const int ta = D == 0 ? j : i, tb = D == 2 ? j : k;
            const unsigned char *wm = B.wmask[side];
            if (wm && wm[(long)ta * B.Db[side] + tb]) fc = 0;
        }
        sf[f] = fc;
        sf[TL::NF + f] = (tx * Sx + ty * Sy + tz * Sz) / Re;
        sf[2 * TL::NF + f] = gx * Sx + gy * Sy + gz * Sz;
    }
    __syncthreads();
    const int tk = t % TK, tj = (t / TK) % TJ, ti = t / (TK * TJ);
    const int i = tl.y + ti + 1, j = tl.z + tj + 1, k = tl.w + tk + 1;
    if (i > B.ni || j > B.nj || k > B.nk) return;
    const int fI0 = (ti * TJ + tj) * TK + tk, fI1 = ((ti + 1) * TJ + tj) * TK + tk;
    const int fJ0 = TL::NFI + (ti * (TJ + 1) + tj) * TK + tk, fJ1 = TL::NFI + (ti * (TJ + 1) + tj + 1) * TK + tk;
    const int fK0 = TL::NFI + TL::NFJ + (ti * TJ + tj) * (TK + 1) + tk, fK1 = fK0 + 1;
    const int l0 = TL::h(ti + 1, tj + 1, tk + 1);
    const long x = cidx(B, i, j, k);
    const double ev = sw[6 * NH + l0], T = sw[3 * NH + l0], rho = sw[5 * NH + l0];
    const double ev_lam = Re * sutherland(g, T) / rho;
    const double X = ev / ev_lam;
    const double fv1 = sa_fv1(X, K.Cv1);
    const double fv2 = F32C(1.0) - X / (F32C(1.0) + X * fv1);
    const double ft2 = K.Ct3 * exp(F32C(0.0) - K.Ct4 * X * X);
    const double Om = B.omega[x], dis = B.dwall[x];
    const double kd = K.kavf * dis;   // POW(kavf*dis, 2.0f) of the reference compiles to a square
    const double sa = Om + ev * fv2 / ((kd * kd) * Re), sb = F32C(0.3) * Om;
    const double Sbar = sa > sb ? sa : sb;
    const double ra = ev / (Sbar * K.kavf * K.kavf * dis * dis * Re), rb = F32C(10.0);
    const double r = ra < rb ? ra : rb;
    const double gg = r + K.Cw2 * (pow(r, F32C(6.0)) - r);
    const double fw = gg * pow((1 + pow(K.Cw3, F32C(6.0))) / (pow(gg, F32C(6.0)) + pow(K.Cw3, F32C(6.0))),
                               (double)(1.0f / 6.0f));
    const double ed = ev / dis;
    const double Qc = (K.Cb1 * (F32C(1.0) - ft2) * Sbar * ev -
                       (K.Cw1() * fw - K.Cb1 * ((1 - ft2) * fv2 + ft2) / (K.kavf * K.kavf)) * (ed * ed) / Re) * B.vol[x];
    const double *Fc = sf, *Fv = sf + TL::NF, *F1 = sf + 2 * TL::NF;
    const double sv = Qc + Fv[fI1] - Fv[fI0] + Fv[fJ1] - Fv[fJ0] + Fv[fK1] - Fv[fK0];
    const double s1 = F1[fI1] - F1[fI0] + F1[fJ1] - F1[fJ0] + F1[fK1] - F1[fK0];
    const double cfv = sv - K.Cb2 * (ev_lam + ev) / K.sigma / Re * s1;
    const double cfc = Fc[fI1] - Fc[fI0] + Fc[fJ1] - Fc[fJ0] + Fc[fK1] - Fc[fK0];
    B.res_ev[x] = cfc - cfv;
}

#define FV3D_SAF(S, A, B, C) template __global__ void k_sa_faces<A, B, C>(const BlockDev *, const int4 *, Gas);
FV3D_FOR_TILES(FV3D_SAF)
#undef FV3D_SAF

// ----------------------------------------------------------- eddy viscosity
// mue_t = fv1 rho nu~ / Re of the interior cells from nu~ as it stands, then
// the clipping of mue_t and of nu~ (source/SA_Model.cpp:47-87).  With save_old
// nu~ as it stood goes to ev_old first (CMesh::LU_SGS_Marching keeps it before
// its residual evaluation, source/LU_SGS_Marching.cpp:38-44).
__global__ void __launch_bounds__(256) k_sa_mut(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                int nblock, Gas g, int save_old)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int n, i, j, k;
    if (!cell_item(blocks, cstart, nblock, t, n, i, j, k)) return;
    const BlockDev &B = blocks[n];
    const long box = B.box, x = cidx(B, i, j, k);
    const SaConst K;
    const double Re = g.re_ref;
    const double T = B.W[5 * box + x], rho = B.W[x];
    double ev = B.ev[x];
    if (save_old) B.ev_old[x] = ev;
    const double ev_lam = Re * sutherland(g, T) / rho;
    double mt = sa_fv1(ev / ev_lam, K.Cv1) * rho * ev / Re;
    if (mt < 0.0) mt = 0.0;
    if (ev < 0.0) ev = 0.0;
    const double mue_lam = sutherland(g, T);
    const double ev_lam2 = Re * mue_lam / rho;
    const double Max = 200.0;
    if (mue_lam > 0 && mt > (Max * mue_lam)) mt = Max * mue_lam;
    if (ev_lam2 > 0 && ev > (Max * ev_lam2)) ev = Max * ev_lam2;
    B.mut[x] = mt;
    B.ev[x] = ev;
}

// mue_t of the ghost cells of the six faces from their own nu~
// (source/SA_Model.cpp:212-257); behind a wall WallFlux then mirrors the first
// cell's value (source/WallFlux.cpp:106-110).  gstart counts the ghost cells
// of the blocks (2 (nj nk + ni nk + ni nj) each).
__global__ void __launch_bounds__(256) k_sa_ghost(const BlockDev *__restrict__ blocks, const long *__restrict__ gstart,
                                                  int nblock, Gas g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= gstart[nblock]) return;
    const int n = find_range(gstart, nblock, t);
    const BlockDev &B = blocks[n];
    long r = t - gstart[n];
    const long nf[3] = {(long)B.nj * B.nk, (long)B.ni * B.nk, (long)B.ni * B.nj};
    int side = 0;
    while (r >= nf[side / 2]) { r -= nf[side / 2]; ++side; }
    const int D = side / 2, hi = side % 2;
    const int nb = D == 2 ? B.nj : B.nk;
    const int ta = 1 + (int)(r / nb), tb = 1 + (int)(r % nb);
    const int N = D == 0 ? B.IM : D == 1 ? B.JM : B.KM;
    const int pg = hi ? N : 0, pi = hi ? N - 1 : 1;
    auto at = [&](int p) { return D == 0 ? cidx(B, p, ta, tb) : D == 1 ? cidx(B, ta, p, tb) : cidx(B, ta, tb, p); };
    const long xg = at(pg);
    const unsigned char *wm = B.wmask[side];
    if (wm && wm[(long)ta * B.Db[side] + tb]) {
        B.mut[xg] = -B.mut[at(pi)];
        return;
    }
    const SaConst K;
    const long box = B.box;
    const double T = B.W[5 * box + xg], rho = B.W[xg], ev = B.ev[xg];
    const double ev_lam = g.re_ref * sutherland(g, T) / rho;
    B.mut[xg] = sa_fv1(ev / ev_lam, K.Cv1) * rho * ev / g.re_ref;
}

// ------------------------------------------------------------ LU-SGS of nu~
// The nu~ equation of CMesh::LU_SGS_Marching (source/LU_SGS_Marching.cpp:
// 79-157, 178-257), one hyperplane per launch.  It runs after the flow sweep of
// the same direction, whose increments of the neighbour cells it reads (the
// forward values in the forward sweep, the final ones in the backward sweep,
// as the reference's interleaved loop sees them).  Diagonal: Vol/dt +
// relax sum(lamc) + 3 sum(lamv); flux difference of the neighbour:
// (nu~ + dnu~) Vn_new - nu~ Vn_old over its two faces' mean normal.
template <bool FORWARD>
__global__ void k_sa_plane(const BlockDev *__restrict__ blocks, const long *__restrict__ pstart, int nblock, int step,
                           double relax, Gas g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
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
    const double alpha2 = B.alpha2[x];   // Vol/dt + relax*sum(lamc) + 3*sum(lamv), from k_spectral_dt
    const long st[3] = {B.si, B.sj, 1};
    const int pos[3] = {i, j, k};
    const int lim[3] = {B.ni, B.nj, B.nk};
    double Ae[3], Ve[3];
    for (int d = 0; d < 3; ++d) {
        const bool ok = FORWARD ? pos[d] != 1 : pos[d] != lim[d];
        if (!ok) { Ae[d] = 0.0; Ve[d] = 0.0; continue; }
        const long nb = FORWARD ? x - st[d] : x + st[d];
        const long fa = nb + st[d], fb = nb;   // the neighbour's upper face first
        const double *S = B.fS[d], *s = B.fs[d];
        const double nx = (S[fa] / s[fa] + S[fb] / s[fb]) / F32C(2.0);
        const double ny = (S[box + fa] / s[fa] + S[box + fb] / s[fb]) / F32C(2.0);
        const double nz = (S[2 * box + fa] / s[fa] + S[2 * box + fb] / s[fb]) / F32C(2.0);
        double o[4], nw[4];
        for (int c = 0; c < 4; ++c) { o[c] = B.Q[c * box + nb]; nw[c] = o[c] + B.dQ[c * box + nb]; }
        const double Vn_old = o[1] / o[0] * nx + o[2] / o[0] * ny + o[3] / o[0] * nz;
        const double Vn_new = nw[1] / nw[0] * nx + nw[2] / nw[0] * ny + nw[3] / nw[0] * nz;
        const double ev_old = B.ev[nb], ev_new = ev_old + B.dev[nb];
        const double dfe = ev_new * Vn_new - ev_old * Vn_old;
        const double sd = F32C(0.5) * (s[nb] + s[nb + st[d]]);
        const double rl = relax * B.lamc[d * box + nb] * B.dev[nb];
        Ae[d] = F32C(0.5) * (FORWARD ? dfe * sd + rl : dfe * sd - rl);
        Ve[d] = B.lamv[d * box + nb] * B.dev[nb];
    }
    if (FORWARD)
        B.dev[x] = (F32C(-1.0) * B.res_ev[x] + Ae[0] + Ae[1] + Ae[2] + Ve[0] + Ve[1] + Ve[2]) / alpha2;
    else
        B.dev[x] = (alpha2 * B.dev[x] - Ae[0] - Ae[1] - Ae[2] + Ve[0] + Ve[1] + Ve[2]) / alpha2;
}

template __global__ void k_sa_plane<true>(const BlockDev *, const long *, int, int, double, Gas);
template __global__ void k_sa_plane<false>(const BlockDev *, const long *, int, int, double, Gas);

// ------------------------------------------------------------- Limit_SA
// CFluidProps::Limit_SA (source/Limit_FlowVars.cpp:122-168).  Raising values
// below 1e-8 is pointwise; replacing a value above 1000 by the mean of its
// admissible neighbours depends on the neighbours' already limited values, so
// a block with such a value (k_update flags it) is limited by one thread in
// the reference's order here, and the floor of the other blocks is applied in
// k_sa_res.
__global__ void k_sa_limit_serial(const BlockDev *__restrict__ blocks, int nblock, int *__restrict__ hi_flag)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double SAmin = F32C(1.0E-8), LSAmax = F32C(1000.0);
    for (int n = 0; n < nblock; ++n) {
        if (!hi_flag[n]) continue;
        hi_flag[n] = 0;
        const BlockDev &B = blocks[n];
        const int IM = B.IM, JM = B.JM, KM = B.KM;
        for (int i = 1; i <= B.ni; ++i)
            for (int j = 1; j <= B.nj; ++j)
                for (int k = 1; k <= B.nk; ++k) {
                    double *e = B.ev + cidx(B, i, j, k);
                    if (*e < SAmin) *e = SAmin;
                    if (*e > LSAmax) {
                        double s0 = 0.0;
                        int kn = 0;
                        for (int ka = -1; ka <= 1; ++ka)
                            for (int ja = -1; ja <= 1; ++ja)
                                for (int ia = -1; ia <= 1; ++ia) {
                                    const int k1 = k + ka, j1 = j + ja, i1 = i + ia;
                                    if (i1 > 0 && i1 < IM && j1 > 0 && j1 < JM && k1 > 0 && k1 < KM) {
                                        const double v = B.ev[cidx(B, i1, j1, k1)];
                                        if (v >= SAmin && v <= LSAmax) { s0 = s0 + v; kn = kn + 1; }
                                    }
                                }
                        *e = kn == 0 ? LSAmax : s0 / kn;
                    }
                }
    }
}

// Res_ev of CMesh::Res_Calculate (source/ComputeAllBlockVars.cpp:352-357,
// 368-371): partial sums of |(nu~_old - nu~)/nu~_old| Vol, after the limiter
// as in the reference, into component 5 of the norm partials (same grid as
// k_update).  With clamp set (limiter on) nu~ below 1e-8 is raised to it
// first; blocks k_sa_limit_serial treated hold no such value any more.
__global__ void __launch_bounds__(256) k_sa_res(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart,
                                                int nblock, double *__restrict__ partial, int clamp)
{
    __shared__ double red[256];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    int n, i, j, k;
    if (cell_item(blocks, cstart, nblock, t, n, i, j, k)) {
        const BlockDev &B = blocks[n];
        const long x = cidx(B, i, j, k);
        double e = B.ev[x];
        if (clamp && e < F32C(1.0E-8)) {
            e = F32C(1.0E-8);
            B.ev[x] = e;
        }
        v = fabs((B.ev_old[x] - e) / B.ev_old[x]) * B.vol[x];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[5L * gridDim.x + blockIdx.x] = red[0];
}

}  // namespace fv3d
