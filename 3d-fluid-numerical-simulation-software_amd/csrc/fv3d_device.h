// Device-side data layout and per-face / per-cell physics of the LU-SGS
// iteration, written for gfx950 (CDNA4).
//
// Layout (per block, structure of arrays, k fastest):
//   cell-centred arrays cover the reference's index box [0..IM]x[0..JM]x[0..KM]
//   (ghost cells 0 and IM included), component c of field F at F[c*box + idx];
//   face-centred arrays use the same box (face i of the I direction at index i);
//   the NDummy=2 layers of each of the six sides live in their own slabs,
//   11 components (5 conservative then 6 primitive) per layer.
//
// All arithmetic follows the reference's expression order; the library is
// compiled with -ffp-contract=off so FP64 results match the CPU oracle except
// where libm pow (MUSCL3 limiter) rounds differently from a*a.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fv3d {

constexpr int kND = 2;        // Dummy layers for the upwind schemes (source/UserInput.cpp:132-155)
constexpr int kMaxBlocks = 4096;

// float literals of the reference widened to double (include/defs.h keeps
// many constants as float literals even when REAL is double)
#define F32C(x) ((double)(x##f))

struct BlockDev {
    int inst;                 // instance (independent passage) the block belongs to
    int IM, JM, KM;           // node counts
    int ni, nj, nk;           // cell counts = IM-1, JM-1, KM-1
    long box;                 // (IM+1)(JM+1)(KM+1)
    int sj, si;               // index strides: k 1, j sj=(KM+1), i si=(JM+1)(KM+1)
    int nblade;
    double w_rot;             // non-dimensional rotation rate
    double *Q;                // 5*box conservative
    double *W;                // 6*box primitive rho,u,v,w,p,T
    double *D[6];             // Dummy slabs IS IP JS JP KS KP
    long Dt[6];               // tangential box size of each slab
    int Db[6];                // stride of the first tangential index
    const double *fS[3];      // face vector, 3*box per direction
    const double *fs[3];      // face area |S|
    const double *fc[3];      // face centroid, 3*box
    const double *vol;        // cell volume
    const double *cc;         // cell centroid, 3*box
    const double *node;       // node coordinates, 3*box
    const unsigned char *wmask[6];  // wall mask per side (tangential box)
    const double *wrot[6];          // wall rotation rate per side
    double *lamc;             // 3*box convective spectral radii
    // laminar Navier-Stokes (null for Euler): coordinate-transformation
    // coefficients of the faces of each direction and of the cells, 9*box each
    // (component m at m*box: gradx/grady/gradz of kesi, then yita, then zita),
    // viscous spectral radii 3*box, vorticity magnitude box
    const double *fg[3];
    const double *cg;
    double *lamv;
    double *omega;
    // Baldwin-Lomax (null without a turbulence model): eddy viscosity over the
    // cell box (ghost cells included) and the minimum over the walls whose lines
    // visit a cell in this evaluation, as the bits of a non-negative double
    // (all ones = not visited; the merge resets it)
    double *mut;
    unsigned long long *blmin;
    // Spalart-Allmaras (null without it): nu~ over the cell box, its value
    // before the update, its LU-SGS increment and residual, and the wall
    // distance of the cells
    double *ev;
    double *ev_old;
    double *dev;
    double *res_ev;
    double *dwall;
    double *alpha2;           // nu~ LU-SGS diagonal Vol/dt + relax*sum(lamc) + 3*sum(lamv)
    int wf0, nwf;             // the wall faces of the block's instance in the wall-distance list
    const double *cfl;        // CFL number of the block's instance (device scalar)
    double *dt;               // local time step
    double *alpha;            // LU-SGS diagonal Vol/dt + relax*sum(lamc) + 2*sum(lamv)
    const double *swg;        // 15*box static sweep metrics per cell: for d = 0..2,
                              // averaged unit normal (3), averaged area, grid normal velocity
    double *R;                // 5*box residual
    double *dQ;               // 5*box LU-SGS increment
    double *fcv;              // 15*box convective face fluxes (direction, component) of the face-parallel
                              // form, face (d, i, j, k) at cell (i, j, k); null unless that form is built
    int *flag_low_dt;         // time-step damping counter (source/Limit_FlowVars.cpp:117)
    // Tile-skewed sweep layout.  The (j,k) plane is cut into tiles of
    // kKbaTJ x kKbaTK columns; tile (J,K) is slot J*ntk+K and owns nr rows.
    // Row r of a tile holds the 64 cells with i + tj + tk = r + 1 (tj, tk the
    // cell's place in the tile), so every cell a wavefront step touches lies
    // in one row and its three upstream neighbours in the row before.
    double *srow;             // per row: kSweepQ quantities x 64 lanes
    double *sdq;              // per row: dQ (5) and dnu~, kSweepD components x 64 lanes
    int ntk;                  // tiles along k
    int nr;                   // rows per tile = ni + kKbaTJ + kKbaTK - 2
    // Edge mailboxes of the row-layout sweeps (forward, backward): per tile
    // row, the dQ of the lanes a downstream tile reads across the j edge
    // (5 x kKbaTK) and across the k edge (5 x kKbaTJ).  Reset to a sentinel
    // before each iteration; a consumer waits until its words are not it.
    double *mbox[2];
    // coarse level of the two-grid cycle (null elsewhere): the restricted state
    // cv_2h and the restricted residual Res_2h, 5*box each
    double *cv2h;
    double *r2h;
    // levels below the first coarse one of the N-level V-cycle: the forcing
    // P = Res_2h - R(Q_2h) of the full approximation storage scheme, 5*box
    double *fasP;
};

__host__ __device__ inline long cidx(const BlockDev &B, int i, int j, int k)
{
    return (long)i * B.si + (long)j * B.sj + k;
}

// The range of work item t in a list of consecutive per-block ranges:
// the last n < count with start[n] <= t (binary search).
__device__ inline int find_range(const long *start, int count, long t)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Cell (i, j, k) of block n for work item t of a cell list (block n owns
// items cstart[n] .. cstart[n+1]-1, its first ni nj nk items are its cells,
// k fastest; items past them pad an instance to a whole number of 256-item
// groups).  False for padding and items past the list.
__device__ inline bool cell_item(const BlockDev *blocks, const long *cstart, int nblock, long t, int &n, int &i, int &j,
                                 int &k)
{
    if (t >= cstart[nblock]) return false;
    // The wave's lanes hold consecutive items: the block of the first active
    // lane's item is searched once for the wave (a uniform search, scalar
    // loads), and each lane steps on from it over the block ends below its
    // own item (none for most waves).
    const long t0 = ((long)__builtin_amdgcn_readfirstlane((int)(t >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)t);
    if (t >= t0) {
        n = find_range(cstart, nblock, t0);
        while (n + 1 < nblock && cstart[n + 1] <= t) ++n;
    } else {
        n = find_range(cstart, nblock, t);
    }
    const BlockDev &B = blocks[n];
    long r = t - cstart[n];
    if (r >= (long)B.ni * B.nj * B.nk) return false;
    k = (int)(r % B.nk) + 1;
    r /= B.nk;
    j = (int)(r % B.nj) + 1;
    i = (int)(r / B.nj) + 1;
    return true;
}

// f(B, n, i, j, k) for the cell of work item t, as cell_item finds it (nothing
// for padding and items past the list).  When every active lane of the wave
// lies in one block (most waves), the block index is wave-uniform and the
// block table entry is read through scalar loads, all issued at once; a wave
// across a block end takes the per-lane form.  f is compiled for both.
template <class F>
__device__ inline void for_cell_item(const BlockDev *__restrict__ blocks, const long *__restrict__ cstart, int nblock,
                                     long t, F &&f)
{
    const long t0 = ((long)__builtin_amdgcn_readfirstlane((int)(t >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)t);
    const int n0 = find_range(cstart, nblock, t0);
    const long s0 = cstart[n0], e0 = cstart[n0 + 1];
    if (__builtin_amdgcn_ballot_w64(t < s0 || t >= e0) == 0) {
        const BlockDev &B = blocks[n0];
        long r = t - s0;
        if (r >= (long)B.ni * B.nj * B.nk) return;   // padding of the instance
        const int k = (int)(r % B.nk) + 1;
        r /= B.nk;
        const int j = (int)(r % B.nj) + 1, i = (int)(r / B.nj) + 1;
        f(B, n0, i, j, k);
    } else {
        int n, i, j, k;
        if (cell_item(blocks, cstart, nblock, t, n, i, j, k)) f(blocks[n], n, i, j, k);
    }
}

// Dummy slab address of component c (0-4 cv, 5-10 pv) of layer l (1..2) on side s
// at tangential indices (a, b): I sides (j,k), J sides (i,k), K sides (i,j).
__device__ inline double *dummy_ptr(const BlockDev &B, int s, int l, int c, int a, int b)
{
    return B.D[s] + ((long)(l - 1) * 11 + c) * B.Dt[s] + (long)a * B.Db[s] + b;
}

struct Gas {
    double gamma, rgas;
    double pt_in, tt_in, p_out;
    double inlet_dir[3];
    double c_timestep, relax;
    int recon_var, recon_scheme, flux_scheme;
    int viscous;                  // Navier-Stokes terms on
    int turbulence;               // 0 laminar, 100 Baldwin-Lomax, 200 Spalart-Allmaras
    double re_ref, pr_lam, pr_tur;
    double ev_ratio;              // inlet nu~ (Spalart-Allmaras)
};

// CFluidProps::SutherLand (include/FluidProps.h:107-112), non-dimensional
__device__ inline double sutherland(const Gas &g, double temp)
{
    return F32C(1.0) / g.re_ref * pow(temp, F32C(3.0) / F32C(2.0)) * (double)(1.0f + 110.4f / 288.15f) /
           (temp + (double)(110.4f / 288.15f));
}

// ---------------------------------------------------------------- minmod
__device__ inline double mm1(double a, double b)
{
    if (a * b < 0.0) return 0.0;
    if (fabs(a) <= fabs(b)) return a;
    return b;
}

// -------------------------------------------------------- reconstruction
// One component of the primitive-variable reconstruction
// (source/Reconstruction.cpp:70-245).  Returns left and right face values.
__device__ inline void recon1_prim(int scheme, double FF, double F, double B, double BB, double &L, double &R)
{
    const double eps = F32C(1E-6);
    if (scheme == 500) {
        double dF = F - FF, dB = B - F, dR = BB - B;
        double SL = ((2.0 * dF) * dB + eps) / ((dF * dF + dB * dB) + eps);
        double SR = ((2.0 * dR) * dB + eps) / ((dR * dR + dB * dB) + eps);
        L = F + (SL / 4.0) * ((1.0 - SL / 3.0) * dF + (SL / 3.0 + 1.0) * dB);
        R = B - (SR / 4.0) * ((1.0 - SR / 3.0) * dR + (SR / 3.0 + 1.0) * dB);
    } else if (scheme == 100) {
        L = F + 0.5 * mm1(B - F, F - FF);
        R = B - 0.5 * mm1(B - F, BB - B);
    } else if (scheme == 200) {
        const double k2u = -1.0;
        L = F + F32C(0.25) * ((F32C(1.0) - k2u) * mm1(F32C(2.0) * (B - F), F - FF) +
                              (F32C(1.0) + k2u) * mm1(F32C(2.0) * (F - FF), B - F));
        R = B - F32C(0.25) * ((F32C(1.0) - k2u) * mm1(F32C(2.0) * (B - F), BB - B) +
                              (F32C(1.0) + k2u) * mm1(F32C(2.0) * (BB - B), B - F));
    } else if (scheme == 300) {
        const double k2c = 1.0;
        L = F + F32C(0.25) * ((F32C(1.0) - k2c) * mm1(F32C(2.0) * (B - F), F - FF) +
                              (F32C(1.0) + k2c) * mm1(B - F, 2.0 * (F - FF)));
        R = B - F32C(0.25) * ((F32C(1.0) - k2c) * mm1(F32C(2.0) * (B - F), BB - B) +
                              (F32C(1.0) + k2c) * mm1(B - F, 2.0 * (BB - B)));
    } else {  // 400 MUSCL2o
        const double eta = F32C(0.2);
        const double a0 = F32C(1.0) - eta, a1 = (eta - F32C(0.55)) / F32C(2.0), a2 = (eta + F32C(0.55)) / F32C(2.0);
        double r1 = ((F - FF) + eps) / ((B - F) + eps);
        double r2 = ((B - F) + eps) / ((BB - B) + eps);
        double f1 = (1.0 * (a1 / r2) + a0) + a2 * r1;
        double m = fmin(fmin(F32C(2.0), f1), F32C(2.0) * r1);
        double f = F32C(0.0) > m ? F32C(0.0) : m;
        L = F + (F32C(0.5) * f) * (B - F);
        r1 = ((BB - B) + eps) / ((B - F) + eps);
        r2 = ((B - F) + eps) / ((F - FF) + eps);
        f1 = (1.0 * (a1 / r2) + a0) + a2 * r1;
        m = fmin(fmin(F32C(2.0), f1), F32C(2.0) * r1);
        f = F32C(0.0) > m ? F32C(0.0) : m;
        R = B - (F32C(0.5) * f) * (B - F);
    }
}

// One component of the conservative-variable reconstruction
// (source/Reconstruction.cpp:7-60).
__device__ inline void recon1_cons(int scheme, double FF, double F, double B, double BB, double &L, double &R)
{
    const double eps = F32C(1E-6);
    if (scheme == 500) {
        double dF = F - FF, dB = B - F, dR = BB - B;
        double SL = ((dF * F32C(2.0)) * dB + eps) / ((dF * dF + dB * dB) + eps);
        double SR = ((dR * F32C(2.0)) * dB + eps) / ((dR * dR + dB * dB) + eps);
        L = F + (SL / F32C(4.0)) * ((F32C(1.0) - SL / F32C(3.0)) * dF + (SL / F32C(3.0) + F32C(1.0)) * dB);
        R = B - (SR / F32C(4.0)) * ((F32C(1.0) - SR / F32C(3.0)) * dR + (SR / F32C(3.0) + F32C(1.0)) * dB);
    } else if (scheme == 100) {
        L = F + mm1(B - F, F - FF) * F32C(0.5);
        R = B - mm1(B - F, BB - B) * F32C(0.5);
    } else if (scheme == 200) {
        const double k2u = F32C(-1.0);
        L = F + (mm1((B - F) * F32C(2.0), F - FF) * (F32C(1.0) - k2u) +
                 mm1((F - FF) * F32C(2.0), B - F) * (F32C(1.0) + k2u)) * 0.25;
        R = B - (mm1((B - F) * F32C(2.0), BB - B) * (F32C(1.0) - k2u) +
                 mm1((BB - B) * F32C(2.0), B - F) * (F32C(1.0) + k2u)) * 0.25;
    } else if (scheme == 300) {
        const double k2c = F32C(1.0);
        L = F + (mm1((B - F) * F32C(2.0), F - FF) * (F32C(1.0) - k2c) +
                 mm1(B - F, (F - FF) * F32C(2.0)) * (F32C(1.0) + k2c)) * 0.25;
        R = B - (mm1((B - F) * F32C(2.0), BB - B) * (F32C(1.0) - k2c) +
                 mm1(B - F, (BB - B) * F32C(2.0)) * (F32C(1.0) + k2c)) * 0.25;
    } else {
        const double eta = F32C(0.05);
        const double a0 = F32C(1.0) - eta, a1 = (eta - F32C(0.55)) / F32C(2.0), a2 = (eta + F32C(0.55)) / F32C(2.0);
        double r1 = ((F - FF) + eps) / ((B - F) + eps);
        double r2 = ((B - F) + eps) / ((BB - B) + eps);
        double f1 = (a1 / r2 + a0) + r1 * a2;
        double m = fmin(fmin(F32C(2.0), f1), F32C(2.0) * r1);
        double f = F32C(0.0) > m ? F32C(0.0) : m;
        L = F + (f * F32C(0.5)) * (B - F);
        r1 = ((BB - B) + eps) / ((B - F) + eps);
        r2 = ((B - F) + eps) / ((F - FF) + eps);
        f1 = (a1 / r2 + a0) + r1 * a2;
        m = fmin(fmin(F32C(2.0), f1), F32C(2.0) * r1);
        f = F32C(0.0) > m ? F32C(0.0) : m;
        R = B - (f * F32C(0.5)) * (B - F);
    }
}

// Characteristic-variable reconstruction (source/Reconstruction.cpp:363-468):
// the left (S) and right (S1) eigenvector matrices of the face state, the
// mean of F's and B's primitive variables; the four conservative states go to
// characteristic variables, are limited component by component as conservative
// ones are, and come back through S1.  The -1/3, 2/3, 1/6 entries and SQRT(3.0f)
// are float values in the reference.  qst holds FF, F, B, BB (conservative),
// wF and wB the primitive F and B.  Positivity is checked by the caller.
__device__ inline void recon_char(int scheme, double Gamma, double Rgas, const double wF[6], const double wB[6],
                                  const double qst[4][5], double cvL[5], double cvR[5])
{
    const double uvel = F32C(0.5) * (wF[1] + wB[1]), vvel = F32C(0.5) * (wF[2] + wB[2]);
    const double wvel = F32C(0.5) * (wF[3] + wB[3]), press = F32C(0.5) * (wF[4] + wB[4]);
    const double dens = F32C(0.5) * (wF[0] + wB[0]);
    const double V2 = F32C(0.5) * (uvel * uvel + vvel * vvel + wvel * wvel);
    const double temp = press / dens / Rgas;
    const double C = sqrt(Gamma * Rgas * temp);
    const double C2 = C * C;
    const double sqrt3f = 0x1.bb67aep+0;   // sqrtf(3.0f)
    const double Med0 = C2 / (Gamma - F32C(1.0));
    const double Med1 = (Gamma - F32C(1.0)) / C * sqrt3f;
    const double Med3 = F32C(1.0) / Med0;
    const double Med4 = F32C(0.5) / C / sqrt3f;
    const double V_bar = (uvel + vvel + wvel);
    const double m3 = (double)(-1.0f / 3.0f), t23 = (double)(2.0f / 3.0f), s6 = (double)(1.0f / 6.0f);
    const double S[5][5] = {
        {V2 - Med0, -uvel, -vvel, -wvel, F32C(1.0)},
        {uvel - vvel, F32C(-1.0), F32C(1.0), 0.0, 0.0},
        {uvel - wvel, F32C(-1.0), 0.0, F32C(1.0), 0.0},
        {-V_bar - V2 * Med1, F32C(1.0) + Med1 * uvel, F32C(1.0) + Med1 * vvel, F32C(1.0) + Med1 * wvel, -Med1},
        {-V_bar + V2 * Med1, F32C(1.0) - Med1 * uvel, F32C(1.0) - Med1 * vvel, F32C(1.0) - Med1 * wvel, Med1}};
    const double S1[5][5] = {
        {-Med3, 0.0, 0.0, -Med4, Med4},
        {-Med3 * uvel, m3, m3, s6 - Med4 * uvel, s6 + Med4 * uvel},
        {-Med3 * vvel, t23, m3, s6 - Med4 * vvel, s6 + Med4 * vvel},
        {-Med3 * wvel, m3, t23, s6 - Med4 * wvel, s6 + Med4 * wvel},
        {-Med3 * V2, vvel - V_bar / F32C(3.0), wvel - V_bar / F32C(3.0), V_bar / F32C(6.0) - Med4 * (V2 + Med0),
         V_bar / F32C(6.0) + Med4 * (V2 + Med0)}};
    double ch[4][5];
    for (int o = 0; o < 4; ++o)
        for (int r = 0; r < 5; ++r)
            ch[o][r] = S[r][0] * qst[o][0] + S[r][1] * qst[o][1] + S[r][2] * qst[o][2] + S[r][3] * qst[o][3] +
                       S[r][4] * qst[o][4];
    double chL[5], chR[5];
    for (int r = 0; r < 5; ++r) recon1_cons(scheme, ch[0][r], ch[1][r], ch[2][r], ch[3][r], chL[r], chR[r]);
    for (int r = 0; r < 5; ++r) {
        cvL[r] = S1[r][0] * chL[0] + S1[r][1] * chL[1] + S1[r][2] * chL[2] + S1[r][3] * chL[3] + S1[r][4] * chL[4];
        cvR[r] = S1[r][0] * chR[0] + S1[r][1] * chR[1] + S1[r][2] * chR[2] + S1[r][3] * chR[3] + S1[r][4] * chR[4];
    }
}

// MIN3 of source/defs.cpp:401-408 keeps the first argument on ties and NaN;
// fmin differs only for NaN operands, which the positivity fallback catches.

// ---------------------------------------------------------- face fluxes
// Roe flux in the face frame (source/Flux_Split_Scheme.cpp:568-635).
__device__ inline void roe_flux(double Gamma, const double QL[5], const double QR[5], double que, double Fc[5])
{
    double densL = QL[0], uvelL = QL[1] / QL[0], vvelL = QL[2] / QL[0], wvelL = QL[3] / QL[0], EL = QL[4] / QL[0];
    double pL = (Gamma - F32C(1.0)) * densL * (EL - F32C(0.5) * (uvelL * uvelL + vvelL * vvelL + wvelL * wvelL));
    double HL = EL + pL / densL;
    double densR = QR[0], uvelR = QR[1] / QR[0], vvelR = QR[2] / QR[0], wvelR = QR[3] / QR[0], ER = QR[4] / QR[0];
    double pR = (Gamma - F32C(1.0)) * densR * (ER - F32C(0.5) * (uvelR * uvelR + vvelR * vvelR + wvelR * wvelR));
    double HR = ER + pR / densR;
    double FL0 = QL[0] * (uvelL - que), FL1 = QL[1] * (uvelL - que) + pL, FL2 = QL[2] * (uvelL - que),
           FL3 = QL[3] * (uvelL - que), FL4 = (QL[4] + pL) * (uvelL - que) + que * pL;
    double FR0 = QR[0] * (uvelR - que), FR1 = QR[1] * (uvelR - que) + pR, FR2 = QR[2] * (uvelR - que),
           FR3 = QR[3] * (uvelR - que), FR4 = (QR[4] + pR) * (uvelR - que) + que * pR;
    double tmpp = sqrt(densR / densL);
    double uav = (uvelL + tmpp * uvelR) / (F32C(1.0) + tmpp);
    double vav = (vvelL + tmpp * vvelR) / (F32C(1.0) + tmpp);
    double wav = (wvelL + tmpp * wvelR) / (F32C(1.0) + tmpp);
    double Hav = (HL + tmpp * HR) / (F32C(1.0) + tmpp);
    double cav = sqrt((Gamma - F32C(1.0)) * (Hav - F32C(0.5) * (uav * uav + vav * vav + wav * wav)));
    const double delta = F32C(0.5);
    double l1 = fabs((uav - que) - cav), l2 = fabs(uav - que), l5 = fabs((uav - que) + cav);
    l1 = l1 > delta ? l1 : F32C(0.5) * (l1 * l1 + delta * delta) / delta;
    l2 = l2 > delta ? l2 : F32C(0.5) * (l2 * l2 + delta * delta) / delta;
    l5 = l5 > delta ? l5 : F32C(0.5) * (l5 * l5 + delta * delta) / delta;
    double a2 = (QR[2] - QL[2]) / (vav + F32C(1E-8)) - (densR - densL);
    double a3 = (QR[3] - QL[3]) / (wav + F32C(1E-8)) - (densR - densL);
    double a4 = (Gamma - F32C(1.0)) *
                ((densR - densL) * (Hav - uav * uav - vav * vav - wav * wav) + uav * (QR[1] - QL[1]) +
                 vav * (QR[2] - QL[2]) + wav * (QR[3] - QL[3]) - (QR[4] - QL[4])) / (cav * cav);
    double a1 = ((uav + cav) * (densR - densL) - (QR[1] - QL[1]) - cav * a4) / (F32C(2.0) * cav);
    double a5 = densR - densL - (a1 + a4);
    Fc[0] = F32C(0.5) * (FL0 + FR0) - F32C(0.5) * (l1 * a1 + l2 * a4 + l5 * a5);
    Fc[1] = F32C(0.5) * (FL1 + FR1) - F32C(0.5) * (l1 * a1 * (uav - cav) + l2 * a4 * uav + l5 * a5 * (uav + cav));
    Fc[2] = F32C(0.5) * (FL2 + FR2) - F32C(0.5) * (l1 * a1 * vav + l2 * a2 * vav + l2 * a4 * vav + l5 * a5 * vav);
    Fc[3] = F32C(0.5) * (FL3 + FR3) - F32C(0.5) * (l1 * a1 * wav + l2 * a3 * wav + l2 * a4 * wav + l5 * a5 * wav);
    Fc[4] = F32C(0.5) * (FL4 + FR4) -
            F32C(0.5) * (l1 * a1 * (Hav - uav * cav) + l2 * a2 * vav * vav + l2 * a3 * wav * wav +
                         l2 * a4 * F32C(0.5) * (uav * uav + vav * vav + wav * wav) + l5 * a5 * (Hav + uav * cav));
}

// One side of the Steger-Warming splitting (source/Flux_Split_Scheme.cpp:82-149);
// sgn = +1 for F+(Q_L), -1 for F-(Q_R).
__device__ inline void sw_half(double Gamma, const double Q[5], double que, double sgn, double f[5])
{
    const double eps = F32C(1E-6);
    double dens = Q[0], u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0], e = Q[4] / Q[0];
    double t = Gamma * (Gamma - F32C(1.0)) * (e - F32C(0.5) * (u * u + v * v + w * w));
    t = sqrt(t * t);
    double c = sqrt(t);
    double l1 = F32C(0.5) * ((u - que) + sgn * sqrt((u - que) * (u - que) + eps * eps));
    double l4 = F32C(0.5) * (((u - que) - c) + sgn * sqrt(((u - que) - c) * ((u - que) - c) + eps * eps));
    double l5 = F32C(0.5) * (((u - que) + c) + sgn * sqrt(((u - que) + c) * ((u - que) + c) + eps * eps));
    double W2 = (F32C(3.0) - Gamma) * (l4 + l5) * c * c / F32C(2.0) / (Gamma - F32C(1.0));
    double A = (u - c) * (u - c) + v * v + w * w;
    double Bq = (u + c) * (u + c) + v * v + w * w;
    double pre = dens / F32C(2.0) / Gamma;
    f[0] = pre * (F32C(2.0) * (Gamma - F32C(1.0)) * l1 + l4 + l5);
    f[1] = pre * (F32C(2.0) * (Gamma - F32C(1.0)) * l1 * u + l4 * (u - c) + l5 * (u + c));
    f[2] = pre * (F32C(2.0) * (Gamma - F32C(1.0)) * l1 * v + l4 * v + l5 * v);
    f[3] = pre * (F32C(2.0) * (Gamma - F32C(1.0)) * l1 * w + l4 * w + l5 * w);
    f[4] = pre * ((Gamma - F32C(1.0)) * l1 * (u * u + v * v + w * w) + F32C(0.5) * l4 * A + F32C(0.5) * l5 * Bq + W2);
}

// ------------------------------------------------- face flux from a stencil
// The interior face of CSpaceDiscr::Upwind_Convective
// (source/Upwind_Convective.cpp:5-360) from its four-cell stencil FF, F, B, BB
// (st: primitive rho u v w p for recon_var 200, conservative otherwise) and, for
// characteristic reconstruction, the primitive F and B (wF, wB); then the
// face frame of Coord_Rotate (source/Flux_Split_Scheme.cpp:728-776) from the
// face vector S, its area s, the node diagonal r and the grid velocity of the
// face centroid (fcx, fcy) times w_rot, the Roe or Steger-Warming flux, and
// the flux rotated back and times the area.
__device__ inline void face_core(const Gas &g, const double st[4][5], const double wF[6], const double wB[6],
                                 double Sx, double Sy, double Sz, double s_, double fcx, double fcy, double w_rot,
                                 double rx, double ry, double rz, double out[5])
{
    const double Gamma = g.gamma;
    const double lim_zero = F32C(1E-20);
    double L[5], R[5], cvL[5], cvR[5];
    if (g.recon_var != 100 && g.recon_var != 200) {
        recon_char(g.recon_scheme, Gamma, g.rgas, wF, wB, st, cvL, cvR);
        double pL = (Gamma - F32C(1.0)) * (cvL[4] - F32C(0.5) * (cvL[1] * cvL[1] + cvL[2] * cvL[2] + cvL[3] * cvL[3]) / cvL[0]);
        double pR = (Gamma - F32C(1.0)) * (cvR[4] - F32C(0.5) * (cvR[1] * cvR[1] + cvR[2] * cvR[2] + cvR[3] * cvR[3]) / cvR[0]);
        if (cvL[0] <= lim_zero || pL <= lim_zero) for (int c = 0; c < 5; ++c) cvL[c] = st[1][c];
        if (cvR[0] <= lim_zero || pR <= lim_zero) for (int c = 0; c < 5; ++c) cvR[c] = st[2][c];
    } else if (g.recon_var == 200) {
        for (int c = 0; c < 5; ++c) recon1_prim(g.recon_scheme, st[0][c], st[1][c], st[2][c], st[3][c], L[c], R[c]);
        if (L[0] <= lim_zero || L[4] <= lim_zero) for (int c = 0; c < 5; ++c) L[c] = st[1][c];
        if (R[0] <= lim_zero || R[4] <= lim_zero) for (int c = 0; c < 5; ++c) R[c] = st[2][c];
        cvL[0] = L[0]; cvL[1] = L[0] * L[1]; cvL[2] = L[0] * L[2]; cvL[3] = L[0] * L[3];
        cvL[4] = L[4] / (Gamma - F32C(1.0)) + L[0] * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) / F32C(2.0);
        cvR[0] = R[0]; cvR[1] = R[0] * R[1]; cvR[2] = R[0] * R[2]; cvR[3] = R[0] * R[3];
        cvR[4] = R[4] / (Gamma - F32C(1.0)) + R[0] * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]) / F32C(2.0);
    } else {
        for (int c = 0; c < 5; ++c) recon1_cons(g.recon_scheme, st[0][c], st[1][c], st[2][c], st[3][c], cvL[c], cvR[c]);
        double pL = (Gamma - F32C(1.0)) * (cvL[4] - F32C(0.5) * (cvL[1] * cvL[1] + cvL[2] * cvL[2] + cvL[3] * cvL[3]) / cvL[0]);
        double pR = (Gamma - F32C(1.0)) * (cvR[4] - F32C(0.5) * (cvR[1] * cvR[1] + cvR[2] * cvR[2] + cvR[3] * cvR[3]) / cvR[0]);
        if (cvL[0] <= lim_zero || pL <= lim_zero) for (int c = 0; c < 5; ++c) cvL[c] = st[1][c];
        if (cvR[0] <= lim_zero || pR <= lim_zero) for (int c = 0; c < 5; ++c) cvR[c] = st[2][c];
    }
    double nx = Sx / s_, ny = Sy / s_, nz = Sz / s_;
    double rl = sqrt(rx * rx + ry * ry + rz * rz);
    double t1x = rx / rl, t1y = ry / rl, t1z = rz / rl;
    double qx = ny * t1z - nz * t1y, qy = nz * t1x - nx * t1z, qz = nx * t1y - ny * t1x;
    double ql = sqrt(qx * qx + qy * qy + qz * qz);
    double t2x = qx / ql, t2y = qy / ql, t2z = qz / ql;
    double fue = F32C(-1.0) * fcy * w_rot, fve = fcx * w_rot;
    double que = fue * nx + fve * ny + F32C(0.0) * nz;
    double QL[5], QR[5];
    QL[0] = cvL[0]; QL[4] = cvL[4];
    QL[1] = cvL[1] * nx + cvL[2] * ny + cvL[3] * nz;
    QL[2] = cvL[1] * t1x + cvL[2] * t1y + cvL[3] * t1z;
    QL[3] = cvL[1] * t2x + cvL[2] * t2y + cvL[3] * t2z;
    QR[0] = cvR[0]; QR[4] = cvR[4];
    QR[1] = cvR[1] * nx + cvR[2] * ny + cvR[3] * nz;
    QR[2] = cvR[1] * t1x + cvR[2] * t1y + cvR[3] * t1z;
    QR[3] = cvR[1] * t2x + cvR[2] * t2y + cvR[3] * t2z;
    double Fc[5];
    if (g.flux_scheme == 200) {
        double fp_[5], fm_[5];
        sw_half(Gamma, QL, que, 1.0, fp_);
        sw_half(Gamma, QR, que, -1.0, fm_);
        for (int c = 0; c < 5; ++c) Fc[c] = fp_[c] + fm_[c];
    } else {
        roe_flux(Gamma, QL, QR, que, Fc);
    }
    out[0] = Fc[0] * s_;
    out[1] = (Fc[1] * nx + Fc[2] * t1x + Fc[3] * t2x) * s_;
    out[2] = (Fc[1] * ny + Fc[2] * t1y + Fc[3] * t2y) * s_;
    out[3] = (Fc[1] * nz + Fc[2] * t1z + Fc[3] * t2z) * s_;
    out[4] = Fc[4] * s_;
}

// The WallFlux override of a wall face (source/WallFlux.cpp:70-175): the flux
// of the state averaged from the primitive variables of the two cells either
// side (a = the cell the face's index names, b = the one below it), with the
// wall's own rotation rate rot, on face vector S and face centroid (fcx, fcy).
__device__ inline void wall_face_core(double Gamma, const double wa[5], const double wb[5], double Sx, double Sy,
                                      double Sz, double fcx, double fcy, double rot, double out[5])
{
    double fd = F32C(0.5) * (wa[0] + wb[0]);
    double fu = F32C(0.5) * (wa[1] + wb[1]);
    double fv = F32C(0.5) * (wa[2] + wb[2]);
    double fw = F32C(0.5) * (wa[3] + wb[3]);
    double fp = F32C(0.5) * (wa[4] + wb[4]);
    double fue = F32C(-1.0) * fcy * rot, fve = fcx * rot;
    double c1 = fd * fu, c2 = fd * fv, c3 = fd * fw;
    double Vn_a = fu * Sx + fv * Sy + fw * Sz;
    double Vn_e = fue * Sx + fve * Sy + F32C(0.0) * Sz;
    double c4 = fp / (Gamma - F32C(1.0)) + fd * (fu * fu + fv * fv + fw * fw) / F32C(2.0);
    out[0] = fd * (Vn_a - Vn_e);
    out[1] = c1 * (Vn_a - Vn_e) + fp * Sx;
    out[2] = c2 * (Vn_a - Vn_e) + fp * Sy;
    out[3] = c3 * (Vn_a - Vn_e) + fp * Sz;
    out[4] = (c4 + fp) * (Vn_a - Vn_e) + fp * Vn_e;
}

}  // namespace fv3d
