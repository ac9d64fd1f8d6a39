// Kernel declarations and the device-side descriptors they share with the
// context code.
#pragma once

#include "fv3d_device.h"

namespace fv3d {

// tile of the residual kernels (k_convective, k_viscous, k_sa_faces; 256
// cells per workgroup): the shapes they are built for, the default first; the
// autotuner picks one per kernel (results do not depend on it: every face is
// evaluated once and every cell sums its faces in a fixed order)
constexpr int kTI = 4, kTJ = 4, kTK = 16;
constexpr int kNumTiles = 5;
constexpr int kTileShape[kNumTiles][3] = {{4, 4, 16}, {2, 8, 16}, {4, 8, 8}, {8, 4, 8}, {8, 8, 4}};
#define FV3D_FOR_TILES(X) X(0, 4, 4, 16) X(1, 2, 8, 16) X(2, 4, 8, 8) X(3, 8, 4, 8) X(4, 8, 8, 4)

// streaming convective residual (k_conv_stream): column tiles TJ x TK marching
// along i in chunks of ci cells.  The convective kernel's configuration number
// q counts the per-tile shapes first (q < kNumTiles, k_convective), then these
// (q - kNumTiles, k_conv_stream); the autotuner times all of them, and results
// do not depend on the choice.
constexpr int kNumStream = 5;
constexpr int kStreamShape[kNumStream][3] = {{16, 16, 4}, {16, 16, 2}, {8, 32, 4}, {8, 32, 2}, {16, 16, 8}};
#define FV3D_FOR_STREAM_TILES(X) X(0, 16, 16) X(1, 8, 32)
// face-parallel convective residual (k_conv_faces, then k_conv_sum): one thread
// per face, configuration kConvFaces after the streaming ones.  Coarse
// multigrid levels of at most kFaceFormCells cells take it by default: their
// tiles are too few to fill the chip, and a tile workgroup evaluates five or
// six faces per thread in turn.
constexpr int kConvFaces = kNumTiles + kNumStream;
constexpr int kNumConv = kConvFaces + 1;   // convective configurations
#ifndef FV3D_FACE_FORM_CELLS
#define FV3D_FACE_FORM_CELLS 1000000
#endif
constexpr long kFaceFormCells = FV3D_FACE_FORM_CELLS;

struct PatchDev {
    int kind;          // 0 inlet, 1 outlet, 2 wall
    int block;         // 0-based block index
    int dir, hi;       // side: direction 0/1/2 and high (1) or low (0) face
    int a0, b0;        // first tangential cell indices
    int na, nb;        // tangential extents
    int centrifugal;   // outlet variant
    int axial;         // radial-equilibrium outlet: applied by k_bc_axial, skipped by k_bc
    int inst;          // instance of the patch's block (its scalars hold the outlet pressure)
    double rot;        // wall rotation rate
};

// Mixing plane: block b1's K-high face against block b2's K-low face, radial
// lines i0 .. i0+ni-1 of both, pitch-wise cells j1 .. j1+nj1-1 of b1 and
// j2 .. j2+nj2-1 of b2; line0 numbers the first line among all planes.
constexpr int kMixMaxJ = 512;
struct MixDev {
    int b1, b2, i0, ni, j1, nj1, j2, nj2, line0;
};

// Per-run scalars kept on the device so iterations need no host round trip.
struct ScalarsDev {
    double cfl, cfl_res;
    double res[5];
    double res_ev;      // Spalart-Allmaras nu~ residual norm
    double total_vol;
    double dens_ref, vel_ref, temp_ref;
    double nblade_in, nblade_out;
    double mass_in, mass_out, pr, eff;
    double p_out;
    double iter;
};

struct PeriodicDev;
// sc: the scalars of each instance (outlet pressure); with nper > 0 thread 0
// first does the work of k_periodic_flip
__global__ void k_bc(const BlockDev *blocks, const PatchDev *patches, int npatch, const long *pstart, Gas g,
                     const ScalarsDev *sc, PeriodicDev *per, int nper, int *pitch_sign);
// radial-equilibrium outlet, one workgroup per patch (list: patch indices);
// the radial lines of a face are at most kAxialMaxI cells long
constexpr int kAxialMaxI = 2048;
__global__ void k_bc_axial(const BlockDev *blocks, const PatchDev *patches, const int *list, Gas g,
                           const ScalarsDev *sc);
// global time step (timestep type "G"): per-block minimum of the local steps
// (dtmin as bits of positive doubles), then dt and the LU-SGS diagonal
__global__ void k_dt_block_min(const BlockDev *blocks, const long *cstart, int nblock, unsigned long long *dtmin);
__global__ void k_dt_global(const BlockDev *blocks, const long *cstart, int nblock, Gas g,
                            const unsigned long long *dtmin);
// one interior/periodic transfer: the state of cell src of block src_block goes
// to layer `layer` of Dummy slab `side` of block dst_block at tangential slot
// `slot` (and to ghost cell `ghost` of that block, -1 for none); rot 0 copies,
// 1 and 2 rotate by the pitch angle of periodic patch `per` with sign +/-
struct ExchDev {
    int src_block, dst_block, side, layer, rot, per;
    long src, slot, ghost;
};
// periodic patch: the block whose pitch angle it uses, whether applying it
// flips that angle's sign (source/BndCondInteriorPeriodic.cpp:54-59), the sign
// in force for this application, and cos/sin of +delta and -delta
struct PeriodicDev {
    int block, flip, sign;
    double cs[4];
};
__global__ void k_periodic_flip(PeriodicDev *per, int nper, int *pitch_sign);
__global__ void k_exchange(const BlockDev *blocks, const ExchDev *el, long n, const PeriodicDev *per);
// the CFL number of each block's instance: BlockDev::cfl
__global__ void k_spectral_dt(const BlockDev *blocks, const long *cstart, int nblock, Gas g);
__global__ void k_dt_flag_decrement(const BlockDev *blocks, int nblock);
template <int TI, int TJ, int TK>
__global__ void k_convective(const BlockDev *blocks, const int4 *tiles, Gas g, int store_faces, double *fI, double *fJ,
                             double *fK);
// fstart: running face counts per (block, direction), 3 * nblock entries + 1
__global__ void k_conv_faces(const BlockDev *blocks, const long *fstart, int nlist, Gas g);
__global__ void k_conv_sum(const BlockDev *blocks, const long *cstart, int nblock);
template <bool FORWARD>
__global__ void k_lusgs_plane(const BlockDev *blocks, const long *pstart, int nblock, int step, double relax, Gas g);
// one tile of the persistent wavefront sweep; dep_*: ticket of the upstream
// tile along j / k, or -1 at the block edge
struct KbaTile {
    int block, j0, k0, dep_j, dep_k;
};
constexpr int kKbaTJ = 4, kKbaTK = 16;
constexpr int kDonePad = 32;   // progress counters one per 128-byte line
// quantities of a sweep-layout row (BlockDev::srow): Q (5), p, lamc (3),
// swg (15), R (5), alpha, lamv (3, Navier-Stokes only), and with
// Spalart-Allmaras nu~, its residual and its LU-SGS diagonal alpha2
constexpr int kSQ_Q = 0, kSQ_P = 5, kSQ_LAM = 6, kSQ_SWG = 9, kSQ_R = 24, kSQ_A = 29, kSQ_LV = 30, kSQ_EV = 33,
              kSQ_REV = 34, kSQ_A2 = 35, kSweepQ = 36;
// components of a dQ row (BlockDev::sdq): dQ (5) and dnu~ (Spalart-Allmaras)
constexpr int kSweepD = 6;
// what k_sweep_fill writes: the state quantities and/or the static metrics
constexpr int kGatherState = 1, kGatherMetrics = 2;
// T: double (plain FP64) or float (the FP32 operator of defect correction)
// SA: the nu~ equation in the same sweep (Navier-Stokes only)
template <bool FWD, int TJ, int TK, bool VISC, bool SA, class T>
__global__ void k_lusgs_sl(const BlockDev *blocks, const KbaTile *tiles, int ntiles, int *ticket, int *err, double relax,
                           Gas g);
// mailbox words per tile row (kSweepD components of the j-edge lanes, then of
// the k-edge lanes), and the bit pattern of an empty word (a NaN payload no
// arithmetic produces)
constexpr int kMboxJ = kSweepD * kKbaTK, kMboxRow = kSweepD * (kKbaTK + kKbaTJ);
constexpr unsigned long long kMboxEmpty = 0x7ff4dead5eed0001ull;
constexpr unsigned int kMboxEmpty32 = 0x7fa5eed1u;   // the same for FP32 words
__global__ void k_fill_u64(unsigned long long *p, long n, unsigned long long v);
__global__ void k_fill_i32(int *p, int n, int v);
__global__ void k_prim_all(const BlockDev *blocks, const long *cstart, int nblock, Gas g, int *bad_flag);
// row storage in T (double, or float for the FP32 operator); unpack widens dQ to double
template <class T>
__global__ void k_sweep_fill(const BlockDev *blocks, int block, long nrows, int what);
// one launch over all blocks: (tile, chunk) items gstart[n] .. gstart[n+1]-1
// serve block n, each split over qsplit workgroups (a grid of
// gstart[nblock] * qsplit) that take equal shares of the nq quantities (15;
// 18 with lamv; 21 with the nu~ rows, when nu~ is carried in the sweep)
template <class T>
__global__ void k_sweep_pack(const BlockDev *blocks, const long *gstart, int nblock, int nq, int qsplit);
// nc: components per cell (5, or 6 with the nu~ increment), split over csplit
// workgroups per item
template <class T>
__global__ void k_sweep_unpack(const BlockDev *blocks, const long *gstart, int nblock, int nc, int csplit);
constexpr int kPackChunk = 16;   // i-planes per k_sweep_pack / k_sweep_unpack workgroup
constexpr int kSecParts = 32;    // workgroups per section in k_sections
__global__ void k_fill_int(int *p, long n, int v);
__global__ void k_update(const BlockDev *blocks, const long *cstart, int nblock, Gas g, double *partial, int *bad_flag,
                         int *sa_hi, int sa);
template <int TJ, int TK, int MODE>
__global__ void k_conv_stream(const BlockDev *blocks, const int4 *tiles, int ntiles, int ci, Gas g);
__global__ void k_limiter_serial(const BlockDev *blocks, int nblock, Gas g, int *bad_flag, const long *cstart,
                                 double *partial, int npart);
__global__ void k_sections(const BlockDev *blocks, const int2 *secs, Gas g, double *sums);
// norms and section integrals of one level (lv) with the CFL ramp and the
// iteration count kept in ramp (the fine level's scalars, shared by the
// levels); one workgroup per instance w: its norm partials are
// parts[w].x .. parts[w].x + parts[w].y - 1, its scalars lv[w] and ramp[w],
// its history row hist_row + 7 w
__global__ void k_finalize(const double *partial, int npart, const int2 *parts, const double *sums, ScalarsDev *lv,
                           ScalarsDev *ramp, Gas g, double *hist_row);
// two-grid cycle (CSolver::Prolong / Interpolate, CMesh::CornerTreatment,
// ComputeForceFunction / AddForceFunction)
__global__ void k_mg_restrict(const BlockDev *fine, const BlockDev *coarse, const long *cstart, int nblock, Gas g,
                              int *bad_flag);
__global__ void k_mg_corner(const BlockDev *coarse);
// coarse tile of the prolongation (tiles[w] = block, first coarse I, J, K)
constexpr int kMgTI = 2, kMgTJ = 4, kMgTK = 8;
template <bool STATE>
__global__ void k_mg_prolong(const BlockDev *fine, const BlockDev *coarse, const int4 *tiles, Gas g, int *bad_flag);
template <bool FORCE>
__global__ void k_fas(const BlockDev *blocks, const long *cstart, int nblock);

// laminar Navier-Stokes terms (fv3d_viscous.hip)
// coordinate-transformation coefficients of every face and cell of one block (set-up)
__global__ void k_metrics(const BlockDev *blocks, int block, double *fg0, double *fg1, double *fg2, double *cg);
// viscous face fluxes of a tile of cells, summed into R (R = Fc - Fv); with
// store set also the face fluxes, face gradients and cell sums
struct ViscDebug {
    double *fv[3];     // 5*box per direction
    double *grad[3];   // 12*box per direction: d/dx, d/dy, d/dz of u v w T
    double *fvc;       // 5*box
};
template <int TI, int TJ, int TK>
__global__ void k_viscous(const BlockDev *blocks, const int4 *tiles, Gas g, int store, ViscDebug dbg);
// vorticity magnitude per cell
__global__ void k_curl(const BlockDev *blocks, const long *cstart, int nblock);
// Baldwin-Lomax along the grid lines normal to a set of wall patches:
// kBlLines lines x G groups of kBlCpt cells per workgroup (so lines of up to
// G*kBlCpt cells); walls[w] = (patch index, line count rounded up to a
// multiple of kBlLines).  Each cell's value is folded into blmin with an
// atomic minimum; k_bl_merge then moves the minima into mut.
constexpr int kBlLines = 16, kBlCpt = 8;
constexpr unsigned long long kBlUnvisited = ~0ull;
template <int G>
__global__ void k_bl_lines(const BlockDev *blocks, const PatchDev *patches, const int2 *walls, int nwalls, Gas g);
__global__ void k_bl_merge(const BlockDev *blocks, const long *cstart, int nblock);

// Spalart-Allmaras (CSpaceDiscr::SA_Model, CMesh::WallDistanceSA,
// CFluidProps::Limit_SA and the nu~ part of CMesh::LU_SGS_Marching)
__global__ void k_wall_distance(const BlockDev *blocks, const long *cstart, int nblock, const double *wf);
template <int TI, int TJ, int TK>
__global__ void k_sa_faces(const BlockDev *blocks, const int4 *tiles, Gas g);
// save_old: keep nu~ as it stands (before the clipping) in ev_old, as the
// LU-SGS step does before its residual evaluation
__global__ void k_sa_mut(const BlockDev *blocks, const long *cstart, int nblock, Gas g, int save_old);
__global__ void k_sa_ghost(const BlockDev *blocks, const long *gstart, int nblock, Gas g);
template <bool FORWARD>
__global__ void k_sa_plane(const BlockDev *blocks, const long *pstart, int nblock, int step, double relax, Gas g);
// Limit_SA: the blocks k_update flagged, serially; then the pointwise floor of
// the others (clamp set) with the Res_ev partials in component 5
__global__ void k_sa_limit_serial(const BlockDev *blocks, int nblock, int *hi_flag);
__global__ void k_sa_res(const BlockDev *blocks, const long *cstart, int nblock, double *partial, int clamp);

__global__ void k_mixing_plane(const BlockDev *blocks, const MixDev *mix, int nmix, Gas g, int sa);

}  // namespace fv3d
