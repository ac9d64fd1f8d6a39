// Context, memory management and C ABI of libfv3d.
//
// The context owns every device allocation.  Metrics are computed once on the
// host in FP64 with the reference's expression order
// (source/GeometryVariables.cpp:10-105) and uploaded; all per-iteration work
// runs on the context's stream without host round trips.
#include "fv3d.h"
#include "fv3d_kernels.h"
#include "fv3d_mlp.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

using namespace fv3d;

namespace {

struct HostPatch {
    int kind, block, face, rotate;
    int rng[6];
};

// two-sided patch: 0 interior, 1 periodic (side 0 = block 1 of the table row)
struct HostConn {
    int kind, orient;
    int block[2], face[2], rng[2][6];
};

struct HostBlock {
    int inst = 0;                // instance (independent passage) of the block
    int IM, JM, KM, nblade;
    double rpm, w_rot, cellvol;
    std::vector<double> nodes;   // (IM, JM, KM, 3) as given
    BlockDev dev;                // device pointers + dims
    std::vector<void *> allocs;
    long sweep_rows = 0;         // rows of the tile-skewed sweep layout
    long pack_groups = 0;        // (tile, chunk) items of k_sweep_pack / k_sweep_unpack
};

// An independent passage of a batched context: its blocks are
// first_block .. first_block + nblock - 1, its outlet static pressure p_out
// (config units).
struct HostInst {
    double p_out;
    int first_block, nblock;
};

}  // namespace

struct fv3d_ctx {
    fv3d_config cfg;
    Gas gas;
    double temp_ref, press_ref, vel_ref, mue_ref, dens_ref, re_ref, len_ref;
    int device;
    hipStream_t stream = nullptr;
    std::vector<HostBlock> blocks;
    // instances: declared with fv3d_instance_add, else one holding every block
    std::vector<HostInst> insts;
    bool explicit_inst = false;
    int ninst = 1;
    std::vector<double> inst_vol;    // total cell volume per instance
    std::vector<int2> parts_host;    // norm partials of each instance (first, count)
    int2 *d_parts = nullptr;
    std::vector<HostPatch> patches;
    std::vector<HostConn> conns;
    std::vector<HostConn> mixes;     // mixing planes: block, rng of both sides (kind, face, orient unused)
    MixDev *d_mix = nullptr;
    int nmix = 0, nmix_lines = 0;
    bool ready = false;
    std::string err;
    // device work lists
    BlockDev *d_blocks = nullptr;
    PatchDev *d_patches = nullptr;
    // interior / periodic exchange: element list, periodic flip table and the
    // per-block pitch-angle sign (device state that persists over iterations)
    ExchDev *d_exch = nullptr;
    long nexch = 0;
    PeriodicDev *d_per = nullptr;
    int nper = 0;
    int *d_pitch_sign = nullptr;
    std::vector<PatchDev> pd_host;   // the device patch list, host copy
    double *d_mbox = nullptr;         // sweep mailboxes of all blocks
    long mbox_words = 0;
    long *d_pack_start = nullptr;     // running pack (tile, chunk) item counts per block
    long pack_groups = 0;
    // Baldwin-Lomax: one launch per group size G (16, 32 or 64 cell groups per
    // line, by line length) over every wall that needs it; d_bl_walls holds
    // (patch index, line count) pairs, launch q at offset bl_launch[q].first
    // with bl_launch_n[q] walls and bl_launch[q].second lines
    int2 *d_bl_walls = nullptr;
    std::vector<std::pair<int, long>> bl_launch;
    std::vector<int> bl_launch_n;
    std::vector<int> bl_groups;
    long *d_pstart_bc = nullptr, *d_cstart = nullptr, *d_pstart_plane = nullptr;
    // residual tiles: one list per tile shape (built on first use), and the
    // shape each residual kernel uses (0 convective, 1 viscous, 2 nu~)
    int4 *d_tiles_s[kNumTiles] = {};
    int ntiles_s[kNumTiles] = {};
    int tile_sel[3] = {0, 0, 0};
    // streaming convective residual: column tiles per configuration (built on
    // first use); tile_sel[0] >= kNumTiles selects configuration
    // tile_sel[0] - kNumTiles of it (face capture always runs the per-tile kernel)
    int4 *d_stiles[kNumStream] = {};
    int nstiles[kNumStream] = {};
    // face-parallel convective form (configuration kConvFaces, built on first
    // use): running face counts per (block, direction), and their total
    long *d_fstart = nullptr;
    long nfaces = 0;
    int2 *d_secs = nullptr;
    int npatch_dev = 0, nplanes = 0;
    long nbc = 0, ncell = 0, nplane_threads = 0;
    long ncell_real = 0;             // interior cells (ncell also counts the instances' padding)
    double *d_partial = nullptr;
    int npart = 0;
    double *d_sums = nullptr;
    int *d_bad = nullptr;
    int *d_flags = nullptr;
    ScalarsDev *d_sc = nullptr;
    double *d_hist = nullptr;
    long hist_cap = 0;
    double *d_stage_row = nullptr;   // history rows (one per instance) written by the captured graph
    double *d_faces[3] = {nullptr, nullptr, nullptr};
    ViscDebug vdbg = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}, nullptr};
    int store_faces = 0;
    double total_vol = 0.0;
    // persistent wavefront sweep: tile lists (forward, backward) and the
    // sync words [err, ticket_f, ticket_b, done_f[ntiles], done_b[ntiles]]
    KbaTile *d_kba[2] = {nullptr, nullptr};
    int nkba = 0;
    int *d_kba_sync = nullptr;
    int sweep_mode = 4;   // 0 one launch per hyperplane (the reference's schedule, a cross-check),
                          // 4 the persistent wavefront on the tile-skewed row layout
    int kba_grid = 0;
    // two-grid cycle: the coarse level is a context of its own on the same
    // stream (coarse is set on the fine level, parent on the coarse one).  The
    // CFL number, its ramp threshold and the iteration count are shared: they
    // live in the fine level's scalars (ramp_sc, one per instance), which the
    // blocks of both levels point at (BlockDev::cfl).
    fv3d_ctx *coarse = nullptr;
    fv3d_ctx *parent = nullptr;
    int4 *d_mg_tiles = nullptr;      // prolongation tiles over the coarse blocks
    int n_mg_tiles = 0;
    int *d_axial = nullptr;          // radial-equilibrium outlet patches (indices into d_patches)
    int n_axial = 0;
    unsigned long long *d_dtmin = nullptr;   // per-block minimum time step (global time step)
    // Spalart-Allmaras: ghost-cell work list (2 (nj nk + ni nk + ni nj) per
    // block), per-block flags of the nu~ limiter, and whether nu~ was given
    // (fv3d_nut_upload or fv3d_initial)
    long *d_gstart = nullptr;
    long nghost = 0;
    int *d_sa_hi = nullptr;
    bool ev_set = false;
    // learned closure: FP32 parameters, Adam moments, gradient, gate bounds,
    // gate counter, training scratch; closure on = mue_t from the network in
    // place of the nu~ transport
    float *d_mlp = nullptr, *d_mlp_m = nullptr, *d_mlp_v = nullptr, *d_mlp_g = nullptr, *d_mlp_bounds = nullptr;
    unsigned long long *d_mlp_tripped = nullptr;
    float *d_mlp_xc = nullptr;   // closure: 8 floats per cell (features, gate flag)
    long mlp_step = 0;
    bool mlp_closure = false;
    double mlp_relax = 1.0;          // under-relaxation of the closure's mue_t
    bool owns_stream = true;
    bool mg_coarse = false;          // allocate cv_2h / Res_2h (coarse level)
    bool mg_fas = false;             // a level of the N-level V-cycle: allocate the forcing too
    int ng = 1;                      // 2 when a coarser level exists
    int nlev = 1;                    // levels from this one down
    long iter_host = 0;              // iterations run (host copy, picks the cycle phase)
    bool initial_done = false;       // fv3d_initial ran (the two-grid cycle needs it)
    ScalarsDev *ramp_sc = nullptr;
    // graph
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_coarse = nullptr;   // coarse start-up phase of the two-grid cycle
    // timing
    int timing = 0;
    hipEvent_t ev[14];
    double t_ms[FV3D_TIMING_SLOTS] = {};
    long t_n[FV3D_TIMING_SLOTS] = {};
};

static int fail(fv3d_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCK(c, x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail((c), FV3D_EHIP, "%s: %s", #x, hipGetErrorString(e_));       \
    } while (0)

template <class T>
static int dalloc(fv3d_ctx *c, T **p, size_t n)
{
    HIPCK(c, hipMalloc((void **)p, n * sizeof(T) + 64));
    HIPCK(c, hipMemsetAsync(*p, 0, n * sizeof(T) + 64, c->stream));
    return 0;
}

// ---------------------------------------------------------------- set-up
// Non-dimensionalisation (source/UserInput.cpp:60-76, 1070-1092).  Config
// values are parsed by the reference with stof, hence the float rounding.
static void nondimensionalize(fv3d_ctx *c)
{
    const fv3d_config &f = c->cfg;
    double gamma = (float)f.gamma, rgas = (float)f.rgas;
    double ttin = (float)f.tt_in, ptin = (float)f.pt_in, pout = (float)f.p_out;
    c->len_ref = (float)f.len_ref;
    c->temp_ref = ttin;
    c->press_ref = ptin;
    c->vel_ref = sqrt(gamma * rgas * c->temp_ref);
    c->mue_ref = F32C(1.716e-5) * pow(c->temp_ref / F32C(288.15), F32C(3.0) / F32C(2.0)) * (double)(288.15f + 110.4f) /
                 (c->temp_ref + F32C(110.4));
    c->dens_ref = c->press_ref / rgas / c->temp_ref;
    c->re_ref = c->dens_ref * c->vel_ref * c->len_ref / c->mue_ref;
    Gas &g = c->gas;
    g.gamma = gamma;
    g.rgas = F32C(1.0) / (F32C(1.0) * F32C(1.0) * gamma);
    g.pt_in = F32C(1.0) / gamma;
    g.tt_in = F32C(1.0);
    g.p_out = pout / (c->dens_ref * c->vel_ref * c->vel_ref);
    for (int d = 0; d < 3; ++d) g.inlet_dir[d] = f.inlet_dir[d];
    g.c_timestep = (float)f.c_timestep;
    g.relax = (float)f.relax;
    g.recon_var = f.recon_var;
    g.recon_scheme = f.recon_scheme;
    g.flux_scheme = f.flux_scheme;
    g.viscous = f.viscous ? 1 : 0;
    g.turbulence = f.turbulence;
    g.re_ref = c->re_ref;
    g.pr_lam = (float)f.pr_lam;
    g.pr_tur = (float)f.pr_tur;
    g.ev_ratio = (float)f.ev_ratio;
}

// Non-dimensional outlet static pressure of instance w (the config's p_out is
// instance 1's unless instances were declared), as nondimensionalize forms it.
static double inst_p_out(const fv3d_ctx *c, int w)
{
    const double pout = (float)c->insts[w].p_out;
    return pout / (c->dens_ref * c->vel_ref * c->vel_ref);
}

// source/GeometryVariables.cpp:10-105
static void face_vector_volume(HostBlock &b, std::vector<double> &fS0, std::vector<double> &fS1, std::vector<double> &fS2,
                               std::vector<double> &fs0, std::vector<double> &fs1, std::vector<double> &fs2,
                               std::vector<double> &fc0, std::vector<double> &fc1, std::vector<double> &fc2,
                               std::vector<double> &vol, std::vector<double> &cc, std::vector<double> &node)
{
    const int IM = b.IM, JM = b.JM, KM = b.KM;
    const long box = (long)(IM + 1) * (JM + 1) * (KM + 1);
    const int sj = KM + 1, si = (JM + 1) * (KM + 1);
    auto X = [&](int i, int j, int k) { return (long)i * si + (long)j * sj + k; };
    auto N = [&](int i, int j, int k, int c) { return b.nodes[(((long)(i - 1) * JM + (j - 1)) * KM + (k - 1)) * 3 + c]; };
    for (auto *v : {&fS0, &fS1, &fS2, &fc0, &fc1, &fc2, &cc, &node}) v->assign(3 * box, 0.0);
    for (auto *v : {&fs0, &fs1, &fs2, &vol}) v->assign(box, 0.0);
    for (int i = 1; i <= IM; ++i)
        for (int j = 1; j <= JM; ++j)
            for (int k = 1; k <= KM; ++k)
                for (int c = 0; c < 3; ++c) node[c * box + X(i, j, k)] = N(i, j, k, c);
    struct V { double x, y, z; };
    auto nd = [&](int i, int j, int k) { return V{N(i, j, k, 0), N(i, j, k, 1), N(i, j, k, 2)}; };
    auto sub = [](V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto add = [](V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; };
    auto face = [&](std::vector<double> &S, std::vector<double> &s, std::vector<double> &cen, long x, V A, V B, V n1, V n2,
                    V n3, V n4) {
        double Sx = F32C(0.5) * (A.y * B.z - B.y * A.z);
        double Sy = F32C(0.5) * (A.z * B.x - A.x * B.z);
        double Sz = F32C(0.5) * (A.x * B.y - B.x * A.y);
        V cs = add(add(add(n1, n2), n3), n4);
        S[x] = Sx; S[box + x] = Sy; S[2 * box + x] = Sz;
        cen[x] = cs.x * 0.25; cen[box + x] = cs.y * 0.25; cen[2 * box + x] = cs.z * 0.25;
        s[x] = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
    };
    for (int i = 1; i <= IM; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k)
                face(fS0, fs0, fc0, X(i, j, k), sub(nd(i, j + 1, k), nd(i, j, k + 1)), sub(nd(i, j + 1, k + 1), nd(i, j, k)),
                     nd(i, j, k), nd(i, j + 1, k), nd(i, j + 1, k + 1), nd(i, j, k + 1));
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM; ++j)
            for (int k = 1; k <= KM - 1; ++k)
                face(fS1, fs1, fc1, X(i, j, k), sub(nd(i, j, k + 1), nd(i + 1, j, k)), sub(nd(i + 1, j, k + 1), nd(i, j, k)),
                     nd(i, j, k), nd(i + 1, j, k), nd(i + 1, j, k + 1), nd(i, j, k + 1));
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM; ++k)
                face(fS2, fs2, fc2, X(i, j, k), sub(nd(i + 1, j + 1, k), nd(i, j, k)), sub(nd(i, j + 1, k), nd(i + 1, j, k)),
                     nd(i, j, k), nd(i + 1, j, k), nd(i + 1, j + 1, k), nd(i, j + 1, k));
    b.cellvol = 0.0;
    auto dotf = [&](std::vector<double> &cen, std::vector<double> &S, long x) {
        return cen[x] * S[x] + cen[box + x] * S[box + x] + cen[2 * box + x] * S[2 * box + x];
    };
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                long x = X(i, j, k);
                double v = dotf(fc0, fS0, X(i + 1, j, k)) - dotf(fc0, fS0, x) + dotf(fc1, fS1, X(i, j + 1, k)) - dotf(fc1, fS1, x) +
                           dotf(fc2, fS2, X(i, j, k + 1)) - dotf(fc2, fS2, x);
                vol[x] = v / F32C(3.0);
                V s = add(nd(i, j, k), nd(i + 1, j, k));
                s = add(s, nd(i, j + 1, k)); s = add(s, nd(i + 1, j + 1, k));
                s = add(s, nd(i, j, k + 1)); s = add(s, nd(i + 1, j, k + 1));
                s = add(s, nd(i, j + 1, k + 1)); s = add(s, nd(i + 1, j + 1, k + 1));
                cc[x] = s.x * 0.125; cc[box + x] = s.y * 0.125; cc[2 * box + x] = s.z * 0.125;
                b.cellvol = b.cellvol + vol[x];
            }
}

template <class T>
static int upload(fv3d_ctx *c, HostBlock &b, T **dst, const std::vector<T> &src)
{
    int rc = dalloc(c, dst, src.size());
    if (rc) return rc;
    b.allocs.push_back(*dst);
    HIPCK(c, hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return 0;
}

// The residual tiles of shape q over every block (block, first i, j, k), on
// the device; built once per shape.
static int ensure_tiles(fv3d_ctx *c, int q)
{
    if (c->d_tiles_s[q]) return FV3D_OK;
    const int TI = kTileShape[q][0], TJ = kTileShape[q][1], TK = kTileShape[q][2];
    std::vector<int4> tiles;
    for (int n = 0; n < (int)c->blocks.size(); ++n) {
        const BlockDev &B = c->blocks[n].dev;
        for (int i0 = 0; i0 < B.ni; i0 += TI)
            for (int j0 = 0; j0 < B.nj; j0 += TJ)
                for (int k0 = 0; k0 < B.nk; k0 += TK) tiles.push_back(make_int4(n, i0, j0, k0));
    }
    // XCD order: the hardware hands workgroup w to XCD w mod 8, so entry w is
    // tile (w mod 8) * nper + w / 8 and each XCD works through one contiguous
    // eighth of the list (neighbouring tiles share their halos through that
    // XCD's L2); entries past the list have block -1 and do nothing
    const size_t n = tiles.size(), nper = (n + 7) / 8;
    std::vector<int4> xt(8 * nper, make_int4(-1, 0, 0, 0));
    for (size_t w = 0; w < xt.size(); ++w) {
        const size_t idx = (w % 8) * nper + w / 8;
        if (idx < n) xt[w] = tiles[idx];
    }
    int rc = dalloc(c, &c->d_tiles_s[q], xt.size());
    if (rc) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_tiles_s[q], xt.data(), xt.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    c->ntiles_s[q] = (int)xt.size();
    return FV3D_OK;
}

// The column tiles of streaming configuration q (block, first i of the chunk,
// first j, first k), in the order the XCD mapping of k_conv_stream assumes.
static int ensure_stream_tiles(fv3d_ctx *c, int q)
{
    if (c->d_stiles[q]) return FV3D_OK;
    const int TJ = kStreamShape[q][0], TK = kStreamShape[q][1], CI = kStreamShape[q][2];
    std::vector<int4> tiles;
    for (int n = 0; n < (int)c->blocks.size(); ++n) {
        const BlockDev &B = c->blocks[n].dev;
        for (int i0 = 0; i0 < B.ni; i0 += CI)
            for (int j0 = 0; j0 < B.nj; j0 += TJ)
                for (int k0 = 0; k0 < B.nk; k0 += TK) tiles.push_back(make_int4(n, i0, j0, k0));
    }
    int rc = dalloc(c, &c->d_stiles[q], tiles.size());
    if (rc) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_stiles[q], tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice,
                            c->stream));
    c->nstiles[q] = (int)tiles.size();
    return FV3D_OK;
}

// The face-parallel convective form: per block the face-flux array
// (15 box-sized slabs, direction-major) and the face work list; the block
// table on the device is uploaded again with the new pointers.  Allocates, so
// it runs at set-up or from fv3d_set_tiles / fv3d_autotune, never inside a
// captured iteration.
static int ensure_face_form(fv3d_ctx *c)
{
    if (c->d_fstart) return FV3D_OK;
    std::vector<long> fst{0};
    for (HostBlock &b : c->blocks) {
        BlockDev &B = b.dev;
        int rc = dalloc(c, &B.fcv, 15 * (size_t)B.box);
        if (rc) return rc;
        b.allocs.push_back(B.fcv);
        fst.push_back(fst.back() + (long)(B.ni + 1) * B.nj * B.nk);
        fst.push_back(fst.back() + (long)B.ni * (B.nj + 1) * B.nk);
        fst.push_back(fst.back() + (long)B.ni * B.nj * (B.nk + 1));
    }
    int rc = dalloc(c, &c->d_fstart, fst.size());
    if (rc) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_fstart, fst.data(), fst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    c->nfaces = fst.back();
    if (c->d_blocks) {
        std::vector<BlockDev> bd;
        for (HostBlock &b : c->blocks) bd.push_back(b.dev);
        HIPCK(c, hipMemcpyAsync(c->d_blocks, bd.data(), bd.size() * sizeof(BlockDev), hipMemcpyHostToDevice, c->stream));
    }
    return FV3D_OK;
}

static int setup_block(fv3d_ctx *c, HostBlock &b, int index)
{
    BlockDev &B = b.dev;
    memset(&B, 0, sizeof B);
    B.IM = b.IM; B.JM = b.JM; B.KM = b.KM;
    B.ni = b.IM - 1; B.nj = b.JM - 1; B.nk = b.KM - 1;
    B.sj = b.KM + 1; B.si = (b.JM + 1) * (b.KM + 1);
    B.box = (long)(b.IM + 1) * (b.JM + 1) * (b.KM + 1);
    B.nblade = b.nblade;
    B.w_rot = b.w_rot;
    std::vector<double> fS[3], fs[3], fc[3], vol, cc, node;
    face_vector_volume(b, fS[0], fS[1], fS[2], fs[0], fs[1], fs[2], fc[0], fc[1], fc[2], vol, cc, node);
    int rc;
    double *p;
    for (int d = 0; d < 3; ++d) {
        if ((rc = upload(c, b, &p, fS[d]))) return rc; B.fS[d] = p;
        if ((rc = upload(c, b, &p, fs[d]))) return rc; B.fs[d] = p;
        if ((rc = upload(c, b, &p, fc[d]))) return rc; B.fc[d] = p;
    }
    if ((rc = upload(c, b, &p, vol))) return rc; B.vol = p;
    if ((rc = upload(c, b, &p, cc))) return rc; B.cc = p;
    if ((rc = upload(c, b, &p, node))) return rc; B.node = p;
    auto alloc = [&](double **q, long n) { int r = dalloc(c, q, n); if (!r) b.allocs.push_back(*q); return r; };
    if ((rc = alloc(&B.Q, 5 * B.box))) return rc;
    if ((rc = alloc(&B.W, 6 * B.box))) return rc;
    if ((rc = alloc(&B.lamc, 3 * B.box))) return rc;
    if ((rc = alloc(&B.dt, B.box))) return rc;
    if (c->cfg.viscous) {
        // filled by k_metrics at the end of set-up
        for (int d = 0; d < 3; ++d) {
            if ((rc = alloc(&p, 9 * B.box))) return rc;
            B.fg[d] = p;
        }
        if ((rc = alloc(&p, 9 * B.box))) return rc;
        B.cg = p;
        if ((rc = alloc(&B.lamv, 3 * B.box))) return rc;
        if ((rc = alloc(&B.omega, B.box))) return rc;
    }
    if (c->cfg.turbulence) if ((rc = alloc(&B.mut, B.box))) return rc;
    if (c->cfg.turbulence == 200) {
        if ((rc = alloc(&B.ev, B.box))) return rc;
        if ((rc = alloc(&B.ev_old, B.box))) return rc;
        if ((rc = alloc(&B.dev, B.box))) return rc;
        if ((rc = alloc(&B.res_ev, B.box))) return rc;
        if ((rc = alloc(&B.dwall, B.box))) return rc;
        if ((rc = alloc(&B.alpha2, B.box))) return rc;
    }
    if (c->cfg.turbulence == 100) {
        unsigned long long *f;
        if ((rc = dalloc(c, &f, B.box))) return rc;
        b.allocs.push_back(f);
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((B.box + 255) / 256)), dim3(256), 0, c->stream, f, B.box,
                           kBlUnvisited);
        HIPCK(c, hipGetLastError());
        B.blmin = f;
    }
    if ((rc = alloc(&B.alpha, B.box))) return rc;
    {
        // static sweep metrics of each cell: averaged unit normal of its two
        // faces along d (upper face first), averaged area and the normal grid
        // velocity at its centroid (source/LU_SGS_Marching.cpp:301-341)
        std::vector<double> swg(15 * B.box, 0.0);
        const long st[3] = {B.si, B.sj, 1};
        for (int i = 1; i <= B.ni; ++i)
            for (int j = 1; j <= B.nj; ++j)
                for (int k = 1; k <= B.nk; ++k) {
                    long x = cidx(B, i, j, k);
                    for (int d = 0; d < 3; ++d) {
                        long fa = x + st[d], fb = x;
                        const std::vector<double> &S = fS[d], &s = fs[d];
                        double nx = (S[fa] / s[fa] + S[fb] / s[fb]) / F32C(2.0);
                        double ny = (S[B.box + fa] / s[fa] + S[B.box + fb] / s[fb]) / F32C(2.0);
                        double nz = (S[2 * B.box + fa] / s[fa] + S[2 * B.box + fb] / s[fb]) / F32C(2.0);
                        double uvel_e = F32C(-1.0) * cc[B.box + x] * B.w_rot, vvel_e = cc[x] * B.w_rot;
                        swg[(d * 5 + 0) * B.box + x] = nx;
                        swg[(d * 5 + 1) * B.box + x] = ny;
                        swg[(d * 5 + 2) * B.box + x] = nz;
                        swg[(d * 5 + 3) * B.box + x] = F32C(0.5) * (s[fb] + s[fa]);
                        swg[(d * 5 + 4) * B.box + x] = uvel_e * nx + vvel_e * ny + F32C(0.0) * nz;
                    }
                }
        if ((rc = upload(c, b, &p, swg))) return rc;
        B.swg = p;
    }
    if ((rc = alloc(&B.R, 5 * B.box))) return rc;
    if ((rc = alloc(&B.dQ, 5 * B.box))) return rc;
    if (c->mg_coarse) {
        if ((rc = alloc(&B.cv2h, 5 * B.box))) return rc;
        if ((rc = alloc(&B.r2h, 5 * B.box))) return rc;
    }
    if (c->mg_fas && (rc = alloc(&B.fasP, 5 * B.box))) return rc;
    {
        // tile-skewed sweep rows (k_sweep_fill at set-up, k_sweep_pack per iteration)
        B.ntk = (B.nk + kKbaTK - 1) / kKbaTK;
        B.nr = B.ni + kKbaTJ + kKbaTK - 2;
        const long ntile = (long)((B.nj + kKbaTJ - 1) / kKbaTJ) * B.ntk;
        b.sweep_rows = ntile * B.nr;
        b.pack_groups = ntile * ((B.ni + kPackChunk - 1) / kPackChunk);
        if ((rc = alloc(&B.srow, b.sweep_rows * kSweepQ * 64))) return rc;
        if ((rc = alloc(&B.sdq, b.sweep_rows * kSweepD * 64))) return rc;
        // the mailboxes of all blocks share one allocation (fv3d_setup), so
        // one fill resets them each iteration
    }
    // Dummy slabs: tangential boxes (JM+1)(KM+1), (IM+1)(KM+1), (IM+1)(JM+1)
    const int ta[3] = {b.JM, b.IM, b.IM}, tb[3] = {b.KM, b.KM, b.JM};
    for (int s = 0; s < 6; ++s) {
        int d = s / 2;
        B.Dt[s] = (long)(ta[d] + 1) * (tb[d] + 1);
        B.Db[s] = tb[d] + 1;
        if ((rc = alloc(&B.D[s], 2 * 11 * B.Dt[s]))) return rc;
    }
    // wall masks per side
    std::vector<unsigned char> mask[6];
    std::vector<double> rot[6];
    bool any[6] = {false, false, false, false, false, false};
    for (int s = 0; s < 6; ++s) { mask[s].assign(B.Dt[s], 0); rot[s].assign(B.Dt[s], 0.0); }
    for (const HostPatch &P : c->patches) {
        if (P.kind != 2 || P.block != index + 1) continue;
        int d = (P.face - 1) / 2, hi = (P.face - 1) % 2;
        int s = 2 * d + hi;
        const int *r = P.rng;
        int aS = d == 0 ? r[2] : r[0], aE = d == 0 ? r[3] : r[1];
        int bS = d == 2 ? r[2] : r[4], bE = d == 2 ? r[3] : r[5];
        double wr = P.rotate ? b.w_rot : 0.0;
        for (int a = aS; a <= aE; ++a)
            for (int bb = bS; bb <= bE; ++bb) {
                mask[s][(long)a * B.Db[s] + bb] = 1;
                rot[s][(long)a * B.Db[s] + bb] = wr;
            }
        any[s] = true;
    }
    for (int s = 0; s < 6; ++s) {
        if (!any[s]) { B.wmask[s] = nullptr; B.wrot[s] = nullptr; continue; }
        unsigned char *m;
        double *rr;
        if ((rc = upload(c, b, &m, mask[s]))) return rc;
        if ((rc = upload(c, b, &rr, rot[s]))) return rc;
        B.wmask[s] = m;
        B.wrot[s] = rr;
    }
    B.flag_low_dt = c->d_flags + index;
    return 0;
}

// Interior and periodic patches (CMesh::BndCondInterior / BndCondPeriodic,
// source/BndCondInteriorPeriodic.cpp, with ShadowOut / Trans / ShadowIn,
// source/ShadowOutIn.cpp).  The reference copies layer nd of one side's cells
// into a 2-D shadow array whose axes depend on the face type, re-indexes it by
// the orientation and writes it into layer nd of the other side's Dummy slab
// (and, for nd = 1, its ghost cells).  Here that index chain is resolved once
// into a list of (source cell, target slot) elements.
static void shadow_axes(int fc, int *al, int *am)
{
    static const int L[6] = {2, 1, 0, 2, 1, 0}, M[6] = {1, 2, 2, 0, 0, 1};
    *al = L[fc - 1];
    *am = M[fc - 1];
}

static int build_exchange(fv3d_ctx *c)
{
    std::vector<ExchDev> el;
    std::vector<PeriodicDev> per;
    // interiors first, then periodics (source/BoundaryCondition.cpp:46-58)
    for (int kind = 0; kind < 2; ++kind)
        for (const HostConn &t : c->conns) {
            if (t.kind != kind) continue;
            int pidx = -1;
            if (kind == 1) {
                // the pitch-angle flip test of source/BndCondInteriorPeriodic.cpp:54-59
                const HostBlock &A = c->blocks[t.block[0] - 1], &B = c->blocks[t.block[1] - 1];
                auto node = [](const HostBlock &b, int i, int j, int k, int q) {
                    return b.nodes[(((long)(i - 1) * b.JM + (j - 1)) * b.KM + (k - 1)) * 3 + q];
                };
                double th_1 = atan(node(A, t.rng[0][0], t.rng[0][2], t.rng[0][4], 1) / node(A, t.rng[0][0], t.rng[0][2], t.rng[0][4], 0));
                double th_2 = atan(node(B, t.rng[1][0], t.rng[1][2], t.rng[1][4], 1) / node(B, t.rng[1][0], t.rng[1][2], t.rng[1][4], 0));
                PeriodicDev q;
                q.block = t.block[0] - 1;
                q.flip = (th_2 - th_1) < 0.0;
                // delta_th = 2*PI/NBlade in float (source/UserInput.cpp:443)
                const double d = (double)(2.0f * 3.1415926535897932f / (float)A.nblade);
                q.cs[0] = cos(d); q.cs[1] = sin(d);
                const double md = F32C(0.0) - d;
                q.cs[2] = cos(md); q.cs[3] = sin(md);
                pidx = (int)per.size();
                per.push_back(q);
            }
            for (int s = 0; s < 2; ++s) {
                const HostBlock &A = c->blocks[t.block[s] - 1], &B = c->blocks[t.block[1 - s] - 1];
                const BlockDev &DA = A.dev, &DB = B.dev;
                const int fa = t.face[s], fb = t.face[1 - s];
                int al, am, bl, bm;
                shadow_axes(fa, &al, &am);
                shadow_axes(fb, &bl, &bm);
                const int L1 = t.rng[s][2 * al + 1] - t.rng[s][2 * al] + 1, M1 = t.rng[s][2 * am + 1] - t.rng[s][2 * am] + 1;
                const int L2 = (t.orient == 2 || t.orient == 4) ? M1 : L1, M2 = (t.orient == 2 || t.orient == 4) ? L1 : M1;
                const int bL = t.rng[1 - s][2 * bl + 1] - t.rng[1 - s][2 * bl] + 1;
                const int bM = t.rng[1 - s][2 * bm + 1] - t.rng[1 - s][2 * bm] + 1;
                if (L2 != bL || M2 != bM)
                    return fail(c, FV3D_EINVAL, "connection sides do not match under orientation %d", t.orient);
                const int da = (fa - 1) / 2, db = (fb - 1) / 2;
                const int NA = da == 0 ? A.IM : da == 1 ? A.JM : A.KM, NB = db == 0 ? B.IM : db == 1 ? B.JM : B.KM;
                for (int nd = 1; nd <= kND; ++nd)
                    for (int l1 = 1; l1 <= L1; ++l1)
                        for (int m1 = 1; m1 <= M1; ++m1) {
                            int l2, m2;
                            switch (t.orient) {
                            case 1: l2 = l1; m2 = M1 + 1 - m1; break;
                            case 2: l2 = m1; m2 = l1; break;
                            case 3: l2 = L1 + 1 - l1; m2 = m1; break;
                            default: l2 = M1 + 1 - m1; m2 = L1 + 1 - l1; break;
                            }
                            int xs[3], xt[3];
                            xs[al] = t.rng[s][2 * al] + l1 - 1;
                            xs[am] = t.rng[s][2 * am] + m1 - 1;
                            xs[da] = fa % 2 == 1 ? nd : NA - nd;
                            xt[bl] = t.rng[1 - s][2 * bl] + l2 - 1;
                            xt[bm] = t.rng[1 - s][2 * bm] + m2 - 1;
                            ExchDev e;
                            e.src_block = t.block[s] - 1;
                            e.dst_block = t.block[1 - s] - 1;
                            e.src = cidx(DA, xs[0], xs[1], xs[2]);
                            e.side = fb - 1;
                            e.layer = nd;
                            // Dummy tangential position: I sides (j,k), J sides (i,k), K sides (i,j)
                            const int ta = db == 0 ? xt[1] : xt[0], tb = db == 2 ? xt[1] : xt[2];
                            e.slot = (long)ta * DB.Db[fb - 1] + tb;
                            xt[db] = fb % 2 == 1 ? 0 : NB;
                            e.ghost = nd == 1 ? cidx(DB, xt[0], xt[1], xt[2]) : -1;
                            e.rot = kind == 1 ? (s == 0 ? 1 : 2) : 0;
                            e.per = pidx;
                            el.push_back(e);
                        }
            }
        }
    c->nexch = (long)el.size();
    c->nper = (int)per.size();
    int rc;
    if (!c->mixes.empty()) {
        std::vector<MixDev> ml;
        int line = 0;
        for (const HostConn &T : c->mixes) {
            MixDev m;
            m.b1 = T.block[0] - 1; m.b2 = T.block[1] - 1;
            m.i0 = T.rng[0][0]; m.ni = T.rng[0][1] - T.rng[0][0] + 1;
            m.j1 = T.rng[0][2]; m.nj1 = T.rng[0][3] - T.rng[0][2] + 1;
            m.j2 = T.rng[1][2]; m.nj2 = T.rng[1][3] - T.rng[1][2] + 1;
            m.line0 = line;
            line += m.ni;
            ml.push_back(m);
        }
        c->nmix = (int)ml.size();
        c->nmix_lines = line;
        if ((rc = dalloc(c, &c->d_mix, ml.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(c->d_mix, ml.data(), ml.size() * sizeof(MixDev), hipMemcpyHostToDevice, c->stream));
    }
    if (!el.empty()) {
        if ((rc = dalloc(c, &c->d_exch, el.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(c->d_exch, el.data(), el.size() * sizeof(ExchDev), hipMemcpyHostToDevice, c->stream));
    }
    if (!per.empty()) {
        if ((rc = dalloc(c, &c->d_per, per.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(c->d_per, per.data(), per.size() * sizeof(PeriodicDev), hipMemcpyHostToDevice, c->stream));
        std::vector<int> sg(c->blocks.size(), 1);
        if ((rc = dalloc(c, &c->d_pitch_sign, sg.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(c->d_pitch_sign, sg.data(), sg.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    return FV3D_OK;
}

// Whether the nu~ equation is solved (Spalart-Allmaras without the learned
// closure in its place).
static bool sa_transport(const fv3d_ctx *c) { return c->cfg.turbulence == 200 && !c->mlp_closure; }

// Spalart-Allmaras set-up: the ghost-cell work list of k_sa_ghost, the
// limiter flags and the wall distance of every cell (CMesh::WallDistanceSA,
// source/WallDistanceSA.cpp:9-132): the centroids of all wall faces of all
// blocks of an instance are listed on the host and k_wall_distance takes the
// minimum over its instance's list.
static int setup_sa(fv3d_ctx *c)
{
    const int nb = (int)c->blocks.size();
    int rc;
    std::vector<long> gst{0};
    for (const HostBlock &b : c->blocks) {
        const BlockDev &B = b.dev;
        gst.push_back(gst.back() + 2L * ((long)B.nj * B.nk + (long)B.ni * B.nk + (long)B.ni * B.nj));
    }
    c->nghost = gst.back();
    if ((rc = dalloc(c, &c->d_gstart, gst.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_gstart, gst.data(), gst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    if ((rc = dalloc(c, &c->d_sa_hi, nb))) return rc;
    std::vector<std::vector<double>> fcen(nb);   // face centroids of the blocks that hold walls, 3 directions
    std::vector<double> wf;
    // the wall faces of each instance, one instance after the other
    std::vector<int> wf0(c->ninst, 0), nwf(c->ninst, 0);
    for (int w = 0; w < c->ninst; ++w) {
        wf0[w] = (int)(wf.size() / 3);
        for (const HostPatch &P : c->patches) {
            if (P.kind != 2 || c->blocks[P.block - 1].inst != w) continue;
            const int n = P.block - 1, d = (P.face - 1) / 2;
            HostBlock &b = c->blocks[n];
            const long box = b.dev.box;
            if (fcen[n].empty()) {
                std::vector<double> fS[3], fs[3], fc[3], vol, cc, node;
                face_vector_volume(b, fS[0], fS[1], fS[2], fs[0], fs[1], fs[2], fc[0], fc[1], fc[2], vol, cc, node);
                fcen[n].resize(9 * box);
                for (int q = 0; q < 3; ++q) std::copy(fc[q].begin(), fc[q].end(), fcen[n].begin() + 3 * q * box);
            }
            const double *F = fcen[n].data() + 3 * d * box;
            for (int i = P.rng[0]; i <= P.rng[1]; ++i)
                for (int j = P.rng[2]; j <= P.rng[3]; ++j)
                    for (int k = P.rng[4]; k <= P.rng[5]; ++k) {
                        const long x = cidx(b.dev, i, j, k);
                        wf.push_back(F[x]);
                        wf.push_back(F[box + x]);
                        wf.push_back(F[2 * box + x]);
                    }
        }
        nwf[w] = (int)(wf.size() / 3) - wf0[w];
    }
    std::vector<BlockDev> bd;
    for (HostBlock &b : c->blocks) {
        b.dev.wf0 = wf0[b.inst];
        b.dev.nwf = nwf[b.inst];
        bd.push_back(b.dev);
    }
    HIPCK(c, hipMemcpyAsync(c->d_blocks, bd.data(), bd.size() * sizeof(BlockDev), hipMemcpyHostToDevice, c->stream));
    const int nf = (int)(wf.size() / 3);
    double *d_wf = nullptr;
    if (nf > 0) {
        if ((rc = dalloc(c, &d_wf, wf.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(d_wf, wf.data(), wf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_wall_distance, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                       c->d_cstart, nb, d_wf);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (d_wf) HIPCK(c, hipFree(d_wf));
    return FV3D_OK;
}

// ------------------------------------------- scratch of fv3d_mlp_grad_noalloc
// The partial gradients and losses of up to kNoAllocChunks 64-sample chunks
// (8833 + 1 floats each) and the loss sum live in a file-scope device array,
// so fv3d_mlp_grad_noalloc allocates nothing.  Every GPU holds its own copy
// (9 MB) until the process ends; host code uses only the per-device address
// hipGetSymbolAddress returns, taken once per device at context creation.
constexpr int kNoAllocChunks = 256;   // 16,384 samples
constexpr long kNoAllocFloats = (long)kNoAllocChunks * (kMlpParams + 1) + 1;
static __device__ __align__(256) float scratch[kNoAllocFloats];
constexpr int kNoAllocDevices = 64;
static float *g_scratch_ptr[kNoAllocDevices];
static std::mutex g_scratch_mu;

static int noalloc_init(fv3d_ctx *c)
{
    if (c->device < 0 || c->device >= kNoAllocDevices)
        return fail(c, FV3D_EINVAL, "device %d beyond the %d the library tracks", c->device, kNoAllocDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return FV3D_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(c, FV3D_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = (float *)p;
    return FV3D_OK;
}

__global__ void k_noalloc_loss(const float *__restrict__ sum, long n, float *__restrict__ loss)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *loss = *sum / (float)n;
}

// ------------------------------------------------------------------ API
extern "C" {

int fv3d_ctx_create(const fv3d_config *cfg, fv3d_ctx **out)
{
    if (!cfg || !out) return FV3D_EINVAL;
    *out = nullptr;
    fv3d_ctx *c = new fv3d_ctx();
    c->cfg = *cfg;
    if (cfg->recon_scheme < 100 || cfg->recon_scheme > 500 || cfg->recon_scheme % 100) { *out = c; return fail(c, FV3D_EINVAL, "recon_scheme %d unsupported", cfg->recon_scheme); }
    if (cfg->flux_scheme != 200 && cfg->flux_scheme != 300) { *out = c; return fail(c, FV3D_EINVAL, "flux_scheme %d unsupported", cfg->flux_scheme); }
    if (cfg->turbulence != 0 && cfg->turbulence != 100 && cfg->turbulence != 200) { *out = c; return fail(c, FV3D_EINVAL, "turbulence model %d not in this build", cfg->turbulence); }
    if (cfg->turbulence && !cfg->viscous) { *out = c; return fail(c, FV3D_EINVAL, "a turbulence model needs the Navier-Stokes equations"); }
    if (cfg->precision != FV3D_PLAIN_FP64 && cfg->precision != FV3D_DEFECT_FP32) { *out = c; return fail(c, FV3D_EINVAL, "precision mode %d", cfg->precision); }
    c->device = cfg->device;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) { *out = c; return fail(c, FV3D_EHIP, "hipSetDevice(%d): %s", c->device, hipGetErrorString(e)); }
    if (int rc = noalloc_init(c)) { *out = c; return rc; }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { *out = c; return fail(c, FV3D_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    for (int n = 0; n < 14; ++n) (void)hipEventCreate(&c->ev[n]);
    nondimensionalize(c);
    *out = c;
    return FV3D_OK;
}

int fv3d_ctx_destroy(fv3d_ctx *c)
{
    if (!c) return FV3D_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->coarse) {
        fv3d_ctx_destroy(c->coarse);
        c->coarse = nullptr;
    }
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph_coarse) (void)hipGraphExecDestroy(c->graph_coarse);
    if (c->d_bl_walls) (void)hipFree(c->d_bl_walls);
    if (c->d_exch) (void)hipFree(c->d_exch);
    if (c->d_mix) (void)hipFree(c->d_mix);
    if (c->d_per) (void)hipFree(c->d_per);
    if (c->d_pitch_sign) (void)hipFree(c->d_pitch_sign);
    if (c->d_mbox) (void)hipFree(c->d_mbox);
    if (c->d_mg_tiles) (void)hipFree(c->d_mg_tiles);
    if (c->d_axial) (void)hipFree(c->d_axial);
    if (c->d_dtmin) (void)hipFree(c->d_dtmin);
    if (c->d_gstart) (void)hipFree(c->d_gstart);
    if (c->d_parts) (void)hipFree(c->d_parts);
    for (void *p : {(void *)c->d_mlp, (void *)c->d_mlp_m, (void *)c->d_mlp_v, (void *)c->d_mlp_g, (void *)c->d_mlp_bounds,
                    (void *)c->d_mlp_tripped, (void *)c->d_mlp_xc})
        if (p) (void)hipFree(p);
    if (c->d_sa_hi) (void)hipFree(c->d_sa_hi);
    if (c->d_pack_start) (void)hipFree(c->d_pack_start);
    if (c->d_fstart) (void)hipFree(c->d_fstart);
    for (int d = 0; d < 3; ++d) {
        if (c->vdbg.fv[d]) (void)hipFree(c->vdbg.fv[d]);
        if (c->vdbg.grad[d]) (void)hipFree(c->vdbg.grad[d]);
    }
    if (c->vdbg.fvc) (void)hipFree(c->vdbg.fvc);
    for (HostBlock &b : c->blocks)
        for (void *p : b.allocs) (void)hipFree(p);
    for (int q = 0; q < kNumTiles; ++q)
        if (c->d_tiles_s[q]) (void)hipFree(c->d_tiles_s[q]);
    for (int q = 0; q < kNumStream; ++q)
        if (c->d_stiles[q]) (void)hipFree(c->d_stiles[q]);
    void *ps[] = {c->d_blocks, c->d_patches, c->d_pstart_bc, c->d_cstart, c->d_pstart_plane, c->d_secs,
                  c->d_partial, c->d_sums, c->d_bad, c->d_flags, c->d_sc, c->d_hist, c->d_stage_row, c->d_kba[0], c->d_kba[1], c->d_kba_sync, c->d_faces[0], c->d_faces[1], c->d_faces[2]};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (c->stream) {
        for (int n = 0; n < 14; ++n) (void)hipEventDestroy(c->ev[n]);
        if (c->owns_stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return FV3D_OK;
}

const char *fv3d_last_error(const fv3d_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fv3d_block_add(fv3d_ctx *c, int IM, int JM, int KM, const double *nodes, int nblade, double rpm)
{
    if (!c || !nodes) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "block_add after setup");
    if (IM < 4 || JM < 4 || KM < 4) return fail(c, FV3D_EINVAL, "block %dx%dx%d: need at least 3 cells per direction", IM, JM, KM);
    if ((int)c->blocks.size() >= kMaxBlocks) return fail(c, FV3D_EINVAL, "more than %d blocks", kMaxBlocks);
    HostBlock b;
    b.IM = IM; b.JM = JM; b.KM = KM; b.nblade = nblade; b.rpm = rpm;
    b.nodes.assign(nodes, nodes + (size_t)IM * JM * KM * 3);
    // source/UserInput.cpp:444 and 1088
    b.w_rot = (double)(2.0f * 3.1415926535897932f) * rpm / F32C(60.0);
    b.w_rot = b.w_rot / (c->vel_ref / c->len_ref);
    if (c->explicit_inst) {
        b.inst = (int)c->insts.size() - 1;
        c->insts.back().nblock++;
    }
    c->blocks.push_back(std::move(b));
    return (int)c->blocks.size();
}

int fv3d_instance_add(fv3d_ctx *c, double p_out)
{
    if (!c) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "instance_add after setup");
    if (!c->explicit_inst && !c->blocks.empty())
        return fail(c, FV3D_ESTATE, "instances are declared before their blocks: %d blocks were added first",
                    (int)c->blocks.size());
    if (!(p_out > 0.0)) return fail(c, FV3D_EINVAL, "instance outlet pressure %g", p_out);
    c->explicit_inst = true;
    c->insts.push_back(HostInst{p_out, (int)c->blocks.size(), 0});
    return (int)c->insts.size();
}

int fv3d_instances(const fv3d_ctx *c)
{
    if (!c) return FV3D_EINVAL;
    return c->ready ? c->ninst : (c->explicit_inst ? (int)c->insts.size() : 1);
}

int fv3d_mixing_plane_add(fv3d_ctx *c, int block1, const int *rng1, int block2, const int *rng2)
{
    if (!c || !rng1 || !rng2) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "mixing_plane_add after setup");
    const int blk[2] = {block1, block2};
    for (int s = 0; s < 2; ++s)
        if (blk[s] < 1 || blk[s] > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "mixing plane block %d", blk[s]);
    const HostBlock &a = c->blocks[block1 - 1], &b = c->blocks[block2 - 1];
    if (rng1[4] != a.KM || rng1[5] != a.KM || rng2[4] != 1 || rng2[5] != 1)
        return fail(c, FV3D_EINVAL, "a mixing plane joins block 1's K-high face to block 2's K-low face");
    if (rng1[0] < 1 || rng1[0] > rng1[1] || rng1[1] > a.IM - 1 || rng1[1] > b.IM - 1)
        return fail(c, FV3D_EINVAL, "mixing plane radial range %d..%d outside both blocks", rng1[0], rng1[1]);
    if (rng1[2] < 1 || rng1[2] > rng1[3] || rng1[3] > a.JM - 1 || rng2[2] < 1 || rng2[2] > rng2[3] ||
        rng2[3] > b.JM - 1)
        return fail(c, FV3D_EINVAL, "mixing plane pitch-wise range outside its block");
    if (rng1[3] - rng1[2] + 1 > kMixMaxJ || rng2[3] - rng2[2] + 1 > kMixMaxJ)
        return fail(c, FV3D_EINVAL, "mixing plane wider than %d cells pitch-wise", kMixMaxJ);
    if (a.inst != b.inst) return fail(c, FV3D_EINVAL, "a mixing plane between blocks of different instances");
    HostConn t;
    t.kind = 2; t.orient = 0;
    t.block[0] = block1; t.block[1] = block2;
    t.face[0] = 6; t.face[1] = 5;
    memcpy(t.rng[0], rng1, sizeof t.rng[0]);
    memcpy(t.rng[1], rng2, sizeof t.rng[1]);
    c->mixes.push_back(t);
    return FV3D_OK;
}

int fv3d_connection_add(fv3d_ctx *c, int kind, int orient, int block1, int face1, const int *rng1, int block2,
                        int face2, const int *rng2)
{
    if (!c || !rng1 || !rng2) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "connection_add after setup");
    if (kind != 0 && kind != 1) return fail(c, FV3D_EINVAL, "connection kind %d (0 interior, 1 periodic)", kind);
    if (orient < 1 || orient > 4) return fail(c, FV3D_EINVAL, "orientation %d (1..4)", orient);
    HostConn t;
    t.kind = kind; t.orient = orient;
    const int blk[2] = {block1, block2}, fc[2] = {face1, face2};
    const int *rg[2] = {rng1, rng2};
    for (int s = 0; s < 2; ++s) {
        if (blk[s] < 1 || blk[s] > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "connection block %d", blk[s]);
        if (fc[s] < 1 || fc[s] > 6) return fail(c, FV3D_EINVAL, "connection face %d", fc[s]);
        const HostBlock &b = c->blocks[blk[s] - 1];
        int d = (fc[s] - 1) / 2, hi = (fc[s] - 1) % 2;
        int N[3] = {b.IM, b.JM, b.KM};
        if (rg[s][2 * d] != rg[s][2 * d + 1] || rg[s][2 * d] != (hi ? N[d] : 1))
            return fail(c, FV3D_EINVAL, "connection normal range does not match face %d", fc[s]);
        for (int e = 0; e < 3; ++e)
            if (e != d && (rg[s][2 * e] < 1 || rg[s][2 * e + 1] > N[e] - 1 || rg[s][2 * e] > rg[s][2 * e + 1]))
                return fail(c, FV3D_EINVAL, "connection tangential range %d..%d outside 1..%d", rg[s][2 * e],
                            rg[s][2 * e + 1], N[e] - 1);
        t.block[s] = blk[s]; t.face[s] = fc[s];
        memcpy(t.rng[s], rg[s], sizeof t.rng[s]);
    }
    c->conns.push_back(t);
    return FV3D_OK;
}

int fv3d_patch_add(fv3d_ctx *c, int kind, int block, int face, const int *rng, int rotate)
{
    if (!c || !rng) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "patch_add after setup");
    if (kind < 0 || kind > 2) return fail(c, FV3D_EINVAL, "patch kind %d not in this build", kind);
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "patch block %d", block);
    if (face < 1 || face > 6) return fail(c, FV3D_EINVAL, "patch face %d", face);
    const HostBlock &b = c->blocks[block - 1];
    int d = (face - 1) / 2, hi = (face - 1) % 2;
    int N[3] = {b.IM, b.JM, b.KM};
    if (rng[2 * d] != rng[2 * d + 1] || rng[2 * d] != (hi ? N[d] : 1))
        return fail(c, FV3D_EINVAL, "patch normal range %d..%d does not match face %d", rng[2 * d], rng[2 * d + 1], face);
    for (int e = 0; e < 3; ++e)
        if (e != d && (rng[2 * e] < 1 || rng[2 * e + 1] > N[e] - 1 || rng[2 * e] > rng[2 * e + 1]))
            return fail(c, FV3D_EINVAL, "patch tangential range %d..%d outside 1..%d", rng[2 * e], rng[2 * e + 1], N[e] - 1);
    if ((kind == 0 || kind == 1) && d != 2) return fail(c, FV3D_EINVAL, "inlet/outlet patches must be K faces");
    HostPatch p;
    p.kind = kind; p.block = block; p.face = face; p.rotate = rotate;
    memcpy(p.rng, rng, sizeof p.rng);
    c->patches.push_back(p);
    return FV3D_OK;
}

// One patch range of CUserInput::NumberOfMultigrid (source/UserInput.cpp:
// 816-844): a tangential cell range s..e becomes the node range s..e+1 (a
// single index stays), and the grid coarsens only if every start and end node
// is odd.  The block loop of the same function (1020-1040) never lowers the
// level count below two, so with the reference's cap of two levels
// (1057-1060) only the patches decide; two-sided patches are read with their
// first side (932, 962).
static bool range_coarsens(const int *r)
{
    int q[6];
    memcpy(q, r, sizeof q);
    for (int a = 0; a < 3; ++a)
        if (q[2 * a] != q[2 * a + 1]) q[2 * a + 1] += 1;
    for (int e = 0; e < 6; ++e)
        if ((q[e] - 1) % 2 != 0) return false;
    return true;
}

// Coarse patch ranges (source/UserInput.cpp:446-797): a single index n becomes
// (n+1)/2, a cell range s..e becomes (s+1)/2 .. e/2.
static void coarse_range(const int *r, int *o)
{
    for (int a = 0; a < 3; ++a) {
        const int s = r[2 * a], e = r[2 * a + 1];
        o[2 * a] = (s + 1) / 2;
        o[2 * a + 1] = s == e ? (e + 1) / 2 : e / 2;
    }
}

// The coarse level of the two-grid cycle (source/UserInput.cpp:241-334 and
// 446-797): nodes (2i-1, 2j-1, 2k-1) of every block, the same patches with
// halved ranges, no turbulence model (source/ComputeAllBlockVars.cpp:230), the
// pitch and rotation rate of the fine blocks.  It shares the fine level's
// stream and CFL state.
static int setup_coarse(fv3d_ctx *c)
{
    // the level count of each instance, as its own run would find it
    std::vector<char> inst_ok(c->ninst, 1);
    for (const HostPatch &P : c->patches)
        if (!range_coarsens(P.rng)) inst_ok[c->blocks[P.block - 1].inst] = 0;
    for (const HostConn &T : c->conns)
        if (!range_coarsens(T.rng[0])) inst_ok[c->blocks[T.block[0] - 1].inst] = 0;
    bool ok = true, any = false;
    for (char v : inst_ok) { ok = ok && v; any = any || v; }
    if (any && !ok)
        return fail(c, FV3D_EINVAL, "two-grid cycle: the instances do not all coarsen (their patch ranges differ)");
    if (!ok) {
        c->ng = 1;
        return FV3D_OK;
    }
    // coarse cell (I,J,K) covers fine cells 2I-1..2I: every block needs an even
    // number of cells along each direction (patches covering its faces ensure
    // it; a block whose faces they do not cover is rejected here, except below
    // the first coarse level, where the V-cycle just stops)
    for (size_t n = 0; n < c->blocks.size(); ++n) {
        const HostBlock &b = c->blocks[n];
        const bool tiny = (b.IM - 1) / 2 < 3 || (b.JM - 1) / 2 < 3 || (b.KM - 1) / 2 < 3;
        if ((b.IM - 1) % 2 || (b.JM - 1) % 2 || (b.KM - 1) % 2 || (tiny && c->mg_coarse)) {
            if (c->mg_coarse) {
                c->ng = 1;
                return FV3D_OK;
            }
            return fail(c, FV3D_EINVAL, "two-grid cycle: block %d has %dx%dx%d cells, not even along every direction",
                        (int)n + 1, b.IM - 1, b.JM - 1, b.KM - 1);
        }
    }
    // levels still wanted below this one: the reference's cycle stops at the
    // first coarse level; the V-cycle goes on while cfg.mg_levels allows
    const int want = c->cfg.mg_levels >= 3 ? c->cfg.mg_levels : 2;
    fv3d_ctx *m = new fv3d_ctx();
    m->cfg = c->cfg;
    m->cfg.mg_levels = want - 1;
    m->cfg.multigrid = want - 1 >= 2 ? 1 : 0;
    m->cfg.turbulence = 0;
    m->mg_fas = c->mg_fas || want >= 3;
    m->device = c->device;
    m->stream = c->stream;
    m->owns_stream = false;
    m->parent = c;
    m->mg_coarse = true;
    for (int n = 0; n < 14; ++n) (void)hipEventCreate(&m->ev[n]);
    nondimensionalize(m);
    auto bail = [&](int rc) {
        int code = fail(c, rc, "coarse level: %s", m->err.c_str());
        fv3d_ctx_destroy(m);
        return code;
    };
    int rc;
    for (const HostBlock &b : c->blocks) {
        if (c->explicit_inst && c->insts[b.inst].first_block == (int)(&b - c->blocks.data()))
            if ((rc = fv3d_instance_add(m, c->insts[b.inst].p_out)) < 0) return bail(rc);
        const int IM = (b.IM + 1) / 2, JM = (b.JM + 1) / 2, KM = (b.KM + 1) / 2;
        std::vector<double> nodes(3 * (size_t)IM * JM * KM);
        for (int i = 1; i <= IM; ++i)
            for (int j = 1; j <= JM; ++j)
                for (int k = 1; k <= KM; ++k)
                    for (int q = 0; q < 3; ++q)
                        nodes[(((size_t)(i - 1) * JM + (j - 1)) * KM + (k - 1)) * 3 + q] =
                            b.nodes[(((size_t)(2 * i - 2) * b.JM + (2 * j - 2)) * b.KM + (2 * k - 2)) * 3 + q];
        if ((rc = fv3d_block_add(m, IM, JM, KM, nodes.data(), b.nblade, b.rpm)) < 0) return bail(rc);
        m->blocks.back().w_rot = b.w_rot;   // source/UserInput.cpp:449-454
    }
    for (const HostPatch &P : c->patches) {
        int o[6];
        coarse_range(P.rng, o);
        if ((rc = fv3d_patch_add(m, P.kind, P.block, P.face, o, P.rotate))) return bail(rc);
    }
    for (const HostConn &T : c->conns) {
        int o1[6], o2[6];
        coarse_range(T.rng[0], o1);
        coarse_range(T.rng[1], o2);
        if ((rc = fv3d_connection_add(m, T.kind, T.orient, T.block[0], T.face[0], o1, T.block[1], T.face[1], o2)))
            return bail(rc);
    }
    for (const HostConn &T : c->mixes) {
        int o1[6], o2[6];
        coarse_range(T.rng[0], o1);
        coarse_range(T.rng[1], o2);
        if ((rc = fv3d_mixing_plane_add(m, T.block[0], o1, T.block[1], o2))) return bail(rc);
    }
    if ((rc = fv3d_setup(m))) return bail(rc);
    // a small coarse level spreads its convective residual one face per thread
    if (m->ncell <= kFaceFormCells) {
        if ((rc = ensure_face_form(m))) return bail(rc);
        m->tile_sel[0] = kConvFaces;
    }
    std::vector<int4> tl;
    for (int n = 0; n < (int)m->blocks.size(); ++n) {
        const BlockDev &B = m->blocks[n].dev;
        for (int I = 1; I <= B.ni; I += kMgTI)
            for (int J = 1; J <= B.nj; J += kMgTJ)
                for (int K = 1; K <= B.nk; K += kMgTK) tl.push_back(make_int4(n, I, J, K));
    }
    if ((rc = dalloc(c, &c->d_mg_tiles, tl.size()))) {
        fv3d_ctx_destroy(m);
        return rc;
    }
    HIPCK(c, hipMemcpyAsync(c->d_mg_tiles, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    c->n_mg_tiles = (int)tl.size();
    c->coarse = m;
    c->ng = 2;
    c->nlev = 1 + m->nlev;
    return FV3D_OK;
}

int fv3d_setup(fv3d_ctx *c)
{
    if (!c) return FV3D_EINVAL;
    if (c->ready) return fail(c, FV3D_ESTATE, "setup called twice");
    if (c->blocks.empty()) return fail(c, FV3D_ESTATE, "no blocks");
    int nb = (int)c->blocks.size();
    int rc;
    if (!c->explicit_inst) c->insts.assign(1, HostInst{c->cfg.p_out, 0, nb});
    c->ninst = (int)c->insts.size();
    for (const HostInst &I : c->insts)
        if (I.nblock == 0) return fail(c, FV3D_EINVAL, "an instance without blocks");
    for (const HostConn &T : c->conns)
        if (c->blocks[T.block[0] - 1].inst != c->blocks[T.block[1] - 1].inst)
            return fail(c, FV3D_EINVAL, "a connection between blocks %d and %d of different instances", T.block[0],
                        T.block[1]);
    if ((rc = dalloc(c, &c->d_flags, nb))) return rc;
    if ((rc = dalloc(c, &c->d_bad, nb))) return rc;
    c->total_vol = 0.0;
    c->inst_vol.assign(c->ninst, 0.0);
    for (int n = 0; n < nb; ++n) {
        if ((rc = setup_block(c, c->blocks[n], n))) return rc;
        // source/ComputeAllBlockVars.cpp:25-36, per instance
        c->inst_vol[c->blocks[n].inst] = c->inst_vol[c->blocks[n].inst] + c->blocks[n].cellvol;
    }
    c->total_vol = c->inst_vol[0];
    {
        // mailboxes (forward and backward sets per block) and the pack work list
        long words = 0;
        std::vector<long> gst{0};
        for (auto &b : c->blocks) {
            words += 2 * b.sweep_rows * kMboxRow;
            gst.push_back(gst.back() + b.pack_groups);
        }
        if ((rc = dalloc(c, &c->d_mbox, words))) return rc;
        c->mbox_words = words;
        long off = 0;
        for (auto &b : c->blocks) {
            b.dev.mbox[0] = c->d_mbox + off;
            b.dev.mbox[1] = b.dev.mbox[0] + b.sweep_rows * kMboxRow;
            off += 2 * b.sweep_rows * kMboxRow;
        }
        c->pack_groups = gst.back();
        if ((rc = dalloc(c, &c->d_pack_start, gst.size()))) return rc;
        HIPCK(c, hipMemcpyAsync(c->d_pack_start, gst.data(), gst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    }
    // per-instance scalars (filled below); the CFL number each block reads is
    // its instance's, kept by the fine level
    if ((rc = dalloc(c, &c->d_sc, c->ninst))) return rc;
    c->ramp_sc = c->parent ? c->parent->ramp_sc : c->d_sc;
    for (auto &b : c->blocks) b.dev.cfl = &c->ramp_sc[b.inst].cfl;
    std::vector<BlockDev> bd;
    for (auto &b : c->blocks) bd.push_back(b.dev);
    if ((rc = dalloc(c, &c->d_blocks, nb))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_blocks, bd.data(), nb * sizeof(BlockDev), hipMemcpyHostToDevice, c->stream));
    for (int n = 0; n < nb; ++n) {
        const BlockDev &B = c->blocks[n].dev;
        if (c->cfg.viscous) {
            // coordinate-transformation coefficients (source/ComputeAllBlockVars.cpp:34-35)
            hipLaunchKernelGGL(k_metrics, dim3((unsigned)((4 * B.box + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                               n, (double *)B.fg[0], (double *)B.fg[1], (double *)B.fg[2], (double *)B.cg);
            HIPCK(c, hipGetLastError());
        }
        const long rows = c->blocks[n].sweep_rows;
        if (c->cfg.precision == FV3D_DEFECT_FP32)
            hipLaunchKernelGGL(k_sweep_fill<float>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, c->stream, c->d_blocks,
                               n, rows, kGatherState | kGatherMetrics);
        else
            hipLaunchKernelGGL(k_sweep_fill<double>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, c->stream, c->d_blocks,
                               n, rows, kGatherState | kGatherMetrics);
        HIPCK(c, hipGetLastError());
    }
    // patch work list
    std::vector<PatchDev> pd;
    std::vector<long> pst{0};
    // inlet and outlet sections of each instance: its first inlet and first
    // outlet patch (source/Ouput.cpp:9-107)
    std::vector<int> nin(c->ninst, 0), nout(c->ninst, 0);
    std::vector<int2> secs(2 * c->ninst, make_int2(0, 0));
    for (const HostPatch &P : c->patches) {
        PatchDev q;
        q.kind = P.kind; q.block = P.block - 1;
        q.inst = c->blocks[P.block - 1].inst;
        q.dir = (P.face - 1) / 2; q.hi = (P.face - 1) % 2;
        int d = q.dir;
        const int *r = P.rng;
        q.a0 = d == 0 ? r[2] : r[0];
        int aE = d == 0 ? r[3] : r[1];
        q.b0 = d == 2 ? r[2] : r[4];
        int bE = d == 2 ? r[3] : r[5];
        q.na = aE - q.a0 + 1; q.nb = bE - q.b0 + 1;
        q.centrifugal = c->cfg.turbo_type == 1;
        q.axial = P.kind == 1 && c->cfg.turbo_type == 0;
        q.rot = P.rotate ? c->blocks[P.block - 1].w_rot : 0.0;
        pd.push_back(q);
        pst.push_back(pst.back() + (long)q.na * q.nb);
        if (P.kind == 0 && nin[q.inst]++ == 0) secs[2 * q.inst] = make_int2(P.block - 1, r[4]);
        if (P.kind == 1 && nout[q.inst]++ == 0) secs[2 * q.inst + 1] = make_int2(P.block - 1, r[4]);
    }
    for (int w = 0; w < c->ninst; ++w)
        if (nin[w] == 0 || nout[w] == 0)
            return fail(c, FV3D_EINVAL, c->ninst > 1 ? "instance %d needs an inlet and an outlet patch"
                                                     : "need an inlet and an outlet patch", w + 1);
    {
        // radial-equilibrium outlets: each covers its block's whole K+ face
        std::vector<int> ax;
        for (int q = 0; q < (int)pd.size(); ++q) {
            if (!pd[q].axial || pd[q].dir != 2 || !pd[q].hi) continue;
            for (int r : ax)
                if (pd[r].block == pd[q].block)
                    return fail(c, FV3D_EINVAL, "two radial-equilibrium outlet patches on block %d", pd[q].block + 1);
            if (c->blocks[pd[q].block].IM - 1 > kAxialMaxI)
                return fail(c, FV3D_EINVAL, "radial-equilibrium outlet: more than %d cells along i", kAxialMaxI);
            ax.push_back(q);
        }
        if (!ax.empty()) {
            if ((rc = dalloc(c, &c->d_axial, ax.size()))) return rc;
            HIPCK(c, hipMemcpyAsync(c->d_axial, ax.data(), ax.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
            c->n_axial = (int)ax.size();
        }
        if (c->cfg.global_dt && (rc = dalloc(c, &c->d_dtmin, nb))) return rc;
    }
    if (c->cfg.turbulence == 100)
        for (const HostPatch &P : c->patches) {
            if (P.kind != 2) continue;
            const HostBlock &hb = c->blocks[P.block - 1];
            const int d = (P.face - 1) / 2;
            // the reference's 1-D model runs over JM cells for K-face walls
            // (source/BL_Model.cpp:225), which is defined only when JM == KM
            if (d == 2 && hb.JM != hb.KM)
                return fail(c, FV3D_EINVAL, "Baldwin-Lomax on a K-face wall needs JM == KM (block %d)", P.block);
        }
    if (c->cfg.turbulence == 100) {
        std::vector<int2> wl;
        for (int G = 16; G <= 64; G *= 2) {
            long lines = 0;
            int n = 0, first = (int)wl.size();
            for (int q = 0; q < (int)pd.size(); ++q) {
                if (c->patches[q].kind != 2) continue;
                const PatchDev &P = pd[q];
                const HostBlock &hb = c->blocks[P.block];
                const int len = (P.dir == 0 ? hb.IM : P.dir == 1 ? hb.JM : hb.KM) - 1;
                int need = 16;
                while (need * kBlCpt < len && need < 64) need *= 2;
                if (need * kBlCpt < len)
                    return fail(c, FV3D_EINVAL, "Baldwin-Lomax lines are limited to %d cells", 64 * kBlCpt);
                if (need != G) continue;
                // lines rounded up to whole workgroups, so a workgroup never spans two walls
                const long nl = ((long)P.na * P.nb + kBlLines - 1) / kBlLines * kBlLines;
                wl.push_back(make_int2(q, (int)nl));
                lines += nl;
                ++n;
            }
            if (!n) continue;
            c->bl_launch.push_back({first, lines});
            c->bl_launch_n.push_back(n);
            c->bl_groups.push_back(G);
        }
        if (!wl.empty()) {
            if ((rc = dalloc(c, &c->d_bl_walls, wl.size()))) return rc;
            HIPCK(c, hipMemcpyAsync(c->d_bl_walls, wl.data(), wl.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
        }
    }
    c->npatch_dev = (int)pd.size();
    c->pd_host = pd;
    if ((rc = build_exchange(c))) return rc;
    c->nbc = pst.back();
    if ((rc = dalloc(c, &c->d_patches, pd.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_patches, pd.data(), pd.size() * sizeof(PatchDev), hipMemcpyHostToDevice, c->stream));
    if ((rc = dalloc(c, &c->d_pstart_bc, pst.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_pstart_bc, pst.data(), pst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    if ((rc = dalloc(c, &c->d_secs, secs.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_secs, secs.data(), secs.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    // cell and plane work lists
    std::vector<long> cst{0}, plst{0};
    c->nplanes = 0;
    c->parts_host.assign(c->ninst, make_int2(0, 0));
    c->ncell_real = 0;
    for (int n = 0; n < nb; ++n) {
        const BlockDev &B = c->blocks[n].dev;
        const int w = c->blocks[n].inst;
        if (n > 0 && c->insts[w].first_block == n) {
            // an instance starts on a whole 256-cell group, so its norm partials
            // are those it has when it runs alone: pad the previous block's range
            cst.back() = (cst.back() + 255) / 256 * 256;
        }
        if (c->insts[w].first_block == n) c->parts_host[w].x = (int)(cst.back() / 256);
        cst.push_back(cst.back() + (long)B.ni * B.nj * B.nk);
        c->ncell_real += (long)B.ni * B.nj * B.nk;
        if (n == c->insts[w].first_block + c->insts[w].nblock - 1)
            c->parts_host[w].y = (int)((cst.back() + 255) / 256) - c->parts_host[w].x;
        plst.push_back(plst.back() + (long)B.ni * B.nj);
        c->nplanes = std::max(c->nplanes, B.IM + B.JM + B.KM - 5);
    }
    c->ncell = cst.back();
    c->nplane_threads = plst.back();
    if ((rc = dalloc(c, &c->d_cstart, cst.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_cstart, cst.data(), cst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    if ((rc = dalloc(c, &c->d_pstart_plane, plst.size()))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_pstart_plane, plst.data(), plst.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    if ((rc = ensure_tiles(c, 0))) return rc;
    // wavefront tiles: forward in increasing, backward in decreasing tile
    // diagonal, so every upstream tile holds a smaller ticket
    {
        struct T { int n, J, K, diag; };
        std::vector<T> lst;
        std::vector<int> nTJ(nb), nTK(nb);
        for (int n = 0; n < nb; ++n) {
            const BlockDev &B = c->blocks[n].dev;
            nTJ[n] = (B.nj + kKbaTJ - 1) / kKbaTJ;
            nTK[n] = (B.nk + kKbaTK - 1) / kKbaTK;
            for (int J = 0; J < nTJ[n]; ++J)
                for (int K = 0; K < nTK[n]; ++K) lst.push_back({n, J, K, J + K});
        }
        for (int dir = 0; dir < 2; ++dir) {
            std::vector<T> o = lst;
            if (dir == 1)
                for (T &t : o) t.diag = (nTJ[t.n] - 1 - t.J) + (nTK[t.n] - 1 - t.K);
            std::stable_sort(o.begin(), o.end(), [](const T &a, const T &b) { return a.diag < b.diag; });
            std::vector<std::vector<int>> tix(nb);
            for (int n = 0; n < nb; ++n) tix[n].assign((size_t)nTJ[n] * nTK[n], -1);
            for (size_t q = 0; q < o.size(); ++q) tix[o[q].n][(size_t)o[q].J * nTK[o[q].n] + o[q].K] = (int)q;
            std::vector<KbaTile> kt;
            for (const T &t : o) {
                int dj = dir == 0 ? t.J - 1 : t.J + 1, dk = dir == 0 ? t.K - 1 : t.K + 1;
                KbaTile x;
                x.block = t.n; x.j0 = t.J * kKbaTJ; x.k0 = t.K * kKbaTK;
                x.dep_j = (dj >= 0 && dj < nTJ[t.n]) ? tix[t.n][(size_t)dj * nTK[t.n] + t.K] : -1;
                x.dep_k = (dk >= 0 && dk < nTK[t.n]) ? tix[t.n][(size_t)t.J * nTK[t.n] + dk] : -1;
                kt.push_back(x);
            }
            if ((rc = dalloc(c, &c->d_kba[dir], kt.size()))) return rc;
            HIPCK(c, hipMemcpyAsync(c->d_kba[dir], kt.data(), kt.size() * sizeof(KbaTile), hipMemcpyHostToDevice, c->stream));
        }
        c->nkba = (int)lst.size();
        if ((rc = dalloc(c, &c->d_kba_sync, 64 + 2 * (size_t)c->nkba * kDonePad))) return rc;
        c->kba_grid = std::min(c->nkba, 4096);
    }
    c->npart = (int)((c->ncell + 255) / 256);
    if ((rc = dalloc(c, &c->d_partial, 6 * (size_t)c->npart))) return rc;
    if ((rc = dalloc(c, &c->d_parts, c->ninst))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_parts, c->parts_host.data(), c->ninst * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    if (c->cfg.turbulence == 200 && (rc = setup_sa(c))) return rc;
    if ((rc = dalloc(c, &c->d_sums, 2 * 3 * kSecParts * (size_t)c->ninst))) return rc;
    if ((rc = dalloc(c, &c->d_stage_row, 7 * (size_t)c->ninst))) return rc;
    // device scalars of each instance
    std::vector<ScalarsDev> scs(c->ninst);
    for (int w = 0; w < c->ninst; ++w) {
        ScalarsDev &sc = scs[w];
        memset(&sc, 0, sizeof sc);
        sc.cfl = (float)c->cfg.cfl;
        sc.cfl_res = F32C(1.0E-3);   // source/Solver.cpp:21
        sc.total_vol = c->inst_vol[w];
        sc.dens_ref = c->dens_ref; sc.vel_ref = c->vel_ref; sc.temp_ref = c->temp_ref;
        sc.nblade_in = c->blocks[secs[2 * w].x].nblade;
        sc.nblade_out = c->blocks[secs[2 * w + 1].x].nblade;
        sc.p_out = inst_p_out(c, w);
    }
    HIPCK(c, hipMemcpyAsync(c->d_sc, scs.data(), scs.size() * sizeof(ScalarsDev), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->ready = true;
    if (c->cfg.multigrid && (rc = setup_coarse(c))) {
        c->ready = false;
        return rc;
    }
    return FV3D_OK;
}

int fv3d_state_upload(fv3d_ctx *c, int block, const double *pv, const double *ener)
{
    if (!c || !pv || !ener) return FV3D_EINVAL;
    if (!c->ready) return fail(c, FV3D_ESTATE, "state_upload before setup");
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "block %d", block);
    const BlockDev &B = c->blocks[block - 1].dev;
    std::vector<double> Q(5 * B.box, 0.0), W(6 * B.box, 0.0);
    HIPCK(c, hipMemcpy(Q.data(), B.Q, Q.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCK(c, hipMemcpy(W.data(), B.W, W.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 1; i <= B.ni; ++i)
        for (int j = 1; j <= B.nj; ++j)
            for (int k = 1; k <= B.nk; ++k) {
                long s = ((long)(i - 1) * B.nj + (j - 1)) * B.nk + (k - 1);
                long x = cidx(B, i, j, k);
                const double *w = pv + 6 * s;
                // source/UserInput.cpp:1139-1150
                Q[x] = w[0]; Q[B.box + x] = w[0] * w[1]; Q[2 * B.box + x] = w[0] * w[2]; Q[3 * B.box + x] = w[0] * w[3];
                Q[4 * B.box + x] = ener[s];
                for (int q = 0; q < 6; ++q) W[q * B.box + x] = w[q];
            }
    HIPCK(c, hipMemcpyAsync(B.Q, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(B.W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return FV3D_OK;
}

// nu~ of one block's cells from a restart file (CMesh::ReadRestart's nut
// field, source/UserInput.cpp:1134-1153)
int fv3d_nut_upload(fv3d_ctx *c, int block, const double *ev)
{
    if (!c || !ev) return FV3D_EINVAL;
    if (!c->ready) return fail(c, FV3D_ESTATE, "nut_upload before setup");
    if (c->cfg.turbulence != 200) return fail(c, FV3D_EINVAL, "nu~ exists with the Spalart-Allmaras model only");
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "block %d", block);
    const BlockDev &B = c->blocks[block - 1].dev;
    std::vector<double> E(B.box, 0.0);
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(E.data(), B.ev, E.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 1; i <= B.ni; ++i)
        for (int j = 1; j <= B.nj; ++j)
            for (int k = 1; k <= B.nk; ++k) E[cidx(B, i, j, k)] = ev[((long)(i - 1) * B.nj + (j - 1)) * B.nk + (k - 1)];
    HIPCK(c, hipMemcpy(B.ev, E.data(), E.size() * sizeof(double), hipMemcpyHostToDevice));
    c->ev_set = true;
    return FV3D_OK;
}

int fv3d_levels(const fv3d_ctx *c) { return c ? c->nlev : FV3D_EINVAL; }

// CMesh::Initial (source/ComputeAllBlockVars.cpp:38-48) on one level: the
// uniform field of CBlocks::Initial_Uniform (source/Initialization.cpp:17-72)
// or CMesh::Initial_Interpolate (74-165), pressure and flow direction linear in
// z between the first inlet's K face and the first outlet's.  Host code: it
// runs once, before the first iteration.
static int initial_instance(fv3d_ctx *m, int type, int w)
{
    const Gas &g = m->gas;
    const double Gamma = g.gamma, Rgas = g.rgas, Ptin = g.pt_in, Ttin = g.tt_in, Pout = inst_p_out(m, w);
    const double *din = m->cfg.inlet_dir, *dout = m->cfg.outlet_dir;
    const int nb = (int)m->blocks.size();
    // cell centroids (and K-face centroids) in the reference's operation order
    std::vector<std::vector<double>> cc(nb), fck(nb);
    for (int n = 0; n < nb; ++n) {
        if (m->blocks[n].inst != w) continue;
        std::vector<double> fS[3], fs[3], fc[3], vol, node;
        face_vector_volume(m->blocks[n], fS[0], fS[1], fS[2], fs[0], fs[1], fs[2], fc[0], fc[1], fc[2], vol, cc[n], node);
        fck[n] = std::move(fc[2]);
    }
    double press0 = 0, temp = 0, dens0 = 0, Vel = 0, n_m0 = 0, zmin = 0, dz = 1, Pin = 0, dp = 0, dB[3] = {0, 0, 0};
    if (type == 0) {
        const double Ma0 = (float)m->cfg.ma0;
        const double m2 = F32C(1.0) + (Gamma - F32C(1.0)) / F32C(2.0) * Ma0 * Ma0;
        press0 = Ptin * pow(m2, F32C(0.0) - Gamma / (Gamma - F32C(1.0)));
        temp = Ttin * pow(m2, (double)(0.0f - 1.0f));
        dens0 = press0 / (Rgas * temp);
        Vel = pow(Gamma * Rgas * temp, F32C(0.5)) * Ma0;
        n_m0 = pow(din[0] * din[0] + din[1] * din[1] + din[2] * din[2], 0.5);
    } else {
        const HostPatch *pin = nullptr, *pout = nullptr;
        for (const HostPatch &P : m->patches) {
            if (m->blocks[P.block - 1].inst != w) continue;
            if (P.kind == 0 && !pin) pin = &P;
            if (P.kind == 1 && !pout) pout = &P;
        }
        if (!pin || !pout) return fail(m, FV3D_EINVAL, "the interpolated start needs an inlet and an outlet patch");
        zmin = F32C(1E+32);
        double zmax = F32C(-1E+32);
        const HostBlock &bi = m->blocks[pin->block - 1], &bo = m->blocks[pout->block - 1];
        auto fz = [&](const HostBlock &b, const std::vector<double> &f, int i, int j, int k) {
            const long box = (long)(b.IM + 1) * (b.JM + 1) * (b.KM + 1);
            return f[2 * box + ((long)i * (b.JM + 1) + j) * (b.KM + 1) + k];
        };
        for (int i = 1; i <= bi.IM - 1; ++i)
            for (int j = 1; j <= bi.JM - 1; ++j) {
                const double z = fz(bi, fck[pin->block - 1], i, j, pin->rng[4]);
                zmin = z < zmin ? z : zmin;
            }
        for (int i = 1; i <= bo.IM - 1; ++i)
            for (int j = 1; j <= bo.JM - 1; ++j) {
                const double z = fz(bo, fck[pout->block - 1], i, j, pout->rng[5]);
                zmax = z > zmax ? z : zmax;
            }
        dz = zmax - zmin;
        Pin = Pout / (double)(float)m->cfg.p21ratio;
        if (Pin >= Ptin) Pin = F32C(0.99) * Ptin;
        dp = Pout - Pin;
        for (int a = 0; a < 3; ++a) dB[a] = dout[a] - din[a];
        temp = Ttin * pow(Pin / Ptin, (Gamma - F32C(1.0)) / Gamma);
        dens0 = Pin / (Rgas * temp);
        const double cs = sqrt(Gamma * Pin / dens0);
        const double Ma = sqrt(F32C(2.0) * ((Ttin / temp) - F32C(1.0)) / (Gamma - F32C(1.0)));
        Vel = Ma * cs;
    }
    for (int n = 0; n < nb; ++n) {
        if (m->blocks[n].inst != w) continue;
        const BlockDev &B = m->blocks[n].dev;
        std::vector<double> Q(5 * B.box, 0.0), W(6 * B.box, 0.0);
        HIPCK(m, hipMemcpy(Q.data(), B.Q, Q.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCK(m, hipMemcpy(W.data(), B.W, W.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 1; i <= B.ni; ++i)
            for (int j = 1; j <= B.nj; ++j)
                for (int k = 1; k <= B.nk; ++k) {
                    const long x = cidx(B, i, j, k);
                    double w[6], ener;
                    if (type == 0) {
                        const double u = Vel * din[0] / n_m0, v = Vel * din[1] / n_m0, ww = Vel * din[2] / n_m0;
                        w[0] = dens0; w[1] = u; w[2] = v; w[3] = ww; w[4] = press0; w[5] = temp;
                        const double V2 = u * u + v * v + ww * ww;
                        ener = dens0 * (press0 / (dens0 * (Gamma - F32C(1.0))) + V2 / F32C(2.0));
                    } else {
                        const double z = cc[n][2 * B.box + x];
                        double Bt[3];
                        for (int a = 0; a < 3; ++a) Bt[a] = din[a] + dB[a] * (z - zmin) / dz;
                        const double press = Pin + dp * (z - zmin) / dz;
                        const double dens = press / Rgas / temp;
                        const double n_m = sqrt(Bt[0] * Bt[0] + Bt[1] * Bt[1] + Bt[2] * Bt[2]);
                        const double u = Vel * Bt[0] / n_m, v = Vel * Bt[1] / n_m, ww = Vel * Bt[2] / n_m;
                        w[0] = dens; w[1] = u; w[2] = v; w[3] = ww; w[4] = press; w[5] = temp;
                        const double V2 = u * u + v * v + ww * ww;
                        ener = press / (Gamma - F32C(1.0)) + dens * V2 / F32C(2.0);
                    }
                    Q[x] = w[0]; Q[B.box + x] = w[0] * w[1]; Q[2 * B.box + x] = w[0] * w[2];
                    Q[3 * B.box + x] = w[0] * w[3]; Q[4 * B.box + x] = ener;
                    for (int q = 0; q < 6; ++q) W[q * B.box + x] = w[q];
                }
        HIPCK(m, hipMemcpy(B.Q, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCK(m, hipMemcpy(B.W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
        if (B.ev) {
            // nu~ = ev_ratio in the interior (source/Initialization.cpp:40, 71, 130, 166)
            std::vector<double> E(B.box, 0.0);
            HIPCK(m, hipMemcpy(E.data(), B.ev, E.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 1; i <= B.ni; ++i)
                for (int j = 1; j <= B.nj; ++j)
                    for (int k = 1; k <= B.nk; ++k) E[cidx(B, i, j, k)] = g.ev_ratio * F32C(1.0);
            HIPCK(m, hipMemcpy(B.ev, E.data(), E.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    return FV3D_OK;
}

// CMesh::Initial on one level, each instance with its own outlet pressure and
// its own first inlet and outlet
static int initial_level(fv3d_ctx *m, int type)
{
    for (int w = 0; w < m->ninst; ++w) {
        int rc = initial_instance(m, type, w);
        if (rc) return rc;
    }
    if (m->cfg.turbulence == 200) m->ev_set = true;
    return FV3D_OK;
}

int fv3d_initial(fv3d_ctx *c)
{
    if (!c) return FV3D_EINVAL;
    if (!c->ready) return fail(c, FV3D_ESTATE, "initial before setup");
    HIPCK(c, hipStreamSynchronize(c->stream));
    fv3d_ctx *m = c->ng == 2 ? c->coarse : c;
    int rc = initial_level(m, c->cfg.initial_type);
    if (rc) {
        if (m != c) fail(c, rc, "coarse level: %s", m->err.c_str());
        return rc;
    }
    // with Spalart-Allmaras the fine level starts from its own field too
    // (source/Structured_FV3D.cpp:89-95); the state interpolated at
    // Coarse_iter replaces its flow variables, not its nu~
    if (m != c && c->cfg.turbulence == 200 && (rc = initial_level(c, c->cfg.initial_type))) return rc;
    c->initial_done = true;
    return FV3D_OK;
}

// ------------------------------------------------------------- stages
static int launch_bc(fv3d_ctx *c)
{
    // the periodic signs are updated by k_bc's first thread when it runs
    if (c->nbc > 0)
        hipLaunchKernelGGL(k_bc, dim3((unsigned)((c->nbc + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks, c->d_patches,
                           c->npatch_dev, c->d_pstart_bc, c->gas, c->d_sc, c->d_per, c->nper, c->d_pitch_sign);
    else if (c->nper > 0)
        hipLaunchKernelGGL(k_periodic_flip, dim3(1), dim3(64), 0, c->stream, c->d_per, c->nper, c->d_pitch_sign);
    if (c->n_axial > 0)
        hipLaunchKernelGGL(k_bc_axial, dim3((unsigned)c->n_axial), dim3(256), 0, c->stream, c->d_blocks, c->d_patches,
                           c->d_axial, c->gas, c->d_sc);
    if (c->nexch > 0)
        hipLaunchKernelGGL(k_exchange, dim3((unsigned)((c->nexch + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_exch, c->nexch, c->d_per);
    if (c->nmix_lines > 0)   // after interior and periodic patches (source/BoundaryCondition.cpp:59-62)
        hipLaunchKernelGGL(k_mixing_plane, dim3((unsigned)c->nmix_lines), dim3(256), 0, c->stream, c->d_blocks, c->d_mix,
                           c->nmix, c->gas, c->cfg.turbulence == 200 ? 1 : 0);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// CFluidProps::SpectralRadii and CTimeDiscr::TimeStep (fused): lambda_c,
// lambda_v, dt and the LU-SGS diagonals; with decrement the Flag_low_dt
// countdown at the end of TimeStep (source/TimeStep.cpp:89-92) runs too.
static int launch_spectral_dt(fv3d_ctx *c, bool decrement = true)
{
    hipLaunchKernelGGL(k_spectral_dt, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                       c->d_cstart, (int)c->blocks.size(), c->gas);
    HIPCK(c, hipGetLastError());
    if (c->cfg.global_dt) {
        const int nb = (int)c->blocks.size();
        const unsigned g = (unsigned)((c->ncell + 255) / 256);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, c->d_dtmin, (long)nb,
                           (unsigned long long)0x4202A05F20000000ull);   // 1E10
        hipLaunchKernelGGL(k_dt_block_min, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb, c->d_dtmin);
        hipLaunchKernelGGL(k_dt_global, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb, c->gas, c->d_dtmin);
        HIPCK(c, hipGetLastError());
    }
    if (!decrement) return FV3D_OK;
    hipLaunchKernelGGL(k_dt_flag_decrement, dim3(1), dim3(256), 0, c->stream, c->d_blocks, (int)c->blocks.size());
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// The TimeStep of a residual evaluation that only a restriction reads, in the
// FP32 defect-correction form: lambda, dt and the LU-SGS diagonals it would form
// are formed again by the level's next evaluation before anything reads them
// (the restriction takes Q, R and the volumes), so only the Flag_low_dt
// countdown at its end runs (source/TimeStep.cpp:89-92).  The plain FP64 path
// forms them, as the reference's ResFlux does, so that they can be read after
// an iteration.
static int launch_restriction_timestep(fv3d_ctx *c)
{
    if (c->cfg.precision != FV3D_DEFECT_FP32) return launch_spectral_dt(c);
    hipLaunchKernelGGL(k_dt_flag_decrement, dim3(1), dim3(256), 0, c->stream, c->d_blocks, (int)c->blocks.size());
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

static int launch_convective(fv3d_ctx *c)
{
    if (!c->store_faces && c->tile_sel[0] == kConvFaces) {
        if (!c->d_fstart) return fail(c, FV3D_ESTATE, "face-parallel convective form not built");
        const int nb = (int)c->blocks.size();
        hipLaunchKernelGGL(k_conv_faces, dim3((unsigned)((c->nfaces + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_fstart, 3 * nb, c->gas);
        hipLaunchKernelGGL(k_conv_sum, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_cstart, nb);
        HIPCK(c, hipGetLastError());
        return FV3D_OK;
    }
    if (!c->store_faces && c->tile_sel[0] >= kNumTiles) {
        const int q = c->tile_sel[0] - kNumTiles;
        int rc = ensure_stream_tiles(c, q);
        if (rc) return rc;
        const int nt = c->nstiles[q], ci = kStreamShape[q][2];
        const int mode = c->cfg.recon_var == 200 ? 0 : c->cfg.recon_var == 100 ? 1 : 2;
        const dim3 g((unsigned)(8 * ((nt + 7) / 8)));
        switch (q < 2 || q == 4 ? 0 : 1) {
#define FV3D_STREAM(S, TJ, TK)                                                                                     \
    case S:                                                                                                         \
        if (mode == 0)                                                                                              \
            hipLaunchKernelGGL((k_conv_stream<TJ, TK, 0>), g, dim3(TJ * TK), 0, c->stream, c->d_blocks,             \
                               c->d_stiles[q], nt, ci, c->gas);                                                     \
        else if (mode == 1)                                                                                         \
            hipLaunchKernelGGL((k_conv_stream<TJ, TK, 1>), g, dim3(TJ * TK), 0, c->stream, c->d_blocks,             \
                               c->d_stiles[q], nt, ci, c->gas);                                                     \
        else                                                                                                        \
            hipLaunchKernelGGL((k_conv_stream<TJ, TK, 2>), g, dim3(TJ * TK), 0, c->stream, c->d_blocks,             \
                               c->d_stiles[q], nt, ci, c->gas);                                                     \
        break;
            FV3D_FOR_STREAM_TILES(FV3D_STREAM)
#undef FV3D_STREAM
        }
        HIPCK(c, hipGetLastError());
        return FV3D_OK;
    }
    const int q = c->tile_sel[0] < kNumTiles ? c->tile_sel[0] : 0;
    if (!c->d_tiles_s[q]) return fail(c, FV3D_ESTATE, "convective tiles of shape %d not built", q);
    const dim3 g((unsigned)c->ntiles_s[q]), b(256);
    switch (q) {
#define FV3D_CONV(S, A, B, C)                                                                                    \
    case S:                                                                                                       \
        hipLaunchKernelGGL((k_convective<A, B, C>), g, b, 0, c->stream, c->d_blocks, c->d_tiles_s[q], c->gas,       \
                           c->store_faces, c->d_faces[0], c->d_faces[1], c->d_faces[2]);                         \
        break;
        FV3D_FOR_TILES(FV3D_CONV)
#undef FV3D_CONV
    }
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// Viscous flux subtracted from the residual (CMesh::ViscousFlux,
// source/ComputeAllBlockVars.cpp:131-160, and Res = Fc - Fv - Fd, 245).
static int launch_viscous(fv3d_ctx *c)
{
    if (!c->cfg.viscous) return FV3D_OK;
    const int q = c->tile_sel[1];
    const dim3 g((unsigned)c->ntiles_s[q]), b(256);
    switch (q) {
#define FV3D_VISC(S, A, B, C)                                                                                    \
    case S:                                                                                                       \
        hipLaunchKernelGGL((k_viscous<A, B, C>), g, b, 0, c->stream, c->d_blocks, c->d_tiles_s[q], c->gas,          \
                           c->store_faces, c->vdbg);                                                             \
        break;
        FV3D_FOR_TILES(FV3D_VISC)
#undef FV3D_VISC
    }
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

static void launch_sa_faces(fv3d_ctx *c)
{
    const int q = c->tile_sel[2];
    const dim3 g((unsigned)c->ntiles_s[q]), b(256);
    switch (q) {
#define FV3D_SAF(S, A, B, C)                                                                                     \
    case S:                                                                                                       \
        hipLaunchKernelGGL((k_sa_faces<A, B, C>), g, b, 0, c->stream, c->d_blocks, c->d_tiles_s[q], c->gas);       \
        break;
        FV3D_FOR_TILES(FV3D_SAF)
#undef FV3D_SAF
    }
}

// Spalart-Allmaras in the residual evaluation: the vorticity, the nu~
// residual (face gradients from nu~ as it stands, fluxes and source from its
// clipped value), then mue_t and the clipping stored, then mue_t of the ghost
// cells.  save_old: the LU-SGS step keeps nu~ before the clipping in ev_old.
static int launch_sa(fv3d_ctx *c, int save_old, bool mut_only)
{
    const int nb = (int)c->blocks.size();
    const unsigned gc = (unsigned)((c->ncell + 255) / 256);
    if (!mut_only) {
        hipLaunchKernelGGL(k_curl, dim3(gc), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb);
        launch_sa_faces(c);
    }
    hipLaunchKernelGGL(k_sa_mut, dim3(gc), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb, c->gas, save_old);
    hipLaunchKernelGGL(k_sa_ghost, dim3((unsigned)((c->nghost + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                       c->d_gstart, nb, c->gas);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// Turbulence model before the time step (source/ComputeAllBlockVars.cpp:
// 228-232): the vorticity, then Baldwin-Lomax on all walls at once (one launch
// per line group size) and the merge of the per-cell minima; or
// Spalart-Allmaras.  mut_only (the cycle's second fine residual, the one the
// restriction takes): Spalart-Allmaras forms mue_t (with the clipping of nu~
// and mue_t and the ghost values) but not the nu~ residual and the vorticity
// it needs, which nothing reads before the next iteration's first residual
// forms them again; the learned closure keeps the mue_t of the iteration's
// first evaluation (it has no reference counterpart; lagging the coarse-grid
// correction's mue_t by half an iteration leaves the fixed point unchanged).
static int launch_turbulence(fv3d_ctx *c, int save_old = 0, bool mut_only = false)
{
    if (c->cfg.turbulence == 200 && c->mlp_closure && mut_only) return FV3D_OK;
    if (c->cfg.turbulence == 200 && c->mlp_closure) {
        // the learned closure in place of the nu~ transport; ghost mue_t as for SA
        const int nb = (int)c->blocks.size();
        const long chunks = (c->ncell + 255) / 256;   // network: 768 workgroups (three per CU), several chunks each
        hipLaunchKernelGGL(k_mlp_cell_features, dim3((unsigned)chunks), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_cstart, nb, (long)c->ncell, c->gas, c->d_mlp_bounds, c->d_mlp_xc);
        hipLaunchKernelGGL(k_mlp_closure, dim3((unsigned)(chunks < 768 ? chunks : 768)), dim3(256), 0, c->stream,
                           c->d_mlp, (long)c->ncell, c->d_mlp_xc);
        hipLaunchKernelGGL(k_mlp_mut, dim3((unsigned)chunks), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb,
                           c->gas, c->d_mlp_bounds, c->d_mlp_xc, c->d_mlp_tripped, c->mlp_relax);
        hipLaunchKernelGGL(k_sa_ghost, dim3((unsigned)((c->nghost + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_gstart, nb, c->gas);
        HIPCK(c, hipGetLastError());
        return FV3D_OK;
    }
    if (c->cfg.turbulence == 200) return launch_sa(c, save_old, mut_only);
    if (c->cfg.turbulence != 100) return FV3D_OK;
    const int nb = (int)c->blocks.size();
    hipLaunchKernelGGL(k_curl, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb);
    for (size_t r = 0; r < c->bl_launch.size(); ++r) {
        const unsigned wgs = (unsigned)(c->bl_launch[r].second / kBlLines);
        const int2 *wl = c->d_bl_walls + c->bl_launch[r].first;
        const int G = c->bl_groups[r], n = c->bl_launch_n[r];
        if (G == 16)
            hipLaunchKernelGGL(k_bl_lines<16>, dim3(wgs), dim3(kBlLines * 16), 0, c->stream, c->d_blocks, c->d_patches, wl, n, c->gas);
        else if (G == 32)
            hipLaunchKernelGGL(k_bl_lines<32>, dim3(wgs), dim3(kBlLines * 32), 0, c->stream, c->d_blocks, c->d_patches, wl, n, c->gas);
        else
            hipLaunchKernelGGL(k_bl_lines<64>, dim3(wgs), dim3(kBlLines * 64), 0, c->stream, c->d_blocks, c->d_patches, wl, n, c->gas);
    }
    if (!c->bl_launch.empty())
        hipLaunchKernelGGL(k_bl_merge, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                           c->d_cstart, nb);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// Vorticity magnitude (CFluidProps::Curl).  Only the turbulence models read it,
// so the laminar iteration does not evaluate it; fv3d_curl does on request.
static int launch_curl(fv3d_ctx *c)
{
    if (!c->cfg.viscous) return fail(c, FV3D_EINVAL, "the vorticity is evaluated for Navier-Stokes only");
    hipLaunchKernelGGL(k_curl, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart,
                       (int)c->blocks.size());
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// The nu~ equation of the LU-SGS sweep in one direction, one launch per
// hyperplane i + j + k, after the flow sweep of that direction has left its
// increments in dQ (the reference interleaves the two per cell; the nu~ update
// of a cell reads the flow increments of its upstream neighbours only, which
// that flow sweep has finished).  The comparison sweep forms use it; the
// default row-layout sweep carries nu~ itself.
static void launch_sa_sweep(fv3d_ctx *c, bool fwd)
{
    if (!sa_transport(c)) return;
    const unsigned g = (unsigned)((c->nplane_threads + 255) / 256);
    const int nb = (int)c->blocks.size();
    for (int s = 0; s < c->nplanes; ++s) {
        if (fwd)
            hipLaunchKernelGGL(k_sa_plane<true>, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_pstart_plane, nb, s,
                               c->gas.relax, c->gas);
        else
            hipLaunchKernelGGL(k_sa_plane<false>, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_pstart_plane, nb, s,
                               c->gas.relax, c->gas);
    }
}

static int launch_sweeps_kba(fv3d_ctx *c)
{
    // layout: err at 0, tickets at 32 and 48, padded counters from 64
    long nsync = 64 + 2L * c->nkba * kDonePad - 32;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((nsync + 255) / 256)), dim3(256), 0, c->stream, c->d_kba_sync + 32, nsync, 0);
    int *err = c->d_kba_sync, *tf = c->d_kba_sync + 32, *tb = c->d_kba_sync + 48;
    if (c->sweep_mode == 4) {
        const bool f32 = c->cfg.precision == FV3D_DEFECT_FP32;
        const bool visc = c->cfg.viscous;
        const bool sa = sa_transport(c);   // the nu~ equation inside the sweep
        const int nb = (int)c->blocks.size();
        // rows packed per cell, increments per cell, and the split of each
        // (tile, chunk) item over workgroups: about 8192 workgroups per launch,
        // so that a coarse level's few tiles still spread over the chip
        const int nq = sa ? 21 : visc ? 18 : 15, ncomp = sa ? 6 : 5;
        const long per = c->pack_groups > 0 ? std::max(1L, (8192 + c->pack_groups - 1) / c->pack_groups) : 1;
        const int qs = (int)std::min<long>(nq, per), cs = (int)std::min<long>(ncomp, per);
        const unsigned pgq = (unsigned)(c->pack_groups * qs), pgc = (unsigned)(c->pack_groups * cs);
        if (f32) {
            // FP32 words: fill every 32-bit word of the shared allocation
            const long nw = 2 * c->mbox_words;
            hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, c->stream, (int *)c->d_mbox, nw,
                               (int)kMboxEmpty32);
            hipLaunchKernelGGL(k_sweep_pack<float>, dim3(pgq), dim3(256), 0, c->stream, c->d_blocks, c->d_pack_start, nb, nq,
                               qs);
        } else {
            const long nw = c->mbox_words;
            hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, c->stream,
                               (unsigned long long *)c->d_mbox, nw, kMboxEmpty);
            hipLaunchKernelGGL(k_sweep_pack<double>, dim3(pgq), dim3(256), 0, c->stream, c->d_blocks, c->d_pack_start, nb, nq,
                               qs);
        }
        // forward and backward sweeps in the operator's precision, with or
        // without the viscous terms
        auto sweep = [&](bool fwd) {
            int *tk = fwd ? tf : tb;
            const KbaTile *tl = c->d_kba[fwd ? 0 : 1];
            dim3 g((unsigned)c->kba_grid), b(256);
#define FV3D_SL_LAUNCH(F, V, S, T) \
    hipLaunchKernelGGL((k_lusgs_sl<F, kKbaTJ, kKbaTK, V, S, T>), g, b, 0, c->stream, c->d_blocks, tl, c->nkba, tk, err, \
                       c->gas.relax, c->gas)
#define FV3D_SL_FORMS(F, T)                      \
    do {                                         \
        if (sa) FV3D_SL_LAUNCH(F, true, true, T);  \
        else if (visc) FV3D_SL_LAUNCH(F, true, false, T); \
        else FV3D_SL_LAUNCH(F, false, false, T); \
    } while (0)
            if (fwd) {
                if (f32) FV3D_SL_FORMS(true, float);
                else FV3D_SL_FORMS(true, double);
            } else {
                if (f32) FV3D_SL_FORMS(false, float);
                else FV3D_SL_FORMS(false, double);
            }
#undef FV3D_SL_FORMS
#undef FV3D_SL_LAUNCH
        };
        auto unpack = [&]() {
            if (f32)
                hipLaunchKernelGGL(k_sweep_unpack<float>, dim3(pgc), dim3(256), 0, c->stream, c->d_blocks, c->d_pack_start, nb,
                                   ncomp, cs);
            else
                hipLaunchKernelGGL(k_sweep_unpack<double>, dim3(pgc), dim3(256), 0, c->stream, c->d_blocks, c->d_pack_start, nb,
                                   ncomp, cs);
        };
        if (c->timing) HIPCK(c, hipEventRecord(c->ev[6], c->stream));
        sweep(true);
        if (c->timing) HIPCK(c, hipEventRecord(c->ev[7], c->stream));
        sweep(false);
        if (c->timing) HIPCK(c, hipEventRecord(c->ev[8], c->stream));
        unpack();
    }
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

static int launch_sweeps(fv3d_ctx *c)
{
    if (c->sweep_mode >= 1) return launch_sweeps_kba(c);
    unsigned g = (unsigned)((c->nplane_threads + 255) / 256);
    int nb = (int)c->blocks.size();
    for (int s = 0; s < c->nplanes; ++s)
        hipLaunchKernelGGL(k_lusgs_plane<true>, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_pstart_plane, nb, s,
                           c->gas.relax, c->gas);
    launch_sa_sweep(c, true);
    for (int s = 0; s < c->nplanes; ++s)
        hipLaunchKernelGGL(k_lusgs_plane<false>, dim3(g), dim3(256), 0, c->stream, c->d_blocks, c->d_pstart_plane, nb, s,
                           c->gas.relax, c->gas);
    launch_sa_sweep(c, false);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

static int launch_update(fv3d_ctx *c)
{
    int nb = (int)c->blocks.size();
    const bool sa = sa_transport(c);
    int *sa_hi = sa && c->cfg.limiter ? c->d_sa_hi : nullptr;
    hipLaunchKernelGGL(k_update, dim3((unsigned)c->npart), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb, c->gas,
                       c->d_partial, c->d_bad, sa_hi, sa ? 1 : 0);
    HIPCK(c, hipGetLastError());
    if (c->cfg.limiter) {
        hipLaunchKernelGGL(k_limiter_serial, dim3(1), dim3(64), 0, c->stream, c->d_blocks, nb, c->gas, c->d_bad,
                           c->d_cstart, c->d_partial, c->npart);
        if (sa) hipLaunchKernelGGL(k_sa_limit_serial, dim3(1), dim3(64), 0, c->stream, c->d_blocks, nb, c->d_sa_hi);
        HIPCK(c, hipGetLastError());
    }
    if (sa) {
        // Limit_SA's floor and the Res_ev partials (component 5)
        hipLaunchKernelGGL(k_sa_res, dim3((unsigned)c->npart), dim3(256), 0, c->stream, c->d_blocks, c->d_cstart, nb,
                           c->d_partial, c->cfg.limiter ? 1 : 0);
        HIPCK(c, hipGetLastError());
    }
    return FV3D_OK;
}

static int launch_finalize(fv3d_ctx *c, double *hist_row)
{
    hipLaunchKernelGGL(k_sections, dim3(2 * kSecParts * c->ninst), dim3(256), 0, c->stream, c->d_blocks, c->d_secs, c->gas,
                       c->d_sums);
    HIPCK(c, hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(c->ninst), dim3(256), 0, c->stream, c->d_partial, c->npart, c->d_parts, c->d_sums,
                       c->d_sc, c->ramp_sc, c->gas, hist_row);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

#define READY(c)                                                       \
    do {                                                               \
        if (!(c)) return FV3D_EINVAL;                                  \
        if (!(c)->ready) return fail((c), FV3D_ESTATE, "not set up"); \
    } while (0)

int fv3d_bc_apply(fv3d_ctx *c) { READY(c); return launch_bc(c); }
// SpectralRadii and TimeStep share one kernel: both calls form lambda and dt
// (idempotent), and only fv3d_timestep counts the low-dt damping down, once,
// as the reference's TimeStep does.
int fv3d_spectral(fv3d_ctx *c) { READY(c); return launch_spectral_dt(c, false); }
int fv3d_timestep(fv3d_ctx *c) { READY(c); return launch_spectral_dt(c, true); }
int fv3d_convective(fv3d_ctx *c) { READY(c); return launch_convective(c); }
int fv3d_viscous(fv3d_ctx *c) { READY(c); return launch_viscous(c); }
int fv3d_curl(fv3d_ctx *c) { READY(c); return launch_curl(c); }
int fv3d_turbulence(fv3d_ctx *c) { READY(c); return launch_turbulence(c); }

int fv3d_resflux(fv3d_ctx *c)
{
    READY(c);
    int rc;
    if ((rc = launch_bc(c))) return rc;
    if ((rc = launch_turbulence(c))) return rc;
    if ((rc = launch_spectral_dt(c))) return rc;
    if ((rc = launch_convective(c))) return rc;
    return launch_viscous(c);
}

// The wavefront sweep bounds every wait; a wait that ran out sets the error
// word instead of hanging the device.
static int check_sweep_error(fv3d_ctx *c)
{
    int e = 0;
    HIPCK(c, hipMemcpy(&e, c->d_kba_sync, sizeof(int), hipMemcpyDeviceToHost));
    if (e) {
        int z = 0;
        HIPCK(c, hipMemcpy(c->d_kba_sync, &z, sizeof(int), hipMemcpyHostToDevice));
        return fail(c, FV3D_ESTATE, "wavefront sweep: a progress wait exceeded its bound");
    }
    return FV3D_OK;
}

static int lusgs_timed(fv3d_ctx *c)
{
    int rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[0], c->stream));
    if ((rc = launch_bc(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[10], c->stream));
    if ((rc = launch_turbulence(c, 1))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[11], c->stream));
    if ((rc = launch_spectral_dt(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[1], c->stream));
    if ((rc = launch_convective(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[12], c->stream));
    if ((rc = launch_viscous(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[2], c->stream));
    if ((rc = launch_sweeps(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[3], c->stream));
    if ((rc = launch_update(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[4], c->stream));
    return FV3D_OK;
}

int fv3d_lusgs(fv3d_ctx *c)
{
    READY(c);
    int rc = lusgs_timed(c);
    if (rc) return rc;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return check_sweep_error(c);
}

// Res_Calculate of instance w from the norm partials of the last update
static int instance_norms(fv3d_ctx *c, int w, double *res6)
{
    const int p0 = c->parts_host[w].x, np = c->parts_host[w].y;
    const int nc = sa_transport(c) ? 6 : 5;
    std::vector<double> part((size_t)np);
    for (int q = 0; q < 6; ++q) {
        res6[q] = 0.0;
        if (q >= nc) continue;
        HIPCK(c, hipMemcpy(part.data(), c->d_partial + (size_t)q * c->npart + p0, np * sizeof(double),
                           hipMemcpyDeviceToHost));
        double s = 0.0;
        for (int p = 0; p < np; ++p) s += part[p];
        res6[q] = s / c->inst_vol[w];
    }
    return FV3D_OK;
}

// COutput::OutputGeneralParameters of instance w (source/Ouput.cpp:282-300)
static int instance_perf(fv3d_ctx *c, int w, double *out5)
{
    hipLaunchKernelGGL(k_sections, dim3(2 * kSecParts * c->ninst), dim3(256), 0, c->stream, c->d_blocks, c->d_secs, c->gas,
                       c->d_sums);
    HIPCK(c, hipGetLastError());
    double part[2 * 3 * kSecParts];
    HIPCK(c, hipMemcpyAsync(part, c->d_sums + (size_t)w * 2 * 3 * kSecParts, sizeof part, hipMemcpyDeviceToHost,
                            c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    // the partials in the order k_finalize adds them
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int sec = 0; sec < 2; ++sec)
        for (int p = 0; p < kSecParts; ++p)
            for (int q = 0; q < 3; ++q) s[sec * 3 + q] += part[(sec * kSecParts + p) * 3 + q];
    ScalarsDev sc;
    HIPCK(c, hipMemcpy(&sc, c->d_sc + w, sizeof sc, hipMemcpyDeviceToHost));
    const double Gamma = c->gas.gamma, Rgas = c->gas.rgas;
    double mi = s[0] * sc.dens_ref * sc.vel_ref * sc.nblade_in;
    double mo = s[3] * sc.dens_ref * sc.vel_ref * sc.nblade_out;
    double pti = pow(s[1] / s[0], Gamma / (Gamma - F32C(1.0))) * sc.dens_ref * sc.vel_ref * sc.vel_ref;
    double pto = pow(s[4] / s[3], Gamma / (Gamma - F32C(1.0))) * sc.dens_ref * sc.vel_ref * sc.vel_ref;
    double tti = s[2] / s[0] * sc.temp_ref, tto = s[5] / s[3] * sc.temp_ref;
    double PR = pto / pti;
    double L_adk = Gamma / (Gamma - F32C(1.0)) * Rgas * tti * (pow(PR, (Gamma - F32C(1.0)) / Gamma) - F32C(1.0));
    double L_k = Gamma / (Gamma - F32C(1.0)) * Rgas * tti * (tto / tti - F32C(1.0));
    out5[0] = mi; out5[1] = mo; out5[2] = PR; out5[3] = L_adk / L_k; out5[4] = fabs((mi - mo) / mi);
    return FV3D_OK;
}

int fv3d_res_calculate(fv3d_ctx *c, double *res6)
{
    READY(c);
    HIPCK(c, hipStreamSynchronize(c->stream));
    double r[6];
    int rc = instance_norms(c, 0, r);
    if (rc) return rc;
    if (res6) memcpy(res6, r, sizeof r);
    return FV3D_OK;
}

int fv3d_perf(fv3d_ctx *c, double *out5)
{
    READY(c);
    double o[5];
    int rc = instance_perf(c, 0, o);
    if (rc) return rc;
    if (out5) memcpy(out5, o, sizeof o);
    return FV3D_OK;
}

int fv3d_instance_results(fv3d_ctx *c, int inst, double *res6, double *perf5)
{
    READY(c);
    if (inst < 1 || inst > c->ninst) return fail(c, FV3D_EINVAL, "instance %d of %d", inst, c->ninst);
    HIPCK(c, hipStreamSynchronize(c->stream));
    double r[6], o[5];
    int rc = instance_norms(c, inst - 1, r);
    if (!rc) rc = instance_perf(c, inst - 1, o);
    if (rc) return rc;
    if (res6) memcpy(res6, r, sizeof r);
    if (perf5) memcpy(perf5, o, sizeof o);
    return FV3D_OK;
}

static int launch_limiter(fv3d_ctx *c)
{
    if (!c->cfg.limiter) return FV3D_OK;
    hipLaunchKernelGGL(k_limiter_serial, dim3(1), dim3(64), 0, c->stream, c->d_blocks, (int)c->blocks.size(), c->gas,
                       c->d_bad, c->d_cstart, (double *)nullptr, 0);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

// One iteration of the two-grid CSolver::Solver (source/Solver.cpp:61-103,
// reuse_solution = N).  Phases: coarse start-up iterations (LU-SGS on the
// coarse level, which reports), the last of them at Coarse_iter (followed by
// the coarse corner treatment and the interpolation of the coarse state onto
// the fine level), and the cycle: fine LU-SGS and its residual norm, the fine
// residual at the new state, restriction with the forced coarse residual,
// coarse LU-SGS on the lambda, dt and diagonal of the last coarse ResFlux, the
// prolonged correction; the fine level reports.
enum { kMgCoarse = 0, kMgSwitch = 1, kMgCycle = 2 };

// The residual of one level (CMesh::ResFlux, source/ComputeAllBlockVars.cpp:
// 223-255) into B.R, with lambda, dt and the LU-SGS diagonals of that state.
// restriction_only: the evaluation whose residual only the restriction to the
// level below reads (launch_restriction_timestep)
static int level_resflux(fv3d_ctx *c, bool restriction_only = false)
{
    int rc;
    if ((rc = launch_bc(c))) return rc;
    if ((rc = launch_turbulence(c))) return rc;
    if ((rc = restriction_only ? launch_restriction_timestep(c) : launch_spectral_dt(c))) return rc;
    if ((rc = launch_convective(c))) return rc;
    return launch_viscous(c);
}

// Level m of the N-level V-cycle below its finer level f, whose B.R holds its
// total residual: restriction (state and residual), the residual of the
// restricted state and the forcing, one forced LU-SGS step, then (above the
// coarsest level) the total residual restricted to the next level down, its
// correction prolonged back, and the limiter.  The first coarse level of the
// reference's two-grid cycle reuses lambda and dt of its start-up phase
// instead (mg_iteration); every level below it is this one.
static int fas_level(fv3d_ctx *f, fv3d_ctx *m)
{
    int rc;
    const dim3 g((unsigned)((m->ncell + 255) / 256)), b(256);
    const int nb = (int)m->blocks.size();
    hipLaunchKernelGGL(k_mg_restrict, g, b, 0, f->stream, f->d_blocks, m->d_blocks, m->d_cstart, nb, m->gas, m->d_bad);
    HIPCK(f, hipGetLastError());
    if ((rc = launch_limiter(m))) return rc;
    if ((rc = level_resflux(m))) return rc;
    hipLaunchKernelGGL(k_fas<true>, g, b, 0, f->stream, m->d_blocks, m->d_cstart, nb);
    HIPCK(f, hipGetLastError());
    if ((rc = launch_sweeps(m))) return rc;
    if ((rc = launch_update(m))) return rc;
    if (m->coarse) {
        if ((rc = level_resflux(m, true))) return rc;
        hipLaunchKernelGGL(k_fas<false>, g, b, 0, f->stream, m->d_blocks, m->d_cstart, nb);
        HIPCK(f, hipGetLastError());
        if ((rc = fas_level(m, m->coarse))) return rc;
        hipLaunchKernelGGL(k_mg_prolong<false>, dim3((unsigned)m->n_mg_tiles), b, 0, f->stream, m->d_blocks,
                           m->coarse->d_blocks, m->d_mg_tiles, m->gas, m->d_bad);
        HIPCK(f, hipGetLastError());
        if ((rc = launch_limiter(m))) return rc;
    }
    return FV3D_OK;
}

static int mg_iteration(fv3d_ctx *c, int phase, double *row)
{
    fv3d_ctx *m = c->coarse;
    const int nb = (int)c->blocks.size();
    int rc;
    if (phase != kMgCycle) {
        if ((rc = lusgs_timed(m))) return rc;
        if (phase == kMgSwitch) {
            hipLaunchKernelGGL(k_mg_corner, dim3((unsigned)nb), dim3(256), 0, c->stream, m->d_blocks);
            hipLaunchKernelGGL(k_mg_prolong<true>, dim3((unsigned)c->n_mg_tiles), dim3(256), 0, c->stream, c->d_blocks,
                               m->d_blocks, c->d_mg_tiles, c->gas, c->d_bad);
            HIPCK(c, hipGetLastError());
            if ((rc = launch_limiter(c))) return rc;
        }
        return launch_finalize(m, row);
    }
    if ((rc = lusgs_timed(c))) return rc;
    // the fine residual at the new state, for the restriction (the coarse
    // levels carry no turbulence model, so the nu~ residual is not needed)
    if ((rc = launch_bc(c))) return rc;
    if ((rc = launch_turbulence(c, 0, true))) return rc;
    if ((rc = launch_restriction_timestep(c))) return rc;
    if ((rc = launch_convective(c))) return rc;
    if ((rc = launch_viscous(c))) return rc;
    if (m->coarse) {
        // N-level V-cycle: the levels below take the full approximation storage form
        if ((rc = fas_level(c, m))) return rc;
    } else {
        hipLaunchKernelGGL(k_mg_restrict, dim3((unsigned)((m->ncell + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_blocks, m->d_blocks, m->d_cstart, nb, m->gas, m->d_bad);
        HIPCK(c, hipGetLastError());
        if ((rc = launch_limiter(m))) return rc;
        if ((rc = launch_sweeps(m))) return rc;
        if ((rc = launch_update(m))) return rc;
    }
    hipLaunchKernelGGL(k_mg_prolong<false>, dim3((unsigned)c->n_mg_tiles), dim3(256), 0, c->stream, c->d_blocks,
                       m->d_blocks, c->d_mg_tiles, c->gas, c->d_bad);
    HIPCK(c, hipGetLastError());
    if ((rc = launch_limiter(c))) return rc;
    if (c->timing) HIPCK(c, hipEventRecord(c->ev[9], c->stream));
    return launch_finalize(c, row);
}

static int mg_phase(const fv3d_ctx *c, long it)
{
    return it < c->cfg.coarse_iter ? kMgCoarse : it == c->cfg.coarse_iter ? kMgSwitch : kMgCycle;
}


static int ensure_hist(fv3d_ctx *c, long n)
{
    if (n <= c->hist_cap) return FV3D_OK;
    if (c->d_hist) HIPCK(c, hipFree(c->d_hist));
    c->d_hist = nullptr;
    c->hist_cap = 0;
    int rc = dalloc(c, &c->d_hist, 7 * (size_t)n);
    if (rc) return rc;
    c->hist_cap = n;
    return FV3D_OK;
}

// ------------------------------------------------- two-grid cycle, by stage
// The stages of one CSolver::Solver iteration (source/Solver.cpp:61-103) as
// separate calls, for a caller that drives the schedule itself; the same
// kernels as fv3d_iterate, so the same bits.
#define TWO_LEVEL(c)                                                                                          \
    do {                                                                                                      \
        READY(c);                                                                                             \
        if ((c)->ng != 2 || (c)->coarse->coarse)                                                              \
            return fail((c), FV3D_ESTATE, "the stage calls drive the two-grid cycle (two levels in use: %d)", \
                        (c)->nlev);                                                                           \
    } while (0)

int fv3d_mg_coarse_lusgs(fv3d_ctx *c, int forced)
{
    TWO_LEVEL(c);
    fv3d_ctx *m = c->coarse;
    int rc;
    if (!forced) return lusgs_timed(m);
    if ((rc = launch_sweeps(m))) return rc;
    return launch_update(m);
}

int fv3d_mg_restrict(fv3d_ctx *c)
{
    TWO_LEVEL(c);
    fv3d_ctx *m = c->coarse;
    hipLaunchKernelGGL(k_mg_restrict, dim3((unsigned)((m->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                       m->d_blocks, m->d_cstart, (int)c->blocks.size(), m->gas, m->d_bad);
    HIPCK(c, hipGetLastError());
    return launch_limiter(m);
}

int fv3d_mg_correct(fv3d_ctx *c)
{
    TWO_LEVEL(c);
    fv3d_ctx *m = c->coarse;
    hipLaunchKernelGGL(k_mg_prolong<false>, dim3((unsigned)c->n_mg_tiles), dim3(256), 0, c->stream, c->d_blocks,
                       m->d_blocks, c->d_mg_tiles, c->gas, c->d_bad);
    HIPCK(c, hipGetLastError());
    return launch_limiter(c);
}

int fv3d_mg_interpolate_state(fv3d_ctx *c)
{
    TWO_LEVEL(c);
    fv3d_ctx *m = c->coarse;
    hipLaunchKernelGGL(k_mg_corner, dim3((unsigned)c->blocks.size()), dim3(256), 0, c->stream, m->d_blocks);
    hipLaunchKernelGGL(k_mg_prolong<true>, dim3((unsigned)c->n_mg_tiles), dim3(256), 0, c->stream, c->d_blocks,
                       m->d_blocks, c->d_mg_tiles, c->gas, c->d_bad);
    HIPCK(c, hipGetLastError());
    return launch_limiter(c);
}

int fv3d_iteration_end(fv3d_ctx *c, double *row7)
{
    READY(c);
    int rc;
    if ((rc = ensure_hist(c, c->ninst))) return rc;
    const long itn = c->iter_host + 1;
    // the coarse start-up iterations report the coarse level (source/Solver.cpp:71-87)
    fv3d_ctx *r = c->ng == 2 && mg_phase(c, itn) != kMgCycle ? c->coarse : c;
    if ((rc = launch_finalize(r, c->d_hist))) return rc;
    c->iter_host = itn;
    if (row7) HIPCK(c, hipMemcpyAsync(row7, c->d_hist, 7L * c->ninst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((rc = check_sweep_error(c))) return rc;
    if (c->coarse && (rc = check_sweep_error(c->coarse))) return fail(c, rc, "coarse level: %s", c->coarse->err.c_str());
    return FV3D_OK;
}

int fv3d_primitive_vars(fv3d_ctx *c)
{
    READY(c);
    hipLaunchKernelGGL(k_prim_all, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, c->stream, c->d_blocks,
                       c->d_cstart, (int)c->blocks.size(), c->gas, c->d_bad);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

int fv3d_limit_vars(fv3d_ctx *c)
{
    READY(c);
    if (!c->cfg.limiter) return FV3D_OK;
    // every block is scanned (the flags of the iteration's kernels are not relied on)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, c->stream, c->d_bad, (int)c->blocks.size(), 1);
    HIPCK(c, hipGetLastError());
    return launch_limiter(c);
}

static int accumulate_timing(fv3d_ctx *c)
{
    HIPCK(c, hipEventSynchronize(c->ev[5]));
    float ms;
    // slots: 0 residual (convective + viscous), 1 LU-SGS operator (pack, sweeps,
    // unpack), 2 update, 3 ghost fill + turbulence + spectral radii and dt,
    // 4 finalize, 5/6 forward/backward sweep kernels, 7 the cycle's coarse
    // correction (second fine residual, transfers, coarse levels), 8 convective
    // residual kernel, 9 viscous (and nu~) residual kernels, 10 ghost fill and
    // patch exchange, 11 turbulence model (vorticity, Baldwin-Lomax,
    // Spalart-Allmaras or the learned closure)
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2])); c->t_ms[0] += ms; c->t_n[0]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3])); c->t_ms[1] += ms; c->t_n[1]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[3], c->ev[4])); c->t_ms[2] += ms; c->t_n[2]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); c->t_ms[3] += ms; c->t_n[3]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[9], c->ev[5])); c->t_ms[4] += ms; c->t_n[4]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[4], c->ev[9])); c->t_ms[7] += ms; c->t_n[7]++;
    if (c->sweep_mode == 4) {
        // the forward and backward row-layout sweep kernels alone
        HIPCK(c, hipEventElapsedTime(&ms, c->ev[6], c->ev[7])); c->t_ms[5] += ms; c->t_n[5]++;
        HIPCK(c, hipEventElapsedTime(&ms, c->ev[7], c->ev[8])); c->t_ms[6] += ms; c->t_n[6]++;
    }
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[12])); c->t_ms[8] += ms; c->t_n[8]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[12], c->ev[2])); c->t_ms[9] += ms; c->t_n[9]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[10])); c->t_ms[10] += ms; c->t_n[10]++;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[10], c->ev[11])); c->t_ms[11] += ms; c->t_n[11]++;
    return FV3D_OK;
}

int fv3d_iterate(fv3d_ctx *c, int n, double *hist)
{
    READY(c);
    if (n < 0) return fail(c, FV3D_EINVAL, "n=%d", n);
    int rc;
    if ((rc = ensure_hist(c, (long)std::max(n, 1) * c->ninst))) return rc;
    if (c->ng == 2 && !c->initial_done)
        return fail(c, FV3D_ESTATE, "the two-grid cycle starts from fv3d_initial (reuse_solution = N): from a restart "
                                    "file the reference's cycle divides by coarse time steps it never computed");
    if (c->cfg.turbulence == 200 && !c->ev_set)
        return fail(c, FV3D_ESTATE, "Spalart-Allmaras needs nu~: fv3d_nut_upload for every block, or fv3d_initial");
    for (int it = 0; it < n; ++it) {
        const long itn = c->iter_host + 1;
        double *row = c->d_hist + 7L * c->ninst * it;
        bool timed = true;
        if (c->ng == 2) {
            const int phase = mg_phase(c, itn);
            timed = phase == kMgCycle;
            if ((rc = mg_iteration(c, phase, row))) return rc;
        } else {
            if ((rc = lusgs_timed(c))) return rc;
            if (c->timing) HIPCK(c, hipEventRecord(c->ev[9], c->stream));
            if ((rc = launch_finalize(c, row))) return rc;
        }
        c->iter_host = itn;
        if (c->timing && timed) {
            HIPCK(c, hipEventRecord(c->ev[5], c->stream));
            if ((rc = accumulate_timing(c))) return rc;
        }
    }
    if (hist && n > 0)
        HIPCK(c, hipMemcpyAsync(hist, c->d_hist, 7L * n * c->ninst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((rc = check_sweep_error(c))) return rc;
    if (c->coarse && (rc = check_sweep_error(c->coarse))) return fail(c, rc, "coarse level: %s", c->coarse->err.c_str());
    if (hist && n > 0)
        for (int w = 0; w < c->ninst; ++w)
            if (!isfinite(hist[7L * (c->ninst * (n - 1L) + w) + 1]))
                return fail(c, FV3D_ENUMERIC, c->ninst > 1 ? "instance %d: residual is not finite" : "residual is not finite",
                            w + 1);
    return FV3D_OK;
}

int fv3d_iterate_graph(fv3d_ctx *c, int n, double *hist)
{
    READY(c);
    if (n < 0) return fail(c, FV3D_EINVAL, "n=%d", n);
    int rc;
    if ((rc = ensure_hist(c, (long)std::max(n, 1) * c->ninst))) return rc;
    if (c->ng == 2 && !c->initial_done)
        return fail(c, FV3D_ESTATE, "the two-grid cycle starts from fv3d_initial (reuse_solution = N): from a restart "
                                    "file the reference's cycle divides by coarse time steps it never computed");
    if (c->cfg.turbulence == 200 && !c->ev_set)
        return fail(c, FV3D_ESTATE, "Spalart-Allmaras needs nu~: fv3d_nut_upload for every block, or fv3d_initial");
    // each graph writes its history row into a fixed staging slot that is
    // copied into the history buffer after each replay; with two levels the
    // coarse start-up phase and the cycle have a graph each, and the one
    // iteration between them (Coarse_iter) runs eagerly
    double *stage_row = c->d_stage_row;
    auto capture = [&](hipGraphExec_t *exec, int phase) -> int {
        if (*exec) return FV3D_OK;
        hipGraph_t g;
        int saved = c->timing;
        c->timing = 0;
        HIPCK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r;
        if (c->ng == 2) {
            r = mg_iteration(c, phase, stage_row);
        } else {
            r = lusgs_timed(c);
            if (!r) r = launch_finalize(c, stage_row);
        }
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->timing = saved;
        if (r) return r;
        if (e != hipSuccess) return fail(c, FV3D_EHIP, "graph capture: %s", hipGetErrorString(e));
        HIPCK(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
        HIPCK(c, hipGraphDestroy(g));
        return FV3D_OK;
    };
    for (int it = 0; it < n; ++it) {
        const long itn = c->iter_host + 1;
        const int phase = c->ng == 2 ? mg_phase(c, itn) : kMgCycle;
        if (phase == kMgSwitch) {
            if ((rc = mg_iteration(c, phase, stage_row))) return rc;
        } else {
            hipGraphExec_t *exec = phase == kMgCoarse ? &c->graph_coarse : &c->graph_exec;
            if ((rc = capture(exec, phase))) return rc;
            HIPCK(c, hipGraphLaunch(*exec, c->stream));
        }
        HIPCK(c, hipMemcpyAsync(c->d_hist + 7L * c->ninst * it, stage_row, 7L * c->ninst * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
        c->iter_host = itn;
    }
    if (hist && n > 0)
        HIPCK(c, hipMemcpyAsync(hist, c->d_hist, 7L * n * c->ninst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if ((rc = check_sweep_error(c))) return rc;
    if (c->coarse && (rc = check_sweep_error(c->coarse))) return fail(c, rc, "coarse level: %s", c->coarse->err.c_str());
    if (hist && n > 0)
        for (int w = 0; w < c->ninst; ++w)
            if (!isfinite(hist[7L * (c->ninst * (n - 1L) + w) + 1]))
                return fail(c, FV3D_ENUMERIC, c->ninst > 1 ? "instance %d: residual is not finite" : "residual is not finite",
                            w + 1);
    return FV3D_OK;
}

int fv3d_sync(fv3d_ctx *c)
{
    if (!c) return FV3D_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return FV3D_OK;
}

int fv3d_set_timing(fv3d_ctx *c, int on)
{
    if (!c) return FV3D_EINVAL;
    c->timing = on;
    for (int q = 0; q < FV3D_TIMING_SLOTS; ++q) { c->t_ms[q] = 0.0; c->t_n[q] = 0; }
    return FV3D_OK;
}

int fv3d_timing(fv3d_ctx *c, double *ms, long *counts)
{
    if (!c) return FV3D_EINVAL;
    for (int q = 0; q < FV3D_TIMING_SLOTS; ++q) {
        if (ms) ms[q] = c->t_n[q] ? c->t_ms[q] / c->t_n[q] : 0.0;
        if (counts) counts[q] = c->t_n[q];
    }
    return FV3D_OK;
}

long fv3d_cell_count(const fv3d_ctx *c) { return c ? c->ncell_real : 0; }

// Invalidates the captured graphs after a change of launch configuration.
static void drop_graphs(fv3d_ctx *c)
{
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph_coarse) { (void)hipGraphExecDestroy(c->graph_coarse); c->graph_coarse = nullptr; }
}

int fv3d_set_tiles(fv3d_ctx *c, const int *shape3)
{
    READY(c);
    if (!shape3) return FV3D_EINVAL;
    for (int k = 0; k < 3; ++k) {
        if (shape3[k] < 0 || shape3[k] >= (k == 0 ? kNumConv : kNumTiles))
            return fail(c, FV3D_EINVAL, "tile shape %d", shape3[k]);
        int rc = ensure_tiles(c, shape3[k] < kNumTiles ? shape3[k] : 0);
        if (!rc && k == 0 && shape3[k] == kConvFaces) rc = ensure_face_form(c);
        else if (!rc && k == 0 && shape3[k] >= kNumTiles) rc = ensure_stream_tiles(c, shape3[k] - kNumTiles);
        if (rc) return rc;
    }
    for (int k = 0; k < 3; ++k) c->tile_sel[k] = shape3[k];
    drop_graphs(c);
    return FV3D_OK;
}

int fv3d_autotune(fv3d_ctx *c, int reps, int *shape3)
{
    READY(c);
    if (reps < 1) return fail(c, FV3D_EINVAL, "reps=%d", reps);
    HIPCK(c, hipStreamSynchronize(c->stream));
    const int saved = c->store_faces;
    c->store_faces = 0;
    const bool use[3] = {true, c->cfg.viscous != 0, c->cfg.turbulence == 200};
    int rc = FV3D_OK;
    for (int kern = 0; kern < 3 && !rc; ++kern) {
        if (shape3) shape3[kern] = -1;
        if (!use[kern]) continue;
        auto run = [&]() -> int {
            if (kern == 0) return launch_convective(c);
            if (kern == 1) return launch_viscous(c);
            launch_sa_faces(c);
            HIPCK(c, hipGetLastError());
            return FV3D_OK;
        };
        float best = 0.0f;
        int bestq = 0;
        for (int q = 0; q < (kern == 0 ? kNumConv : kNumTiles) && !rc; ++q) {
            if (q == kConvFaces) {
                // the face-parallel form is for small grids (as on the coarse
                // levels); on a large one it writes and reads back every face
                if (c->ncell > kFaceFormCells) continue;
                if ((rc = ensure_face_form(c))) break;
            } else if (q >= kNumTiles) {
                if ((rc = ensure_stream_tiles(c, q - kNumTiles))) break;
            } else if ((rc = ensure_tiles(c, q))) break;
            c->tile_sel[kern] = q;
            if ((rc = run())) break;   // warm-up
            HIPCK(c, hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps && !rc; ++r) rc = run();
            HIPCK(c, hipEventRecord(c->ev[1], c->stream));
            HIPCK(c, hipEventSynchronize(c->ev[1]));
            float ms = 0.0f;
            HIPCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            if (q == 0 || ms < best) { best = ms; bestq = q; }
        }
        c->tile_sel[kern] = bestq;
        if (shape3) shape3[kern] = bestq;
    }
    c->store_faces = saved;
    drop_graphs(c);
    return rc;
}

// ---------------------------------------------------- learned closure
static int mlp_alloc(fv3d_ctx *c)
{
    if (c->d_mlp) return FV3D_OK;
    int rc;
    if ((rc = dalloc(c, &c->d_mlp, kMlpParams))) return rc;
    if ((rc = dalloc(c, &c->d_mlp_m, kMlpParams))) return rc;
    if ((rc = dalloc(c, &c->d_mlp_v, kMlpParams))) return rc;
    if ((rc = dalloc(c, &c->d_mlp_g, kMlpParams + 1))) return rc;
    if ((rc = dalloc(c, &c->d_mlp_bounds, 2 * kMlpIn + 1))) return rc;
    if ((rc = dalloc(c, &c->d_mlp_tripped, (size_t)kTripSlots * kTripStride))) return rc;
    return FV3D_OK;
}

int fv3d_mlp_set(fv3d_ctx *c, const float *params, const float *bounds)
{
    READY(c);
    if (!params) return FV3D_EINVAL;
    int rc = mlp_alloc(c);
    if (rc) return rc;
    std::vector<float> b(2 * kMlpIn + 1);
    for (int q = 0; q < kMlpIn; ++q) { b[2 * q] = -1e30f; b[2 * q + 1] = 1e30f; }
    b[2 * kMlpIn] = 1e30f;
    if (bounds) std::copy(bounds, bounds + 2 * kMlpIn + 1, b.begin());
    HIPCK(c, hipMemcpyAsync(c->d_mlp, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->d_mlp_bounds, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    // a fresh optimiser state
    hipLaunchKernelGGL(k_fill_int, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, (int *)c->d_mlp_m,
                       (long)kMlpParams, 0);
    hipLaunchKernelGGL(k_fill_int, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, (int *)c->d_mlp_v,
                       (long)kMlpParams, 0);
    HIPCK(c, hipGetLastError());
    c->mlp_step = 0;
    HIPCK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    return FV3D_OK;
}

int fv3d_mlp_get(fv3d_ctx *c, float *params)
{
    READY(c);
    if (!params) return FV3D_EINVAL;
    if (!c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(params, c->d_mlp, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return FV3D_OK;
}

// Per-cell closure data of one block: the features (what = 0, 6 per cell) or
// the training target (what = 1).
static long mlp_block_data(fv3d_ctx *c, int block, float *out, int what)
{
    READY(c);
    if (!out) return FV3D_EINVAL;
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "block %d", block);
    if (!c->cfg.viscous || !c->blocks[block - 1].dev.dwall)
        return fail(c, FV3D_EINVAL, "closure features need the Navier-Stokes terms and a wall distance (Spalart-Allmaras)");
    const BlockDev &B = c->blocks[block - 1].dev;
    const long ncb = (long)B.ni * B.nj * B.nk;
    const int per = what == 0 ? kMlpIn : 1;
    float *d = nullptr;
    int rc = dalloc(c, &d, (size_t)ncb * per);
    if (rc) return rc;
    // the block alone: a one-block cell list
    long *ds = nullptr;
    const long st[2] = {0, ncb};
    if ((rc = dalloc(c, &ds, 2))) { (void)hipFree(d); return rc; }
    HIPCK(c, hipMemcpyAsync(ds, st, sizeof st, hipMemcpyHostToDevice, c->stream));
    const long fs = 0;
    long *dfs = nullptr;
    if ((rc = dalloc(c, &dfs, 1))) { (void)hipFree(d); (void)hipFree(ds); return rc; }
    HIPCK(c, hipMemcpyAsync(dfs, &fs, sizeof fs, hipMemcpyHostToDevice, c->stream));
    if (what == 0)
        hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((ncb + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_blocks + (block - 1), ds, 1, c->gas, dfs, d);
    else
        hipLaunchKernelGGL(k_mlp_targets, dim3((unsigned)((ncb + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_blocks + (block - 1), ds, 1, c->gas, dfs, d);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(out, d, (size_t)ncb * per * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(d); (void)hipFree(ds); (void)hipFree(dfs);
    return ncb * per;
}

long fv3d_mlp_features(fv3d_ctx *c, int block, float *out) { return mlp_block_data(c, block, out, 0); }

long fv3d_mlp_targets(fv3d_ctx *c, int block, float *out)
{
    if (c && c->ready && c->cfg.turbulence == 0) return fail(c, FV3D_EINVAL, "targets need an eddy viscosity");
    return mlp_block_data(c, block, out, 1);
}

// Device copies of n samples (features, and targets when t is given).
static int mlp_upload(fv3d_ctx *c, const float *x, const float *t, long n, float **dx, float **dt)
{
    int rc = dalloc(c, dx, (size_t)n * kMlpIn);
    if (rc) return rc;
    HIPCK(c, hipMemcpyAsync(*dx, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (t) {
        if ((rc = dalloc(c, dt, (size_t)n))) return rc;
        HIPCK(c, hipMemcpyAsync(*dt, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    return FV3D_OK;
}

int fv3d_mlp_forward(fv3d_ctx *c, const float *x, long n, float *y)
{
    READY(c);
    if (!x || !y || n < 1) return FV3D_EINVAL;
    if (!c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    float *dx = nullptr, *dy = nullptr, *unused = nullptr;
    int rc = mlp_upload(c, x, nullptr, n, &dx, &unused);
    if (!rc) rc = dalloc(c, &dy, (size_t)n);
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_mlp, dx, n, dy,
                           (float *)nullptr);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, FV3D_EHIP, "mlp forward: %s", hipGetErrorString(e));
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

// One gradient of the mean loss over samples [off, off + m) into d_mlp_g
// (and the summed squared error into d_mlp_g[kMlpParams]).
static void mlp_gradient(fv3d_ctx *c, const float *dx, const float *dt, long off, long m, float *dpart, float *dloss)
{
    const int nchunk = (int)((m + 63) / 64);
    hipLaunchKernelGGL(k_mlp_train_chunk, dim3((unsigned)nchunk), dim3(256), 0, c->stream, c->d_mlp, dx + off * kMlpIn,
                       dt + off, m, 1.0f / (float)m, dpart, dloss);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, c->stream, dpart, dloss, nchunk,
                       c->d_mlp_g, c->d_mlp_g + kMlpParams);
}

int fv3d_mlp_grad(fv3d_ctx *c, const float *x, const float *t, long n, float *grad, float *loss)
{
    READY(c);
    if (!x || !t || n < 1) return FV3D_EINVAL;
    if (!c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    float *dx = nullptr, *dt = nullptr, *dpart = nullptr, *dloss = nullptr;
    const int nchunk = (int)((n + 63) / 64);
    int rc = mlp_upload(c, x, t, n, &dx, &dt);
    if (!rc) rc = dalloc(c, &dpart, (size_t)nchunk * kMlpParams);
    if (!rc) rc = dalloc(c, &dloss, (size_t)nchunk);
    if (!rc) {
        mlp_gradient(c, dx, dt, 0, n, dpart, dloss);
        std::vector<float> g(kMlpParams + 1);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(g.data(), c->d_mlp_g, g.size() * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, FV3D_EHIP, "mlp gradient: %s", hipGetErrorString(e));
        else {
            if (grad) std::copy(g.begin(), g.begin() + kMlpParams, grad);
            if (loss) *loss = g[kMlpParams] / (float)n;
        }
    }
    for (float *p : {dx, dt, dpart, dloss})
        if (p) (void)hipFree(p);
    return rc;
}

int fv3d_mlp_grad_noalloc(fv3d_ctx *c, const float *x_dev, const float *t_dev, long n, float *grad_dev,
                          float *loss_dev)
{
    READY(c);
    if (!x_dev || !t_dev || !grad_dev || !loss_dev || n < 1) return fail(c, FV3D_EINVAL, "mlp_grad_noalloc arguments");
    if (n > 64L * kNoAllocChunks)
        return fail(c, FV3D_EINVAL, "mlp_grad_noalloc: %ld samples, at most %d", n, 64 * kNoAllocChunks);
    if (!c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    float *s = g_scratch_ptr[c->device];
    if (!s) return fail(c, FV3D_ESTATE, "scratch address of device %d not taken", c->device);
    // every chunk's partial gradient and loss is written before it is read, so
    // nothing in the scratch needs clearing between calls
    float *part = s, *loss = s + (long)kNoAllocChunks * kMlpParams, *sum = s + kNoAllocFloats - 1;
    const int nchunk = (int)((n + 63) / 64);
    hipLaunchKernelGGL(k_mlp_train_chunk, dim3((unsigned)nchunk), dim3(256), 0, c->stream, c->d_mlp, x_dev, t_dev, n,
                       1.0f / (float)n, part, loss);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, c->stream, part, loss, nchunk,
                       grad_dev, sum);
    hipLaunchKernelGGL(k_noalloc_loss, dim3(1), dim3(64), 0, c->stream, sum, n, loss_dev);
    HIPCK(c, hipGetLastError());
    return FV3D_OK;
}

int fv3d_mlp_train(fv3d_ctx *c, const float *x, const float *t, long n, int steps, int batch, float lr,
                   float *loss_hist)
{
    READY(c);
    if (!x || !t || n < 1 || steps < 1 || batch < 64 || batch % 64) return fail(c, FV3D_EINVAL, "training arguments");
    if (!c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    const long m = std::min<long>(batch, n);
    const int nchunk = (int)((m + 63) / 64);
    float *dx = nullptr, *dt = nullptr, *dpart = nullptr, *dloss = nullptr, *dhist = nullptr;
    int rc = mlp_upload(c, x, t, n, &dx, &dt);
    if (!rc) rc = dalloc(c, &dpart, (size_t)nchunk * kMlpParams);
    if (!rc) rc = dalloc(c, &dloss, (size_t)nchunk);
    if (!rc) rc = dalloc(c, &dhist, (size_t)steps);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps && !rc; ++s) {
        // consecutive minibatches of the sample order, wrapping (the caller shuffles)
        const long off = (long)s * m % std::max<long>(1, n - m + 1);
        mlp_gradient(c, dx, dt, off, m, dpart, dloss);
        ++c->mlp_step;
        const float c1 = 1.0f - powf(b1, (float)c->mlp_step), c2 = 1.0f - powf(b2, (float)c->mlp_step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, c->d_mlp, c->d_mlp_m,
                           c->d_mlp_v, c->d_mlp_g, lr, b1, b2, eps, c1, c2);
        HIPCK(c, hipMemcpyAsync(dhist + s, c->d_mlp_g + kMlpParams, sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, FV3D_EHIP, "mlp training: %s", hipGetErrorString(e));
    }
    if (!rc) {
        std::vector<float> h(steps);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(h.data(), dhist, steps * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, FV3D_EHIP, "mlp training: %s", hipGetErrorString(e));
        else if (loss_hist)
            for (int s = 0; s < steps; ++s) loss_hist[s] = h[s] / (float)m;
    }
    for (float *p : {dx, dt, dpart, dloss, dhist})
        if (p) (void)hipFree(p);
    drop_graphs(c);
    return rc;
}

int fv3d_mlp_closure(fv3d_ctx *c, int on)
{
    READY(c);
    if (c->cfg.turbulence != 200)
        return fail(c, FV3D_EINVAL, "the learned closure replaces the Spalart-Allmaras transport (turbulence 200)");
    if (on && !c->d_mlp) return fail(c, FV3D_ESTATE, "no network parameters (fv3d_mlp_set)");
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (on && !c->d_mlp_xc) {
        const int rc = dalloc(c, &c->d_mlp_xc, (size_t)8 * c->ncell);
        if (rc) return rc;
    }
    c->mlp_closure = on != 0;
    drop_graphs(c);
    return FV3D_OK;
}

int fv3d_mlp_relax(fv3d_ctx *c, double omega)
{
    READY(c);
    if (!(omega > 0.0 && omega <= 1.0)) return fail(c, FV3D_EINVAL, "closure relaxation %g not in (0, 1]", omega);
    c->mlp_relax = omega;
    drop_graphs(c);
    return FV3D_OK;
}

long fv3d_mlp_tripped(fv3d_ctx *c)
{
    READY(c);
    if (!c->d_mlp_tripped) return 0;
    // the sum of the counter slots, which are then cleared
    std::vector<unsigned long long> v((size_t)kTripSlots * kTripStride), z(v.size(), 0);
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(v.data(), c->d_mlp_tripped, v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    HIPCK(c, hipMemcpy(c->d_mlp_tripped, z.data(), z.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    unsigned long long sum = 0;
    for (int q = 0; q < kTripSlots; ++q) sum += v[(size_t)q * kTripStride];
    return (long)sum;
}

int fv3d_set_debug(fv3d_ctx *c, int flags)
{
    READY(c);
    c->store_faces = flags & 1;
    if (flags & ~3) return fail(c, FV3D_EINVAL, "debug flags 0x%x: only bits 0 (face capture) and 1 (one launch "
                                                 "per hyperplane) exist", flags);
    c->sweep_mode = (flags & 2) ? 0 : 4;
    if (c->cfg.precision == FV3D_DEFECT_FP32 && c->sweep_mode != 4) {
        c->sweep_mode = 4;
        return fail(c, FV3D_EINVAL, "the FP32 operator runs with the default sweep form only");
    }
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->store_faces && !c->d_faces[0]) {
        long mx = 0;
        for (auto &b : c->blocks) mx = std::max(mx, b.dev.box);
        if (c->blocks.size() != 1) return fail(c, FV3D_EINVAL, "face-flux capture supports one block");
        for (int d = 0; d < 3; ++d) {
            int rc = dalloc(c, &c->d_faces[d], 5 * (size_t)mx);
            if (rc) return rc;
            if (c->cfg.viscous) {
                if ((rc = dalloc(c, &c->vdbg.fv[d], 5 * (size_t)mx))) return rc;
                if ((rc = dalloc(c, &c->vdbg.grad[d], 12 * (size_t)mx))) return rc;
            }
        }
        if (c->cfg.viscous) {
            int rc = dalloc(c, &c->vdbg.fvc, 5 * (size_t)mx);
            if (rc) return rc;
        }
    }
    return FV3D_OK;
}

long fv3d_level_field_download(fv3d_ctx *c, int level, int block, const char *name, double *out)
{
    if (!c) return FV3D_EINVAL;
    if (level == 1) return fv3d_field_download(c, block, name, out);
    if (level != 2 || !c->coarse) return fail(c, FV3D_EINVAL, "grid level %d not in use", level);
    long r = fv3d_field_download(c->coarse, block, name, out);
    if (r < 0) fail(c, (int)r, "coarse level: %s", c->coarse->err.c_str());
    return r;
}

int fv3d_wall_distance_upload(fv3d_ctx *c, int block, const double *d)
{
    READY(c);
    if (!d) return FV3D_EINVAL;
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "block %d", block);
    const BlockDev &B = c->blocks[block - 1].dev;
    if (!B.dwall) return fail(c, FV3D_EINVAL, "the wall distance exists with Spalart-Allmaras only");
    std::vector<double> h((size_t)B.box);
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipMemcpy(h.data(), B.dwall, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    long q = 0;
    for (int i = 1; i <= B.ni; ++i)
        for (int j = 1; j <= B.nj; ++j)
            for (int k = 1; k <= B.nk; ++k) h[cidx(B, i, j, k)] = d[q++];
    HIPCK(c, hipMemcpy(B.dwall, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return FV3D_OK;
}

long fv3d_field_download(fv3d_ctx *c, int block, const char *name, double *out)
{
    READY(c);
    if (!name || !out) return FV3D_EINVAL;
    if (block < 1 || block > (int)c->blocks.size()) return fail(c, FV3D_EINVAL, "block %d", block);
    const BlockDev &B = c->blocks[block - 1].dev;
    HIPCK(c, hipStreamSynchronize(c->stream));
    auto fetch = [&](const double *src, long n, std::vector<double> &dst) -> int {
        dst.resize(n);
        HIPCK(c, hipMemcpy(dst.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    };
    std::vector<double> h;
    long cnt = 0;
    std::string nm(name);
    if (nm == "mut" || nm == "ev") {
        const double *p = nm == "mut" ? B.mut : B.ev;
        if (!p) return fail(c, FV3D_EINVAL, "%s exists with %s only", name, nm == "mut" ? "a turbulence model" : "Spalart-Allmaras");
        if (fetch(p, B.box, h)) return FV3D_EHIP;
        for (int i = 0; i <= B.IM; ++i)
            for (int j = 0; j <= B.JM; ++j)
                for (int k = 0; k <= B.KM; ++k) out[cnt++] = h[cidx(B, i, j, k)];
        return cnt;
    }
    if (nm == "cv" || nm == "pv") {
        int C = nm == "cv" ? 5 : 6;
        if (fetch(nm == "cv" ? B.Q : B.W, C * B.box, h)) return FV3D_EHIP;
        for (int i = 0; i <= B.IM; ++i)
            for (int j = 0; j <= B.JM; ++j)
                for (int k = 0; k <= B.KM; ++k)
                    for (int q = 0; q < C; ++q) out[cnt++] = h[q * B.box + cidx(B, i, j, k)];
        return cnt;
    }
    const double *src = nullptr;
    int C = 0, hiI = 0, hiJ = 0, hiK = 0;
    if (nm == "res") { src = B.R; C = 5; }
    else if (nm == "dq") { src = B.dQ; C = 5; }
    else if (nm == "dt") { src = B.dt; C = 1; }
    else if (nm == "vol") { src = B.vol; C = 1; }
    else if (nm == "resev" || nm == "dev" || nm == "dwall") {
        src = nm == "resev" ? B.res_ev : nm == "dev" ? B.dev : B.dwall;
        if (!src) return fail(c, FV3D_EINVAL, "%s exists with Spalart-Allmaras only", name);
        C = 1;
    }
    else if (nm == "lamcI") { src = B.lamc; C = 1; }
    else if (nm == "lamcJ") { src = B.lamc + B.box; C = 1; }
    else if (nm == "lamcK") { src = B.lamc + 2 * B.box; C = 1; }
    else if (nm == "lamvI" || nm == "lamvJ" || nm == "lamvK" || nm == "omega") {
        if (!B.lamv) return fail(c, FV3D_EINVAL, "%s exists for Navier-Stokes only", name);
        src = nm == "omega" ? B.omega : B.lamv + (nm[4] - 'I') * B.box;
        C = 1;
    }
    else if (nm == "fcI" || nm == "fcJ" || nm == "fcK") {
        int d = nm[2] - 'I';
        if (!c->d_faces[d]) return fail(c, FV3D_ESTATE, "face fluxes were not captured (fv3d_set_debug)");
        src = c->d_faces[d]; C = 5; hiI = d == 0; hiJ = d == 1; hiK = d == 2;
    }
    else if (nm == "fvI" || nm == "fvJ" || nm == "fvK" || nm == "gradI" || nm == "gradJ" || nm == "gradK" || nm == "fv") {
        int d = nm == "fv" ? -1 : nm[nm.size() - 1] - 'I';
        const double *p = d < 0 ? c->vdbg.fvc : nm[0] == 'f' ? c->vdbg.fv[d] : c->vdbg.grad[d];
        if (!p) return fail(c, FV3D_ESTATE, "viscous fluxes were not captured (Navier-Stokes and fv3d_set_debug)");
        src = p; C = nm[0] == 'g' ? 12 : 5; hiI = d == 0; hiJ = d == 1; hiK = d == 2;
    }
    if (src) {
        if (fetch(src, (long)C * B.box, h)) return FV3D_EHIP;
        for (int i = 1; i <= B.ni + hiI; ++i)
            for (int j = 1; j <= B.nj + hiJ; ++j)
                for (int k = 1; k <= B.nk + hiK; ++k)
                    for (int q = 0; q < C; ++q) out[cnt++] = h[q * B.box + cidx(B, i, j, k)];
        return cnt;
    }
    static const char *dn[6] = {"dIS", "dIP", "dJS", "dJP", "dKS", "dKP"};
    for (int s = 0; s < 6; ++s) {
        if (nm.compare(0, 3, dn[s]) != 0) continue;
        bool pvq = nm.size() > 3 && nm.substr(3) == "_pv";
        int C2 = pvq ? 6 : 5, c0 = pvq ? 5 : 0;
        if (fetch(B.D[s], 22 * B.Dt[s], h)) return FV3D_EHIP;
        int d = s / 2;
        int ra = d == 0 ? 2 : B.ni, rb = d == 0 ? B.nj : d == 1 ? 2 : B.nj, rc = d == 2 ? 2 : B.nk;
        for (int a = 1; a <= ra; ++a)
            for (int b2 = 1; b2 <= rb; ++b2)
                for (int c2 = 1; c2 <= rc; ++c2) {
                    int layer = d == 0 ? a : d == 1 ? b2 : c2;
                    int ta = d == 0 ? b2 : a, tb = d == 2 ? b2 : c2;
                    for (int q = 0; q < C2; ++q)
                        out[cnt++] = h[((long)(layer - 1) * 11 + c0 + q) * B.Dt[s] + (long)ta * B.Db[s] + tb];
                }
        return cnt;
    }
    return fail(c, FV3D_EINVAL, "unknown field %s", name);
}

}  // extern "C"
