// C++ host layer of libfv3d, for callers written against the reference's class
// surface: CMesh / CSpaceDiscr / CTimeDiscr / CSolver with the reference's
// method names, plus readers for the reference's case files.  Everything here
// forwards to the C ABI in fv3d.h; no HIP type appears.
//
// Reference surfaces mirrored (paths relative to the reference tree):
//   CMesh        include/Mesh.h:31-90       (Geometry, BoundaryCondition, ResFlux,
//                                             Time_Marching, Res_Calculate)
//   CSpaceDiscr  include/SpaceDiscr.h:62-69 (ConvectiveFlux, ViscousFlux)
//   CFluidProps  include/FluidProps.h (Curl)
//   CTimeDiscr   include/TimeDiscr.h:43-49  (TimeStep)
//   CSolver      include/Solver.h:46-52     (Solver: one iteration with the CFL ramp)
//   CUserInput   source/UserInput.cpp:17-239 (config.dat), 241-286 (grid),
//                336-446 (boundary table), 1118-1159 (restart solution)
//
// Errors: where the reference prints and exits (include/Error.h:11-16), these
// classes throw fv3d::Error carrying the library's status code and message.
#pragma once

#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv3d.h"

namespace fv3d {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &msg) : std::runtime_error(msg), status(status) {}
    int status;
};

inline void check(const fv3d_ctx *c, long rc, const char *what)
{
    if (rc < 0) throw Error((int)rc, std::string(what) + ": " + (c ? fv3d_last_error(c) : "no context"));
}

// ------------------------------------------------------------ case files

// Comment and blank stripping of the reference's line reader
// (include/fstringIO.h:15-42): text after '#' is dropped, then spaces.
inline std::string strip_line(std::string s)
{
    size_t h = s.find('#');
    if (h != std::string::npos) s.erase(h);
    while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

inline std::vector<std::string> read_lines(const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot open " + path);
    std::vector<std::string> out;
    std::string s;
    while (std::getline(f, s)) out.push_back(strip_line(s));
    return out;
}

// config.dat, read by line number as CUserInput::ReadConfig does
// (source/UserInput.cpp:17-239).  Values are kept as written; the library
// applies the reference's float parsing and non-dimensionalisation.
struct CaseConfig {
    fv3d_config raw{};
    std::string grid_file, bc_file;
    int max_iter = 0, coarse_iter = 0, n_threads = 1;
    double conv_tol = 0.0;
    bool multigrid = false, reuse_solution = false, reuse_wall_distance = false;
    int time_scheme = 0;
};

inline CaseConfig ReadConfig(const std::string &path)
{
    std::vector<std::string> L = read_lines(path);
    auto v = [&](int n) -> const std::string & {
        if (n < 1 || n > (int)L.size()) throw Error(FV3D_EINVAL, path + ": missing line " + std::to_string(n));
        return L[n - 1];
    };
    auto num = [&](int n) { return std::strtod(v(n).c_str(), nullptr); };
    auto inum = [&](int n) { return (int)std::strtol(v(n).c_str(), nullptr, 10); };
    auto flag = [&](int n) { return !v(n).empty() && (v(n)[0] == 'Y' || v(n)[0] == 'y'); };
    CaseConfig c;
    fv3d_config &r = c.raw;
    c.grid_file = v(2);
    c.bc_file = v(5);
    const char tt = v(9).empty() ? 'G' : v(9)[0];
    r.turbo_type = (tt == 'A' || tt == 'a') ? 0 : (tt == 'C' || tt == 'c') ? 1 : 2;
    r.viscous = !(v(10).size() && (v(10)[0] == 'E' || v(10)[0] == 'e'));
    r.initial_type = (v(11).size() && (v(11)[0] == 'U' || v(11)[0] == 'u')) ? 0 : 1;
    r.rgas = num(12); r.gamma = num(13); r.pr_lam = num(14); r.pr_tur = num(15);
    r.pt_in = num(19); r.tt_in = num(20); r.p_out = num(21);
    r.p21ratio = num(22);
    r.ma0 = num(23); r.len_ref = num(24); r.ev_ratio = num(25);
    {
        std::istringstream s(v(26));
        s >> r.inlet_dir[0] >> r.inlet_dir[1] >> r.inlet_dir[2];
    }
    {
        std::istringstream s(v(27));
        s >> r.outlet_dir[0] >> r.outlet_dir[1] >> r.outlet_dir[2];
    }
    // UseFlowMassImposed (source/UserInput.cpp:79-80): the outlet pressure
    // rescaled towards a target mass flow during the run
    // (source/Structured_FV3D.cpp:186-195) is not in this build
    if (flag(28))
        throw Error(FV3D_EINVAL, path + ": UseFlowMassImposed = Y (outlet pressure driven to a mass flow) is not "
                                        "supported; set it to N");
    c.max_iter = inum(33); c.coarse_iter = inum(34);
    c.conv_tol = num(36); c.n_threads = inum(37);
    c.multigrid = flag(38); c.reuse_solution = flag(39); c.reuse_wall_distance = flag(40);
    r.multigrid = c.multigrid;
    r.coarse_iter = c.coarse_iter;
    r.cfl = num(44); r.c_timestep = num(45); r.relax = num(46);
    r.limiter = flag(51);
    r.global_dt = !(v(52).size() && (v(52)[0] == 'L' || v(52)[0] == 'l'));
    r.recon_var = inum(53); r.recon_scheme = inum(54); r.flux_scheme = inum(55);
    c.time_scheme = inum(56);
    r.turbulence = inum(57);
    r.device = 0;
    r.precision = FV3D_PLAIN_FP64;
    return c;
}

// Boundary-definition file (source/UserInput.cpp:336-446): patch tables then
// the block table.  kind: 0 inlet, 1 outlet, 2 wall, 3 periodic, 4 interior,
// 5 mixing plane.
struct PatchRow {
    int kind = 0, block = 0, face = 0, rotate = 0;
    int rng[6] = {0, 0, 0, 0, 0, 0};
    int orient = 0, block2 = 0, face2 = 0;
    int rng2[6] = {0, 0, 0, 0, 0, 0};
};
struct BlockRow {
    int bid = 0, IM = 0, JM = 0, KM = 0, nblade = 1;
    double rpm = 0.0;
};
struct BndTable {
    std::vector<PatchRow> patches;
    std::vector<BlockRow> blocks;
};

inline BndTable ReadBndGrid(const std::string &path)
{
    std::vector<std::string> L = read_lines(path);
    size_t pos = 0;
    auto next = [&]() -> const std::string & {
        if (pos >= L.size()) throw Error(FV3D_EINVAL, path + ": unexpected end");
        return L[pos++];
    };
    auto ints = [](const std::string &s) {
        std::istringstream is(s);
        std::vector<int> t;
        double x;
        while (is >> x) t.push_back((int)x);
        return t;
    };
    BndTable T;
    for (int i = 0; i < 4; ++i) next();
    for (int kind = 0; kind < 6; ++kind) {
        if (kind > 0)
            for (int i = 0; i < 3; ++i) next();
        int n = (int)std::strtol(next().c_str(), nullptr, 10);
        next();
        for (int r = 0; r < n; ++r) {
            std::vector<int> t = ints(next());
            PatchRow p;
            p.kind = kind;
            if (kind <= 1 && t.size() >= 9) {
                p.block = t[1]; p.face = t[2];
                for (int q = 0; q < 6; ++q) p.rng[q] = t[3 + q];
            } else if (kind == 2 && t.size() >= 10) {
                p.block = t[1]; p.face = t[2]; p.rotate = t[3];
                for (int q = 0; q < 6; ++q) p.rng[q] = t[4 + q];
            } else if (t.size() >= 18) {
                p.orient = t[1]; p.block = t[2]; p.face = t[3];
                for (int q = 0; q < 6; ++q) p.rng[q] = t[4 + q];
                p.block2 = t[10]; p.face2 = t[11];
                for (int q = 0; q < 6; ++q) p.rng2[q] = t[12 + q];
            } else {
                throw Error(FV3D_EINVAL, path + ": short patch row");
            }
            T.patches.push_back(p);
        }
    }
    for (int i = 0; i < 3; ++i) next();
    int nb = (int)std::strtol(next().c_str(), nullptr, 10);
    next();
    for (int b = 0; b < nb; ++b) {
        std::istringstream is(next());
        BlockRow r;
        is >> r.bid >> r.IM >> r.JM >> r.KM >> r.nblade >> r.rpm;
        T.blocks.push_back(r);
    }
    return T;
}

// Grid file (source/UserInput.cpp:241-286): per block IM JM KM then the nodes
// with i fastest.  Returned per block in the fv3d_block_add order (k fastest).
inline std::vector<std::vector<double>> ReadGrids(const std::string &path, std::vector<std::array<int, 3>> *dims)
{
    std::ifstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot open " + path);
    std::string s;
    std::getline(f, s);
    int nb = (int)std::strtol(strip_line(s).c_str(), nullptr, 10);
    std::vector<std::vector<double>> out;
    for (int b = 0; b < nb; ++b) {
        std::getline(f, s);
        std::getline(f, s);
        int IM = 0, JM = 0, KM = 0;
        std::istringstream(strip_line(s)) >> IM >> JM >> KM;
        std::vector<double> xyz((size_t)IM * JM * KM * 3);
        for (int k = 0; k < KM; ++k)
            for (int j = 0; j < JM; ++j)
                for (int i = 0; i < IM; ++i) {
                    double x, y, z;
                    if (!(f >> x >> y >> z)) throw Error(FV3D_EINVAL, path + ": short node list");
                    size_t n = (((size_t)i * JM + j) * KM + k) * 3;
                    xyz[n] = x; xyz[n + 1] = y; xyz[n + 2] = z;
                }
        std::getline(f, s);
        out.push_back(std::move(xyz));
        if (dims) dims->push_back({IM, JM, KM});
    }
    return out;
}

// RestartSolution.dat (source/UserInput.cpp:1118-1159): rows of
// "block i j k p u v w T rho rhoE", with Spalart-Allmaras (with_nut) one more
// column nu~.  Returned per block as the fv3d_state_upload pair pv
// (ni,nj,nk,6 = rho,u,v,w,p,T) and ener (ni,nj,nk), and ev (ni,nj,nk) for
// fv3d_nut_upload (empty without with_nut).
// The five conservative components of a residual norm, as the reference's
// TConsVars (include/defs.h:58-64) names them.
struct TConsVars {
    double dens = 0.0, xmom = 0.0, ymom = 0.0, zmom = 0.0, ener = 0.0;
};

struct BlockState {
    std::vector<double> pv, ener, ev;
};

inline std::vector<BlockState> ReadSolution(const std::string &path, const std::vector<std::array<int, 3>> &dims,
                                            bool with_nut = false)
{
    std::ifstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot open " + path);
    std::vector<BlockState> out(dims.size());
    for (size_t b = 0; b < dims.size(); ++b) {
        size_t n = (size_t)(dims[b][0] - 1) * (dims[b][1] - 1) * (dims[b][2] - 1);
        out[b].pv.assign(n * 6, 0.0);
        out[b].ener.assign(n, 0.0);
        if (with_nut) out[b].ev.assign(n, 0.0);
    }
    std::string s;
    while (std::getline(f, s)) {
        std::istringstream is(s);
        int b, i, j, k;
        double p, u, v, w, T, rho, rhoE, ev = 0.0;
        if (!(is >> b >> i >> j >> k >> p >> u >> v >> w >> T >> rho >> rhoE)) continue;
        if (with_nut && !(is >> ev)) throw Error(FV3D_EINVAL, path + ": a row without the nu~ column");
        if (b < 1 || b > (int)dims.size()) throw Error(FV3D_EINVAL, path + ": block id out of range");
        const int nj = dims[b - 1][1] - 1, nk = dims[b - 1][2] - 1, ni = dims[b - 1][0] - 1;
        if (i < 1 || i > ni || j < 1 || j > nj || k < 1 || k > nk) throw Error(FV3D_EINVAL, path + ": cell out of range");
        size_t x = ((size_t)(i - 1) * nj + (j - 1)) * nk + (k - 1);
        double *q = &out[b - 1].pv[x * 6];
        q[0] = rho; q[1] = u; q[2] = v; q[3] = w; q[4] = p; q[5] = T;
        out[b - 1].ener[x] = rhoE;
        if (with_nut) out[b - 1].ev[x] = ev;
    }
    return out;
}

// ------------------------------------------------------------ classes

class CSpaceDiscr;
class CTimeDiscr;

// All blocks of one grid level on one GPU (one libfv3d context).
class CMesh {
public:
    explicit CMesh(const fv3d_config &cfg)
    {
        check(nullptr, fv3d_ctx_create(&cfg, &ctx_), "fv3d_ctx_create");
    }
    ~CMesh()
    {
        if (ctx_) fv3d_ctx_destroy(ctx_);
    }
    CMesh(const CMesh &) = delete;
    CMesh &operator=(const CMesh &) = delete;

    // Block and patch tables (CUserInput::ReadGrids / ReadBndGrid).
    int AddBlock(int IM, int JM, int KM, const std::vector<double> &nodes, int nblade, double rpm)
    {
        if (nodes.size() != (size_t)IM * JM * KM * 3) throw Error(FV3D_EINVAL, "AddBlock: node array size");
        int id = fv3d_block_add(ctx_, IM, JM, KM, nodes.data(), nblade, rpm);
        check(ctx_, id, "fv3d_block_add");
        return id;
    }
    void AddPatch(int kind, int block, int face, const int rng[6], int rotate)
    {
        check(ctx_, fv3d_patch_add(ctx_, kind, block, face, rng, rotate), "fv3d_patch_add");
    }
    // Interior (kind 0) or periodic (kind 1) patch (CUserInput::ReadBndGrid).
    void AddConnection(int kind, int orient, int block1, int face1, const int rng1[6], int block2, int face2,
                       const int rng2[6])
    {
        check(ctx_, fv3d_connection_add(ctx_, kind, orient, block1, face1, rng1, block2, face2, rng2),
              "fv3d_connection_add");
    }
    // Mixing plane (CUserInput::ReadBndGrid's MixingPlane rows,
    // source/UserInput.cpp:420-432): block1's K-high face against block2's
    // K-low face (source/BndCondMixingPlane.cpp:287-522).
    void AddMixingPlane(int block1, const int rng1[6], int block2, const int rng2[6])
    {
        check(ctx_, fv3d_mixing_plane_add(ctx_, block1, rng1, block2, rng2), "fv3d_mixing_plane_add");
    }
    // CMesh::Geometry: metrics and device set-up, and with UseMultigrid the
    // coarse level when the patch ranges allow it.
    void Geometry() { check(ctx_, fv3d_setup(ctx_), "fv3d_setup"); }
    int Levels() const { return fv3d_levels(ctx_); }
    // CUserInput::ReadSolution: one block's restart state.
    void Initial(int block, const BlockState &st)
    {
        check(ctx_, fv3d_state_upload(ctx_, block, st.pv.data(), st.ener.data()), "fv3d_state_upload");
        if (!st.ev.empty()) check(ctx_, fv3d_nut_upload(ctx_, block, st.ev.data()), "fv3d_nut_upload");
    }
    // CMesh::Initial: the uniform or interpolated start field (on the coarse
    // level when two levels run).
    void Initial() { check(ctx_, fv3d_initial(ctx_), "fv3d_initial"); }

    void BoundaryCondition() { check(ctx_, fv3d_bc_apply(ctx_), "fv3d_bc_apply"); }
    void ResFlux() { check(ctx_, fv3d_resflux(ctx_), "fv3d_resflux"); }
    // LU-SGS is the time scheme of this path (CMesh::LU_SGS_Marching).
    void Time_Marching() { check(ctx_, fv3d_lusgs(ctx_), "fv3d_lusgs"); }
    // res: dens, xmom, ymom, zmom, ener, ev (source/ComputeAllBlockVars.cpp:318-372)
    std::array<double, 6> Res_Calculate()
    {
        std::array<double, 6> r{};
        check(ctx_, fv3d_res_calculate(ctx_, r.data()), "fv3d_res_calculate");
        return r;
    }
    // The reference's signature: Res_Calculate(TConsVars &Res, REAL &Res_ev)
    // (include/Mesh.h:34).
    void Res_Calculate(TConsVars &res, double &res_ev)
    {
        const std::array<double, 6> r = Res_Calculate();
        res.dens = r[0]; res.xmom = r[1]; res.ymom = r[2]; res.zmom = r[3]; res.ener = r[4];
        res_ev = r[5];
    }
    // CFluidProps::PrimitiveVars and Limit_Vars of every block
    // (source/PrimitiveVars.cpp:7-29, source/Limit_FlowVars.cpp:8-119; the
    // latter does nothing with the limiter off in config.dat).
    void PrimitiveVars() { check(ctx_, fv3d_primitive_vars(ctx_), "fv3d_primitive_vars"); }
    void Limit_Vars() { check(ctx_, fv3d_limit_vars(ctx_), "fv3d_limit_vars"); }
    // CMesh::WallDistanceSA (source/WallDistanceSA.cpp:9-132): computed on the
    // device at Geometry(); the distances of one block's cells.
    std::vector<double> WallDistanceSA(int block, size_t ncell) { return Field(block, "dwall", ncell); }
    // CUserInput::ReadWallDistance: distances of a block read from WallDistance.dat.
    void SetWallDistance(int block, const std::vector<double> &d)
    {
        check(ctx_, fv3d_wall_distance_upload(ctx_, block, d.data()), "fv3d_wall_distance_upload");
    }
    // Host mirror of a device field, for output (names as fv3d_field_download).
    std::vector<double> Field(int block, const char *name, size_t count)
    {
        std::vector<double> out(count);
        long n = fv3d_field_download(ctx_, block, name, out.data());
        check(ctx_, n, "fv3d_field_download");
        out.resize((size_t)n);
        return out;
    }
    long CellCount() const { return fv3d_cell_count(ctx_); }
    fv3d_ctx *ctx() const { return ctx_; }

private:
    fv3d_ctx *ctx_ = nullptr;
};

class CSpaceDiscr {
public:
    explicit CSpaceDiscr(CMesh &m) : m_(m) {}
    // CSpaceDiscr::SA_Model / CBlocks::BL_Model: the turbulence stage of ResFlux
    // (after BoundaryCondition).  CSpaceDiscr::WallFlux runs inside
    // ConvectiveFlux (wall faces take the face-averaged flux there).
    void SA_Model() { check(m_.ctx(), fv3d_turbulence(m_.ctx()), "fv3d_turbulence"); }
    // CMesh::ConvectiveFlux incl. WallFlux, the source term and Res = Fc.
    void ConvectiveFlux() { check(m_.ctx(), fv3d_convective(m_.ctx()), "fv3d_convective"); }
    // CMesh::ViscousFlux with the face gradients: Res -= Fv (after ConvectiveFlux;
    // nothing for Euler).
    void ViscousFlux() { check(m_.ctx(), fv3d_viscous(m_.ctx()), "fv3d_viscous"); }

private:
    CMesh &m_;
};

class CFluidProps {
public:
    explicit CFluidProps(CMesh &m) : m_(m) {}
    // CFluidProps::Curl: vorticity magnitude, field "omega" (Navier-Stokes).
    void Curl() { check(m_.ctx(), fv3d_curl(m_.ctx()), "fv3d_curl"); }

private:
    CMesh &m_;
};

class CTimeDiscr {
public:
    explicit CTimeDiscr(CMesh &m) : m_(m) {}
    // Spectral radii and the local time step (fused on the device).
    void TimeStep()
    {
        check(m_.ctx(), fv3d_spectral(m_.ctx()), "fv3d_spectral");
        check(m_.ctx(), fv3d_timestep(m_.ctx()), "fv3d_timestep");
    }

private:
    CMesh &m_;
};

// One history row: iter res_dens mass_in mass_out PR eff CFL.
using HistoryRow = std::array<double, 7>;

class CSolver {
public:
    explicit CSolver(CMesh &m, bool graph = true) : m_(m), graph_(graph) {}
    // CSolver::Solver: one iteration with the CFL ramp (single grid, or the
    // coarse start-up phase / two-grid cycle).
    HistoryRow Solver()
    {
        HistoryRow r{};
        int rc = graph_ ? fv3d_iterate_graph(m_.ctx(), 1, r.data()) : fv3d_iterate(m_.ctx(), 1, r.data());
        check(m_.ctx(), rc, "fv3d_iterate");
        return r;
    }
    // The stages of one iteration of CSolver::Solver (source/Solver.cpp:61-103)
    // as the reference's private methods: the two-grid cycle's Prolong
    // (restriction with the forcing), the forced coarse Time_Marching,
    // ComputeDeltConsVars + Interpolate (the correction), and the state
    // Interpolate at Coarse_iter.
    void Prolong() { check(m_.ctx(), fv3d_mg_restrict(m_.ctx()), "fv3d_mg_restrict"); }
    void CoarseTime_Marching(bool forced)
    {
        check(m_.ctx(), fv3d_mg_coarse_lusgs(m_.ctx(), forced ? 1 : 0), "fv3d_mg_coarse_lusgs");
    }
    void Interpolate() { check(m_.ctx(), fv3d_mg_correct(m_.ctx()), "fv3d_mg_correct"); }
    void InterpolateState() { check(m_.ctx(), fv3d_mg_interpolate_state(m_.ctx()), "fv3d_mg_interpolate_state"); }
    // One iteration driven stage by stage (single grid or two-grid cycle);
    // the same bits as Solver().  iter is the iteration number (1-based).
    HistoryRow SolverByStage(long iter, int coarse_iter)
    {
        fv3d_ctx *c = m_.ctx();
        if (m_.Levels() == 1) {
            m_.Time_Marching();
        } else if (iter <= coarse_iter) {
            CoarseTime_Marching(false);
            if (iter == coarse_iter) InterpolateState();
        } else {
            m_.Time_Marching();
            m_.ResFlux();
            Prolong();
            CoarseTime_Marching(true);
            Interpolate();
        }
        HistoryRow r{};
        check(c, fv3d_iteration_end(c, r.data()), "fv3d_iteration_end");
        return r;
    }
    // n iterations in one call (one synchronisation at the end).
    std::vector<HistoryRow> Run(int n)
    {
        std::vector<HistoryRow> h((size_t)n);
        int rc = graph_ ? fv3d_iterate_graph(m_.ctx(), n, h[0].data()) : fv3d_iterate(m_.ctx(), n, h[0].data());
        check(m_.ctx(), rc, "fv3d_iterate");
        return h;
    }

private:
    CMesh &m_;
    bool graph_;
};

// Loads a reference case directory (config.dat, grid, boundary table and,
// when reuse_solution is set, RestartSolution.dat) into a mesh; without
// restart states the start field is CMesh::Initial's.
inline void LoadCase(CMesh &mesh, const std::string &dir, const BndTable &bnd,
                     const std::vector<std::vector<double>> &nodes, const std::vector<std::array<int, 3>> &dims,
                     const std::vector<BlockState> &states)
{
    (void)dir;
    for (size_t b = 0; b < nodes.size(); ++b) {
        int nblade = b < bnd.blocks.size() ? bnd.blocks[b].nblade : 1;
        double rpm = b < bnd.blocks.size() ? bnd.blocks[b].rpm : 0.0;
        mesh.AddBlock(dims[b][0], dims[b][1], dims[b][2], nodes[b], nblade, rpm);
    }
    for (const PatchRow &p : bnd.patches) {
        if (p.kind == 5)
            mesh.AddMixingPlane(p.block, p.rng, p.block2, p.rng2);
        else if (p.kind == 3 || p.kind == 4)
            mesh.AddConnection(p.kind == 3 ? 1 : 0, p.orient, p.block, p.face, p.rng, p.block2, p.face2, p.rng2);
        else
            mesh.AddPatch(p.kind, p.block, p.face, p.rng, p.rotate);
    }
    mesh.Geometry();
    if (states.empty())
        mesh.Initial();
    for (size_t b = 0; b < states.size(); ++b) mesh.Initial((int)b + 1, states[b]);
}

// COutput::OutputRestartSolution (source/Ouput.cpp:227-259): rows of
// "block i j k p u v w T rho rhoE" (and nu~ with Spalart-Allmaras), k outermost,
// i innermost, in the solver's non-dimensional units; written with 17
// significant digits (the reference's stream default is 6), so ReadSolution
// gives back the same bits.
inline void WriteSolution(const std::string &path, CMesh &mesh, const std::vector<std::array<int, 3>> &dims,
                          bool with_nut)
{
    std::ofstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot write " + path);
    f.precision(17);
    for (size_t b = 0; b < dims.size(); ++b) {
        const int IM = dims[b][0], JM = dims[b][1], KM = dims[b][2];
        const size_t box = (size_t)(IM + 1) * (JM + 1) * (KM + 1);
        std::vector<double> pv = mesh.Field((int)b + 1, "pv", 6 * box), cv = mesh.Field((int)b + 1, "cv", 5 * box);
        std::vector<double> ev;
        if (with_nut) ev = mesh.Field((int)b + 1, "ev", box);
        for (int k = 1; k <= KM - 1; ++k)
            for (int j = 1; j <= JM - 1; ++j)
                for (int i = 1; i <= IM - 1; ++i) {
                    const size_t x = ((size_t)i * (JM + 1) + j) * (KM + 1) + k;
                    const double *w = &pv[6 * x];
                    f << b + 1 << " " << i << " " << j << " " << k << " " << w[4] << " " << w[1] << " " << w[2] << " "
                      << w[3] << " " << w[5] << " " << w[0] << " " << cv[5 * x + 4];
                    if (with_nut) f << " " << ev[x];
                    f << "\n";
                }
    }
    if (!f) throw Error(FV3D_EINVAL, "error writing " + path);
}

// WallDistance.dat (source/WallDistanceSA.cpp:162-185, read back by
// CUserInput::ReadWallDistance, source/UserInput.cpp:1094-1116): per block its
// number, its cell counts, then "i j k d" rows, k outermost.  Distances are
// returned per block with k fastest (the fv3d_wall_distance_upload order).
inline std::vector<std::vector<double>> ReadWallDistance(const std::string &path,
                                                         const std::vector<std::array<int, 3>> &dims)
{
    std::ifstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot open " + path);
    std::vector<std::vector<double>> out(dims.size());
    for (size_t b = 0; b < dims.size(); ++b) {
        const int ni = dims[b][0] - 1, nj = dims[b][1] - 1, nk = dims[b][2] - 1;
        int id, a, c, d;
        if (!(f >> id >> a >> c >> d)) throw Error(FV3D_EINVAL, path + ": missing block header");
        if (a != ni || c != nj || d != nk) throw Error(FV3D_EINVAL, path + ": block cell counts differ from the grid");
        out[b].assign((size_t)ni * nj * nk, 0.0);
        for (int k = 1; k <= nk; ++k)
            for (int j = 1; j <= nj; ++j)
                for (int i = 1; i <= ni; ++i) {
                    int ii, jj, kk;
                    double v;
                    if (!(f >> ii >> jj >> kk >> v)) throw Error(FV3D_EINVAL, path + ": short wall-distance list");
                    out[b][((size_t)(i - 1) * nj + (j - 1)) * nk + (k - 1)] = v;
                }
    }
    return out;
}

inline void WriteWallDistance(const std::string &path, CMesh &mesh, const std::vector<std::array<int, 3>> &dims)
{
    std::ofstream f(path);
    if (!f) throw Error(FV3D_EINVAL, "cannot write " + path);
    f.precision(17);
    for (size_t b = 0; b < dims.size(); ++b) {
        const int ni = dims[b][0] - 1, nj = dims[b][1] - 1, nk = dims[b][2] - 1;
        std::vector<double> d = mesh.Field((int)b + 1, "dwall", (size_t)ni * nj * nk);
        f << b + 1 << "\n" << ni << " " << nj << " " << nk << "\n";
        for (int k = 1; k <= nk; ++k)
            for (int j = 1; j <= nj; ++j)
                for (int i = 1; i <= ni; ++i)
                    f << i << " " << j << " " << k << " " << d[((size_t)(i - 1) * nj + (j - 1)) * nk + (k - 1)] << "\n";
    }
    if (!f) throw Error(FV3D_EINVAL, "error writing " + path);
}

}  // namespace fv3d
