// fv3d_run: runs a reference case directory on the GPU through the C++ class
// layer, in place of the reference's executable (source/Structured_FV3D.cpp:19-214).
//
//   fv3d_run <case_dir> [iterations] [--history FILE] [--eager] [--by-stage] [--device N] [--fp32]
//            [--no-output]
//   fv3d_run <case_dir> --parse-only
//
// The case directory holds config.dat, the grid and boundary-definition files
// it names, RestartSolution.dat when reuse_solution = Y (else the start field
// is CMesh::Initial's, on the coarse level with UseMultigrid), and
// WallDistance.dat when ReuseWallDistance = Y with Spalart-Allmaras.  Without
// an iteration count the config's max_iter is used.  The run stops at the
// first iteration whose density residual is at or below the config's
// convergence tolerance with inlet and outlet mass flows within 0.3 %, as the
// reference's main loop does: it iterates in chunks while the residual is
// above 100 x the tolerance and one iteration at a time below, so the final
// state is that iteration's.  It then writes RestartSolution.dat (and, with
// Spalart-Allmaras, WallDistance.dat) into the case directory unless
// --no-output.  --by-stage drives each iteration through the stage calls of
// the two-grid cycle instead of one fused call; --eager skips graph replay;
// --parse-only reads the case files and prints a JSON summary without
// touching the GPU.  --fp32 runs the LU-SGS operator in FP32 over the FP64
// residual (defect correction; same fixed point).
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "fv3d_classes.h"

namespace {

double checksum(const std::vector<double> &a)
{
    double s = 0.0;
    for (size_t n = 0; n < a.size(); ++n) s += a[n] * (double)(1 + n % 7);
    return s;
}

void print_summary(const fv3d::CaseConfig &cc, const fv3d::BndTable &bnd, const std::vector<std::array<int, 3>> &dims,
                   const std::vector<std::vector<double>> &nodes, const std::vector<fv3d::BlockState> &st)
{
    const fv3d_config &r = cc.raw;
    std::printf("{\"config\": {\"rgas\": %.17g, \"gamma\": %.17g, \"pr_lam\": %.17g, \"pr_tur\": %.17g, "
                "\"pt_in\": %.17g, \"tt_in\": %.17g, \"p_out\": %.17g, \"ma0\": %.17g, \"len_ref\": %.17g, "
                "\"ev_ratio\": %.17g, \"inlet_dir\": [%.17g, %.17g, %.17g], \"cfl\": %.17g, \"c_timestep\": %.17g, "
                "\"relax\": %.17g, \"turbo_type\": %d, \"viscous\": %d, \"recon_var\": %d, \"recon_scheme\": %d, "
                "\"flux_scheme\": %d, \"turbulence\": %d, \"limiter\": %d, \"global_dt\": %d, \"max_iter\": %d, "
                "\"conv_tol\": %.17g, \"reuse_solution\": %d, \"time_scheme\": %d, \"multigrid\": %d, "
                "\"coarse_iter\": %d, \"initial_type\": %d, \"p21ratio\": %.17g, \"outlet_dir\": [%.17g, %.17g, %.17g]},\n",
                r.rgas, r.gamma, r.pr_lam, r.pr_tur, r.pt_in, r.tt_in, r.p_out, r.ma0, r.len_ref, r.ev_ratio,
                r.inlet_dir[0], r.inlet_dir[1], r.inlet_dir[2], r.cfl, r.c_timestep, r.relax, r.turbo_type,
                r.viscous, r.recon_var, r.recon_scheme, r.flux_scheme, r.turbulence, r.limiter, r.global_dt,
                cc.max_iter, cc.conv_tol, (int)cc.reuse_solution, cc.time_scheme, r.multigrid, r.coarse_iter,
                r.initial_type, r.p21ratio, r.outlet_dir[0], r.outlet_dir[1], r.outlet_dir[2]);
    std::printf(" \"patches\": [");
    for (size_t n = 0; n < bnd.patches.size(); ++n) {
        const fv3d::PatchRow &p = bnd.patches[n];
        std::printf("%s[%d, %d, %d, %d, %d, %d, %d, %d, %d, %d]", n ? ", " : "", p.kind, p.block, p.face, p.rotate,
                    p.rng[0], p.rng[1], p.rng[2], p.rng[3], p.rng[4], p.rng[5]);
    }
    std::printf("],\n \"blocks\": [");
    for (size_t b = 0; b < bnd.blocks.size(); ++b) {
        const fv3d::BlockRow &q = bnd.blocks[b];
        std::printf("%s[%d, %d, %d, %d, %d, %.17g]", b ? ", " : "", q.bid, q.IM, q.JM, q.KM, q.nblade, q.rpm);
    }
    std::printf("],\n \"grids\": [");
    for (size_t b = 0; b < dims.size(); ++b)
        std::printf("%s[%d, %d, %d, %.17g]", b ? ", " : "", dims[b][0], dims[b][1], dims[b][2], checksum(nodes[b]));
    std::printf("],\n \"states\": [");
    for (size_t b = 0; b < st.size(); ++b)
    {
        std::printf("%s[%.17g, %.17g", b ? ", " : "", checksum(st[b].pv), checksum(st[b].ener));
        if (!st[b].ev.empty()) std::printf(", %.17g", checksum(st[b].ev));
        std::printf("]");
    }
    std::printf("]}\n");
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <case_dir> [iterations] [--history FILE] [--eager] [--by-stage] [--device N] [--fp32] "
                             "[--no-output] [--parse-only]\n",
                     argv[0]);
        return 2;
    }
    std::string dir = argv[1], history;
    int n_iter = -1, device = 0;
    bool eager = false, parse_only = false, fp32 = false, by_stage = false, no_output = false;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--history") && a + 1 < argc) history = argv[++a];
        else if (!std::strcmp(argv[a], "--eager")) eager = true;
        else if (!std::strcmp(argv[a], "--parse-only")) parse_only = true;
        else if (!std::strcmp(argv[a], "--fp32")) fp32 = true;
        else if (!std::strcmp(argv[a], "--by-stage")) by_stage = true;
        else if (!std::strcmp(argv[a], "--no-output")) no_output = true;
        else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) device = std::atoi(argv[++a]);
        else n_iter = std::atoi(argv[a]);
    }
    try {
        fv3d::CaseConfig cc = fv3d::ReadConfig(dir + "/config.dat");
        fv3d::BndTable bnd = fv3d::ReadBndGrid(dir + "/" + cc.bc_file);
        std::vector<std::array<int, 3>> dims;
        std::vector<std::vector<double>> nodes = fv3d::ReadGrids(dir + "/" + cc.grid_file, &dims);
        std::vector<fv3d::BlockState> st;
        if (cc.reuse_solution)
            st = fv3d::ReadSolution(dir + "/RestartSolution.dat", dims, cc.raw.turbulence == 200);
        if (parse_only) {
            print_summary(cc, bnd, dims, nodes, st);
            return 0;
        }
        cc.raw.device = device;
        if (fp32) cc.raw.precision = FV3D_DEFECT_FP32;
        fv3d::CMesh mesh(cc.raw);
        fv3d::LoadCase(mesh, dir, bnd, nodes, dims, st);
        const bool sa = cc.raw.turbulence == 200;
        if (sa && cc.reuse_wall_distance) {
            // CUserInput::ReadWallDistance in place of the computed distances
            std::vector<std::vector<double>> wd = fv3d::ReadWallDistance(dir + "/WallDistance.dat", dims);
            for (size_t b = 0; b < wd.size(); ++b) mesh.SetWallDistance((int)b + 1, wd[b]);
        }
        fv3d::CSolver solver(mesh, !eager);
        if (n_iter < 0) n_iter = cc.max_iter;
        FILE *hf = history.empty() ? nullptr : std::fopen(history.c_str(), "w");
        if (!history.empty() && !hf) throw fv3d::Error(FV3D_EINVAL, "cannot write " + history);
        std::printf("cells %ld, iterations %d\n", mesh.CellCount(), n_iter);
        // iterations in chunks, one host synchronisation per chunk
        const int chunk = 50;
        int done = 0;
        bool converged = false;
        double last_res = 0.0;
        while (done < n_iter && !converged) {
            int m = n_iter - done < chunk ? n_iter - done : chunk;
            // near the stop test the iterations run one at a time, so the state
            // ends where the reference's loop would break
            const bool near = cc.conv_tol > 0.0 && last_res > 0.0 && last_res <= 100.0 * cc.conv_tol;
            if (near || by_stage) m = 1;
            std::vector<fv3d::HistoryRow> h;
            if (by_stage)
                h.push_back(solver.SolverByStage(done + 1, mesh.Levels() == 2 ? cc.coarse_iter : 0));
            else
                h = solver.Run(m);
            size_t last = 0;
            for (size_t q = 0; q < h.size() && !converged; ++q) {
                const fv3d::HistoryRow &r = h[q];
                if (hf)
                    std::fprintf(hf, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", r[0], r[1], r[2], r[3], r[4], r[5],
                                 r[6]);
                last = q;
                // source/Structured_FV3D.cpp:214: residual and mass-flow balance
                if (r[1] <= cc.conv_tol && std::fabs((r[2] - r[3]) / r[2]) <= 0.003) converged = true;
            }
            done += m;
            const fv3d::HistoryRow &r = h[last];
            last_res = r[1];
            std::printf("iter %6d  res %.6e  mass in %.6e out %.6e  PR %.6f  eff %.6f  CFL %.3f\n", (int)r[0], r[1],
                        r[2], r[3], r[4], r[5], r[6]);
        }
        if (hf) std::fclose(hf);
        if (!no_output) {
            // COutput::OutputRestartSolution and the wall distances, in the case directory
            fv3d::WriteSolution(dir + "/RestartSolution.dat", mesh, dims, sa);
            if (sa && !cc.reuse_wall_distance) fv3d::WriteWallDistance(dir + "/WallDistance.dat", mesh, dims);
        }
        return 0;
    } catch (const fv3d::Error &e) {
        std::fprintf(stderr, "fv3d_run: %s (status %d)\n", e.what(), e.status);
        return 1;
    }
}
