// Test-infrastructure driver for the unmodified reference solver.
//
// This file is part of the oracle, not of the product: it is linked against
// the reference sources compiled by oracle/Makefile (outputs under
// oracle/_ref/) and is only run by the test suite, by the fixture generator
// under tests/golden/ and by bench.py's CPU-baseline leg.
//
// It repeats the set-up sequence of the reference main()
// (source/Structured_FV3D.cpp:36-103) and its iteration loop (142-219), but
// prints the history with 17 significant digits, times the loop with
// std::chrono::steady_clock, and dumps fields as .npy files.
//
// One fix-up is applied before iterating: with equationType == Euler the
// reference never allocates lambda_vI/J/K (source/FluidProps.cpp:767) but
// LU_SGS_Marching reads them unconditionally (source/LU_SGS_Marching.cpp:77,
// 90, 176, 189), which dereferences a null pointer.  The harness allocates
// them zero-filled, which is the value TimeStep implies for Euler
// (source/TimeStep.cpp:54-58).
//
// Usage: fv3d_ref CASE_DIR N_ITER OUT_DIR [--threads N] [--dump-first]
//                 [--no-dump] [--dump-coarse] [--quiet]
// --dump-coarse also writes the coarse level's final fields (final_c<block>_*)
// when the two-grid cycle runs.
#include "Solver.h"
#include "Output.h"
#include "UserInput.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

namespace {

void write_npy(const std::string &path, const std::vector<double> &data, const std::vector<int> &shape)
{
    std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (";
    for (size_t n = 0; n < shape.size(); ++n) {
        hdr += std::to_string(shape[n]);
        hdr += (shape.size() == 1 || n + 1 < shape.size()) ? ", " : "";
    }
    hdr += "), }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr.append(pad, ' ');
    hdr += '\n';
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(3); }
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    unsigned short hl = (unsigned short)hdr.size();
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    std::fwrite(data.data(), sizeof(double), data.size(), f);
    std::fclose(f);
}

// Cell-array dump over index ranges [i0,i1] x [j0,j1] x [k0,k1].
template <class T, int C, class F>
void dump3(const std::string &path, T ***a, int i0, int i1, int j0, int j1, int k0, int k1, F get)
{
    std::vector<double> v;
    v.reserve((size_t)(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1) * C);
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j)
            for (int k = k0; k <= k1; ++k) {
                double tmp[C];
                get(a[i][j][k], tmp);
                for (int c = 0; c < C; ++c) v.push_back(tmp[c]);
            }
    write_npy(path, v, {i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1, C});
}

void cv5(const TConsVars &q, double *o) { o[0] = q.dens; o[1] = q.xmom; o[2] = q.ymom; o[3] = q.zmom; o[4] = q.ener; }
void pv6(const TPrimVars &w, double *o) { o[0] = w.dens; o[1] = w.uvel; o[2] = w.vvel; o[3] = w.wvel; o[4] = w.press; o[5] = w.temp; }
void r1(const REAL &x, double *o) { o[0] = x; }

void dump_block(const std::string &dir, const std::string &tag, CBlocks &B)
{
    CGeometry &G = B.Geometry;
    CFluidProps &F = B.FluidProp;
    int IM = G.IM, JM = G.JM, KM = G.KM, ND = G.NDummy;
    std::string p = dir + "/" + tag;
    dump3<TConsVars, 5>(p + "_cv.npy", F.Cell_cv, 0, IM, 0, JM, 0, KM, cv5);
    dump3<TPrimVars, 6>(p + "_pv.npy", F.Cell_pv, 0, IM, 0, JM, 0, KM, pv6);
    dump3<TConsVars, 5>(p + "_res.npy", B.SpaceDiscr.Res_cv, 1, IM - 1, 1, JM - 1, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_dq.npy", B.TimeDiscr.delt_cv, 1, IM - 1, 1, JM - 1, 1, KM - 1, cv5);
    dump3<REAL, 1>(p + "_dt.npy", B.TimeDiscr.dt, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<TConsVars, 5>(p + "_fcI.npy", B.SpaceDiscr.FaceI_Fc, 1, IM, 1, JM - 1, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_fcJ.npy", B.SpaceDiscr.FaceJ_Fc, 1, IM - 1, 1, JM, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_fcK.npy", B.SpaceDiscr.FaceK_Fc, 1, IM - 1, 1, JM - 1, 1, KM, cv5);
    dump3<REAL, 1>(p + "_lamcI.npy", F.lambda_cI, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<REAL, 1>(p + "_lamcJ.npy", F.lambda_cJ, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<REAL, 1>(p + "_lamcK.npy", F.lambda_cK, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<REAL, 1>(p + "_lamvI.npy", F.lambda_vI, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<REAL, 1>(p + "_lamvJ.npy", F.lambda_vJ, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    dump3<REAL, 1>(p + "_lamvK.npy", F.lambda_vK, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
    // Dummy slabs, layers 1..NDummy, tangential cell ranges 1..N-1.
    dump3<TConsVars, 5>(p + "_dIS_cv.npy", F.DummyIS_cv, 1, ND, 1, JM - 1, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_dIP_cv.npy", F.DummyIP_cv, 1, ND, 1, JM - 1, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_dJS_cv.npy", F.DummyJS_cv, 1, IM - 1, 1, ND, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_dJP_cv.npy", F.DummyJP_cv, 1, IM - 1, 1, ND, 1, KM - 1, cv5);
    dump3<TConsVars, 5>(p + "_dKS_cv.npy", F.DummyKS_cv, 1, IM - 1, 1, JM - 1, 1, ND, cv5);
    dump3<TConsVars, 5>(p + "_dKP_cv.npy", F.DummyKP_cv, 1, IM - 1, 1, JM - 1, 1, ND, cv5);
    dump3<TPrimVars, 6>(p + "_dIS_pv.npy", F.DummyIS_pv, 1, ND, 1, JM - 1, 1, KM - 1, pv6);
    dump3<TPrimVars, 6>(p + "_dIP_pv.npy", F.DummyIP_pv, 1, ND, 1, JM - 1, 1, KM - 1, pv6);
    dump3<TPrimVars, 6>(p + "_dJS_pv.npy", F.DummyJS_pv, 1, IM - 1, 1, ND, 1, KM - 1, pv6);
    dump3<TPrimVars, 6>(p + "_dJP_pv.npy", F.DummyJP_pv, 1, IM - 1, 1, ND, 1, KM - 1, pv6);
    dump3<TPrimVars, 6>(p + "_dKS_pv.npy", F.DummyKS_pv, 1, IM - 1, 1, JM - 1, 1, ND, pv6);
    dump3<TPrimVars, 6>(p + "_dKP_pv.npy", F.DummyKP_pv, 1, IM - 1, 1, JM - 1, 1, ND, pv6);
    if (CFluidProps::equationType == EquationTypes::NavierStokes) {
        dump3<TVisVars, 1>(p + "_mut.npy", F.Cell_tur, 0, IM, 0, JM, 0, KM,
                           [](const TVisVars &t, double *o) { o[0] = t.mue; });
        dump3<REAL, 1>(p + "_omega.npy", F.Omega, 1, IM - 1, 1, JM - 1, 1, KM - 1, r1);
        dump3<TConsVars, 5>(p + "_fv.npy", B.SpaceDiscr.Cell_Fv, 1, IM - 1, 1, JM - 1, 1, KM - 1, cv5);
        dump3<TConsVars, 5>(p + "_fvI.npy", B.SpaceDiscr.FaceI_Fv, 1, IM, 1, JM - 1, 1, KM - 1, cv5);
        dump3<TConsVars, 5>(p + "_fvJ.npy", B.SpaceDiscr.FaceJ_Fv, 1, IM - 1, 1, JM, 1, KM - 1, cv5);
        dump3<TConsVars, 5>(p + "_fvK.npy", B.SpaceDiscr.FaceK_Fv, 1, IM - 1, 1, JM - 1, 1, KM, cv5);
        // face gradients: d/dx, d/dy, d/dz of u v w T
        TPrimVars ***gx[3] = {F.gradx_I, F.gradx_J, F.gradx_K}, ***gy[3] = {F.grady_I, F.grady_J, F.grady_K},
                  ***gz[3] = {F.gradz_I, F.gradz_J, F.gradz_K};
        for (int d = 0; d < 3; ++d) {
            std::vector<double> v;
            for (int i = 1; i <= IM - (d != 0); ++i)
                for (int j = 1; j <= JM - (d != 1); ++j)
                    for (int k = 1; k <= KM - (d != 2); ++k)
                        for (TPrimVars ***g : {gx[d], gy[d], gz[d]}) {
                            const TPrimVars &w = g[i][j][k];
                            v.insert(v.end(), {w.uvel, w.vvel, w.wvel, w.temp});
                        }
            write_npy(p + "_grad" + std::string(1, (char)('I' + d)) + ".npy", v,
                      {IM - (d != 0), JM - (d != 1), KM - (d != 2), 12});
        }
    }
    if (CFluidProps::turbulenceType == TurbulenceTypes::SA) {
        dump3<TEddyVars, 1>(p + "_ev.npy", F.Cell_ev, 0, IM, 0, JM, 0, KM,
                            [](const TEddyVars &t, double *o) { o[0] = t.ev; });
        dump3<TEddyVars, 1>(p + "_resev.npy", B.SpaceDiscr.Res_ev, 1, IM - 1, 1, JM - 1, 1, KM - 1,
                            [](const TEddyVars &t, double *o) { o[0] = t.ev; });
        dump3<TEddyVars, 1>(p + "_dev.npy", B.TimeDiscr.delt_ev, 1, IM - 1, 1, JM - 1, 1, KM - 1,
                            [](const TEddyVars &t, double *o) { o[0] = t.ev; });
        dump3<TEddyVars, 1>(p + "_cfctur.npy", B.SpaceDiscr.Cell_Fc_tur, 1, IM - 1, 1, JM - 1, 1, KM - 1,
                            [](const TEddyVars &t, double *o) { o[0] = t.ev; });
        dump3<TEddyVars, 1>(p + "_cfvtur.npy", B.SpaceDiscr.Cell_Fv_tur, 1, IM - 1, 1, JM - 1, 1, KM - 1,
                            [](const TEddyVars &t, double *o) { o[0] = t.ev; });
        dump3<TCell, 1>(p + "_dwall.npy", G.Cell, 1, IM - 1, 1, JM - 1, 1, KM - 1,
                        [](const TCell &t, double *o) { o[0] = t.dis_Wall; });
        // face gradients of nu~ (x, y, z), as CFluidProps::Gradient stores them
        REAL ***ex[3] = {F.gradx_I_ev, F.gradx_J_ev, F.gradx_K_ev}, ***ey[3] = {F.grady_I_ev, F.grady_J_ev, F.grady_K_ev},
             ***ez[3] = {F.gradz_I_ev, F.gradz_J_ev, F.gradz_K_ev};
        for (int d = 0; d < 3; ++d) {
            std::vector<double> v;
            for (int i = 1; i <= IM - (d != 0); ++i)
                for (int j = 1; j <= JM - (d != 1); ++j)
                    for (int k = 1; k <= KM - (d != 2); ++k)
                        v.insert(v.end(), {(double)ex[d][i][j][k], (double)ey[d][i][j][k], (double)ez[d][i][j][k]});
            write_npy(p + "_gev" + std::string(1, (char)('I' + d)) + ".npy", v,
                      {IM - (d != 0), JM - (d != 1), KM - (d != 2), 3});
        }
    }
}

void zero_alloc_lambda_v(CFluidProps &F)
{
    REAL ***(*alloc)(int, int, int) = [](int IM, int JM, int KM) {
        REAL ***a = new REAL **[IM];
        for (int i = 0; i < IM; ++i) {
            a[i] = new REAL *[JM];
            for (int j = 0; j < JM; ++j) {
                a[i][j] = new REAL[KM];
                for (int k = 0; k < KM; ++k) a[i][j][k] = 0.0;
            }
        }
        return a;
    };
    F.lambda_vI = alloc(F.IM, F.JM, F.KM);
    F.lambda_vJ = alloc(F.IM, F.JM, F.KM);
    F.lambda_vK = alloc(F.IM, F.JM, F.KM);
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s CASE_DIR N_ITER OUT_DIR [--threads N] [--dump-first] [--no-dump] [--quiet]\n", argv[0]);
        return 2;
    }
    std::string case_dir = argv[1];
    int n_iter = std::atoi(argv[2]);
    std::string out_dir = argv[3];
    int threads = -1;
    bool dump_first = false, dump = true, quiet = false, dump_coarse = false;
    for (int a = 4; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--threads") && a + 1 < argc) threads = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--dump-first")) dump_first = true;
        else if (!std::strcmp(argv[a], "--no-dump")) dump = false;
        else if (!std::strcmp(argv[a], "--quiet")) quiet = true;
        else if (!std::strcmp(argv[a], "--dump-coarse")) dump_coarse = true;
    }
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return 3;
    if (out_dir[0] != '/') out_dir = std::string(cwd) + "/" + out_dir;
    if (chdir(case_dir.c_str()) != 0) { std::perror("chdir"); return 3; }

    // Silence the reference's console output unless asked for it.
    if (quiet) { if (!std::freopen("/dev/null", "w", stdout)) return 3; }

    CSolver Solver;
    CUserInput::ReadConfig(Solver);
    CUserInput::NumberOfMultigrid(Solver);
    CUserInput::ReadGrids(Solver);
    CUserInput::ReadBndGrid(Solver);
    CUserInput::NonDimensionalization(Solver);
    if (threads > 0) NT = threads;

    for (int g = 1; g <= Solver.NG; ++g) Solver.Mesh[g].AllocateMemory();
    if (Solver.NG == 2) Solver.Mesh[2].AllocateMemory_Pro_Inter_Vars();
    if (CFluidProps::equationType == EquationTypes::Euler)
        for (int g = 1; g <= Solver.NG; ++g)
            for (int nb = 1; nb <= CMesh::NB; ++nb) zero_alloc_lambda_v(Solver.Mesh[g].Block[nb].FluidProp);
    for (int g = 1; g <= Solver.NG; ++g) Solver.Mesh[g].Geometry();
    if (CFluidProps::turbulenceType == TurbulenceTypes::SA) {
        if (!Solver.reUseWallDistance) Solver.Mesh[1].WallDistanceSA();
        else CUserInput::ReadWallDistance(Solver.Mesh[1]);
    }
    if (!Solver.reUseSolution) {
        if (Solver.NG == 1) Solver.Mesh[1].Initial();
        else {
            Solver.Mesh[2].Initial();
            if (CFluidProps::turbulenceType == TurbulenceTypes::SA) Solver.Mesh[1].Initial();
        }
    } else {
        CUserInput::ReadSolution(Solver.Mesh[1]);
    }

    FILE *hist = std::fopen((out_dir + "/history.txt").c_str(), "w");
    if (!hist) { std::perror("history"); return 3; }
    double loop_s = 0.0;
    int last = 0;
    for (Solver.iter = 1; Solver.iter <= n_iter; Solver.iter++) {
        auto t0 = std::chrono::steady_clock::now();
        Solver.Solver();
        auto t1 = std::chrono::steady_clock::now();
        loop_s += std::chrono::duration<double>(t1 - t0).count();
        CMesh &M = (Solver.NG == 1 || Solver.iter > Solver.Coarse_iter) ? Solver.Mesh[1] : Solver.Mesh[2];
        COutput::OutputGeneralParameters(Solver.MassFlow_in, Solver.MassFlow_out, Solver.Efficiency,
                                         Solver.MassFlow_ratio, Solver.Press_ratio, M);
        std::fprintf(hist, "%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", Solver.iter, (double)Solver.Res.dens,
                     (double)Solver.MassFlow_in, (double)Solver.MassFlow_out, (double)Solver.Press_ratio,
                     (double)Solver.Efficiency, (double)CTimeDiscr::CFL, (double)Solver.Res_ev);
        last = Solver.iter;
        if (dump && dump_first && Solver.iter == 1)
            for (int nb = 1; nb <= CMesh::NB; ++nb)
                dump_block(out_dir, "it1_b" + std::to_string(nb), Solver.Mesh[1].Block[nb]);
        if (Solver.Res.dens <= Solver.ConvTol && Solver.MassFlow_ratio <= 0.003) break;
    }
    std::fclose(hist);
    if (dump)
        for (int nb = 1; nb <= CMesh::NB; ++nb)
            dump_block(out_dir, "final_b" + std::to_string(nb), Solver.Mesh[1].Block[nb]);
    if (dump && dump_coarse && Solver.NG == 2)
        for (int nb = 1; nb <= CMesh::NB; ++nb)
            dump_block(out_dir, "final_c" + std::to_string(nb), Solver.Mesh[2].Block[nb]);
    std::FILE *lv = std::fopen((out_dir + "/levels.txt").c_str(), "w");
    if (lv) { std::fprintf(lv, "%d\n", Solver.NG); std::fclose(lv); }

    long cells = 0;
    for (int nb = 1; nb <= CMesh::NB; ++nb) {
        CGeometry &G = Solver.Mesh[1].Block[nb].Geometry;
        cells += (long)(G.IM - 1) * (G.JM - 1) * (G.KM - 1);
    }
    FILE *tf = std::fopen((out_dir + "/timing.json").c_str(), "w");
    if (tf) {
        std::fprintf(tf, "{\"iterations\": %d, \"loop_seconds\": %.9g, \"cells\": %ld, \"threads\": %d}\n",
                     last, loop_s, cells, NT);
        std::fclose(tf);
    }
    std::fflush(nullptr);
    // The reference's CMesh destructor reads Block after deleting it when
    // NG == 2 (source/Mesh.cpp); leave without running destructors.
    std::_Exit(0);
}
