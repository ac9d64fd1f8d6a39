/* CPU restatement of the reference solver's per-iteration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg.  The product never
 * links, loads or calls it.
 *
 * It restates, in plain C with the reference's array extents and expression
 * order, the functions of the reference's LU-SGS iteration: ghost fill, spectral
 * radii, local time step, MUSCL/NND reconstruction with the Roe and
 * Steger-Warming fluxes, wall flux, the laminar Navier-Stokes terms (metric
 * coefficients, face gradients, curl, Sutherland viscosity, viscous flux and
 * the viscous spectral radii), residual assembly, LU-SGS sweeps, update,
 * primitive recovery, variable limiting, residual norm and the section
 * performance integrals.  Each function cites the reference file:line it
 * follows.  Indices are the reference's: cells 1..IM-1, ghost cells 0 and IM,
 * faces 1..IM, Dummy layers 1..NDummy.
 */
#ifndef FV3D_ORACLE_H
#define FV3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* raw config.dat values (parsed as float by the reference) */
    double rgas, gamma, pr_lam, pr_tur;
    double pt_in, tt_in, p_out, ma0, len_ref, ev_ratio;
    double inlet_dir[3];
    double cfl, c_timestep, relax;
    int turbo_type;      /* 0 axial, 1 centrifugal, 2 general */
    int viscous;         /* 0 Euler, 1 Navier-Stokes */
    int recon_var;       /* 100 conservative, 200 primitive */
    int recon_scheme;    /* 100 NND, 200 MUSCL2u, 300 MUSCL2c, 400 MUSCL2o, 500 MUSCL3 */
    int flux_scheme;     /* 200 Steger-Warming, 300 Roe */
    int turbulence;      /* 0 none, 100 Baldwin-Lomax, 200 Spalart-Allmaras */
    int limiter;         /* limit flow variables */
    int global_dt;       /* 0 local time step */
    int nthreads;
    /* two-grid cycle and start field (source/UserInput.cpp:50-58, 86-92) */
    int multigrid;       /* UseMultigrid */
    int coarse_iter;     /* Coarse_iter */
    int initial_type;    /* 0 uniform, 1 interpolate (CMesh::Initial) */
    double p21ratio;     /* outlet/inlet static pressure ratio of the interpolated start */
    double outlet_dir[3];
} orc_config;

typedef struct {
    double gamma, rgas, pr_lam, pr_tur;
    double pt_in, tt_in, p_out;      /* non-dimensional */
    double temp_ref, press_ref, vel_ref, mue_ref, dens_ref, re_ref;
    double ev_ratio;
    double inlet_dir[3];
} orc_nondim;

typedef struct orc_ctx orc_ctx;

/* Non-dimensionalisation of source/UserInput.cpp:1070-1092. */
void orc_nondimensionalize(const orc_config *cfg, orc_nondim *nd);

orc_ctx *orc_create(const orc_config *cfg);
void orc_destroy(orc_ctx *c);

/* nodes: IM*JM*KM*3 doubles, node (i,j,k) (1-based) at ((i-1)*JM+(j-1))*KM+(k-1).
 * rpm as written in the block table; w_rot = 2*PI*rpm/60 non-dimensionalised. */
int orc_block_add(orc_ctx *c, int IM, int JM, int KM, const double *nodes, int nblade, double rpm);

/* kind: 0 inlet, 1 outlet, 2 wall.  rng = is ie js je ks ke as in the
 * boundary-definition file. */
int orc_patch_add(orc_ctx *c, int kind, int block, int face, const int *rng, int rotate);

/* Two-sided patch (source/UserInput.cpp:386-417): kind 0 interior, 1 periodic;
 * orient 1..4 as in CMesh::Trans; ranges as in orc_patch_add, start <= end. */
int orc_connection_add(orc_ctx *c, int kind, int orient, int block1, int face1, const int *rng1, int block2, int face2,
                       const int *rng2);

/* Metrics of every block.  With cfg.multigrid it also decides the number of
 * grid levels as CUserInput::NumberOfMultigrid does (source/UserInput.cpp:
 * 799-1067, at most two) and builds the coarse level: nodes (2i-1, 2j-1,
 * 2k-1) (241-334), patch ranges halved (446-797). */
/* Mixing plane (source/UserInput.cpp:420-432): block 1's K-high face (range
 * rng1, cells is..ie, js..je, ks = ke = KM) against block 2's K-low face
 * (rng2, ks = ke = 1); both sides share the i range of rng1. */
int orc_mixing_plane_add(orc_ctx *c, int block1, const int *rng1, int block2, const int *rng2);
int orc_geometry(orc_ctx *c);

/* Grid levels in use (1 or 2) and the coarse level's context (NULL for one). */
int orc_levels(orc_ctx *c);
orc_ctx *orc_coarse(orc_ctx *c);

/* CMesh::Initial (source/ComputeAllBlockVars.cpp:38-48, source/Initialization.cpp)
 * on the level the reference starts from without a restart file: the coarse
 * level when two levels run, else the only one (source/Structured_FV3D.cpp:86-96). */
int orc_initial(orc_ctx *c);

/* Restart semantics of source/UserInput.cpp:1118-1159: pv is (ni,nj,nk,6)
 * rho,u,v,w,p,T and ener (ni,nj,nk), cell (i,j,k) at ((i-1)*nj+(j-1))*nk+(k-1). */
int orc_set_state(orc_ctx *c, int block, const double *pv, const double *ener);
/* nu~ of a block from the restart file (Spalart-Allmaras), same cell order */
int orc_set_nut(orc_ctx *c, int block, const double *ev);

int orc_bc(orc_ctx *c);
int orc_spectral(orc_ctx *c);
int orc_timestep(orc_ctx *c);
int orc_convective(orc_ctx *c);
int orc_resflux(orc_ctx *c);
int orc_lusgs(orc_ctx *c);
int orc_res_calculate(orc_ctx *c, double *res6);
int orc_perf(orc_ctx *c, double *out5);   /* mass_in, mass_out, PR, eff, mass_ratio */
double orc_cfl(orc_ctx *c);

/* n iterations of CSolver::Solver (source/Solver.cpp:47-135: single grid, or
 * the coarse start-up phase and two-grid cycle of reuse_solution = N) and the
 * performance parameters of the level the reference reports
 * (source/Structured_FV3D.cpp:148-158).  hist receives n rows of: iter
 * res_dens mass_in mass_out PR eff CFL. */
int orc_iterate(orc_ctx *c, int n, double *hist);

/* Copies a field out.  Names: cv pv mut (full (IM+1)(JM+1)(KM+1)xC incl.
 * ghost cells), res dq dt lamcI lamcJ lamcK lamvI lamvJ lamvK omega fv
 * (interior (ni,nj,nk)xC), dIS_cv dIP_cv dJS_cv dJP_cv dKS_cv dKP_cv and the
 * _pv variants (layers 1..NDummy), vol, fcI fcJ fcK (convective face fluxes),
 * fvI fvJ fvK (viscous face fluxes), gradI gradJ gradK (face gradients: d/dx,
 * d/dy, d/dz of u v w T), metI metJ metK metC (face and cell transformation
 * coefficients, 9 each), SI SJ SK (face vectors); Spalart-Allmaras: ev (nu~,
 * full box), resev dev dwall (interior), gevI gevJ gevK (face gradients of
 * nu~, x y z). */
long orc_get(orc_ctx *c, int block, const char *name, double *out);

/* Counters of the repair branches since the last call (then cleared):
 * [0] face states replaced by the first-order fallback (density or pressure
 * not positive after reconstruction), [1] cells Limit_Vars replaced, [2] of
 * those, cells with no admissible neighbour (the last-neighbour clamp),
 * [3] nu~ values Limit_SA replaced (above 1000), [4] of those, with no
 * admissible neighbour, [5] blocks whose Flag_low_dt Limit_Vars set. */
#define ORC_NSTATS 6
#define ORC_STAT_RECON_FALLBACK 0
#define ORC_STAT_LIMIT_VARS 1
#define ORC_STAT_LIMIT_VARS_KN0 2
#define ORC_STAT_LIMIT_SA 3
#define ORC_STAT_LIMIT_SA_KN0 4
#define ORC_STAT_LOW_DT 5
void orc_stats(long *out);

/* OpenMP threads of the cell and face loops (n > 0 sets them); returns the
 * count in use.  Results do not depend on it: every parallel loop writes only
 * its own cells or faces, and every sum keeps the serial order. */
int orc_set_threads(int n);

/* Test hooks: the face flux (scheme 200 Steger-Warming, 300 Roe) on face-frame
 * states, and the reconstruction (var 100 conservative / 200 primitive) of one
 * face from its four-cell stencil FF, F | B, BB.  Also orc_get "SI" "SJ" "SK":
 * face vectors (faces x 3). */
int orc_test_flux(int scheme, double gamma, const double *ql, const double *qr, double que, double *out);
int orc_test_recon(int var, int scheme, double gamma, const double *ff, const double *f, const double *b,
                   const double *bb, double *l, double *r);

/* Viscous stages on a given primitive box (ghost cells included), without
 * boundary conditions; then orc_get "omega" "gradI/J/K" "fvI/J/K" "fv".  And
 * the Sutherland viscosity of a non-dimensional temperature. */
int orc_test_viscous(orc_ctx *c, int block, const double *pv_box);
double orc_test_sutherland(orc_ctx *c, double temp);

#ifdef __cplusplus
}
#endif
#endif
