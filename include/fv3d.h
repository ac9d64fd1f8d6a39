/* libfv3d -- C ABI of the MI355X per-iteration path of a structured
 * multi-block finite-volume turbomachinery solver.
 *
 * Every entry point returns 0 on success and a negative fv3d_status code on
 * failure; fv3d_last_error() then holds a message.  The library never exits
 * the process.  Calls on one context are stream-ordered on that context's HIP
 * stream; only the functions that return host data synchronise.  One context
 * per host thread.  Host pointers are borrowed for the duration of the call.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   the CMesh / CSpaceDiscr / CTimeDiscr / CSolver method surface of
 *   include/Mesh.h:31-90, include/SpaceDiscr.h:62-69, include/TimeDiscr.h:43-49,
 *   include/Solver.h:46-52.  The reference has no FFI; its classes call each
 *   other directly.  The C++ classes in include/fv3d_classes.h keep those
 *   names and forward to the entry points below.
 */
#ifndef FV3D_H
#define FV3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fv3d_ctx fv3d_ctx;

enum fv3d_status {
    FV3D_OK = 0,
    FV3D_EINVAL = -1,      /* bad argument or unsupported option */
    FV3D_ESTATE = -2,      /* call out of order (e.g. before fv3d_setup) */
    FV3D_EHIP = -3,        /* HIP runtime error */
    FV3D_ENOMEM = -4,
    FV3D_ENUMERIC = -5     /* non-finite residual (the reference's divergence print) */
};

enum fv3d_precision {
    FV3D_PLAIN_FP64 = 0,   /* every kernel in FP64 */
    FV3D_DEFECT_FP32 = 1   /* FP64 residual/update/norms, FP32 LU-SGS operator (both
                              sweeps, default sweep form); converges to the FP64
                              fixed point.  The cycle's residual evaluations that
                              only a restriction reads skip lambda, dt and the
                              diagonals, so after an iteration those fields are the
                              iteration's first evaluation's */
};

/* Raw config.dat values, as parsed by CUserInput::ReadConfig
 * (source/UserInput.cpp:17-239); the library applies the reference's float
 * parsing and non-dimensionalisation (source/UserInput.cpp:1070-1092). */
typedef struct {
    double rgas, gamma, pr_lam, pr_tur;
    double pt_in, tt_in, p_out, ma0, len_ref, ev_ratio;
    double inlet_dir[3];
    double cfl, c_timestep, relax;
    int turbo_type;     /* 0 axial, 1 centrifugal, 2 general outlet */
    int viscous;        /* 0 Euler, 1 Navier-Stokes */
    int recon_var;      /* 100 conservative, 200 primitive, any other value characteristic (source/UserInput.cpp:117-130) */
    int recon_scheme;   /* 100 NND, 200 MUSCL2u, 300 MUSCL2c, 400 MUSCL2o, 500 MUSCL3 */
    int flux_scheme;    /* 200 Steger-Warming, 300 Roe */
    int turbulence;     /* 0 none (laminar), 100 Baldwin-Lomax (source/BL_Model.cpp),
                           200 Spalart-Allmaras (source/SA_Model.cpp) */
    int limiter;        /* limit flow variables (source/Limit_FlowVars.cpp:8-119) */
    int global_dt;      /* 0 local time step, 1 global: the minimum over each block (source/TimeStep.cpp:64-86) */
    int device;         /* HIP device ordinal */
    int precision;      /* enum fv3d_precision, for both grid levels */
    /* two-grid cycle and start field (source/UserInput.cpp:50-58, 72-92) */
    int multigrid;      /* UseMultigrid: a second, coarse level when every patch range allows it */
    int coarse_iter;    /* Coarse_iter: coarse-level iterations before the cycle */
    int initial_type;   /* start field of fv3d_initial: 0 uniform, 1 interpolated along z */
    double p21ratio;    /* P21ratio: outlet/inlet static pressure of the interpolated start */
    double outlet_dir[3];   /* OutletAirAngle: flow direction at the outlet of the interpolated start */
    /* Grid levels with multigrid (no config.dat counterpart): 0 or 2 runs the
     * reference's two-grid cycle, capped at two levels by
     * source/UserInput.cpp:1057-1058; 3 to 5 run an N-level V-cycle (full
     * approximation storage below the first coarse level) over as many of
     * them as every block and patch range can be halved to. */
    int mg_levels;
} fv3d_config;

/* Context life cycle. */
int fv3d_ctx_create(const fv3d_config *cfg, fv3d_ctx **out);
int fv3d_ctx_destroy(fv3d_ctx *ctx);
const char *fv3d_last_error(const fv3d_ctx *ctx);

/* Adds a block (CUserInput::ReadGrids, source/UserInput.cpp:241-286 and the
 * block table, 439-445).  nodes: IM*JM*KM*3 doubles, node (i,j,k) 1-based at
 * ((i-1)*JM+(j-1))*KM+(k-1).  Returns the 1-based block id (>0) or an error. */
int fv3d_block_add(fv3d_ctx *ctx, int IM, int JM, int KM, const double *nodes, int nblade, double rpm);

/* Adds a patch (CUserInput::ReadBndGrid, source/UserInput.cpp:336-432).
 * kind 0 inlet, 1 outlet, 2 wall; rng = is ie js je ks ke. */
int fv3d_patch_add(fv3d_ctx *ctx, int kind, int block, int face, const int *rng, int rotate);

/* Mixing plane between two blade rows (replaces the MixingPlane rows of the
 * boundary table, source/UserInput.cpp:420-432, applied by
 * CMesh::BndCondMixingPlane, source/BndCondMixingPlane.cpp:287-522): block1's
 * K-high face (rng1 = {is, ie, js, je, KM, KM}, cell ranges) against block2's
 * K-low face (rng2 = {is, ie, js, je, 1, 1}); both sides use rng1's radial
 * range, at most 512 cells pitch-wise per side.  Per radial line the
 * flux-averaged state of each side sets the other side's ghost cells through
 * characteristic relations; nu~ ghosts take the upstream average.  Applied
 * after the interior and periodic patches, on every grid level. */
int fv3d_mixing_plane_add(fv3d_ctx *c, int block1, const int *rng1, int block2, const int *rng2);

/* Adds an interior (kind 0) or periodic (kind 1) patch between face1 of block1
 * and face2 of block2 (CUserInput::ReadBndGrid, source/UserInput.cpp:386-417;
 * applied by CMesh::BndCondInterior / BndCondPeriodic,
 * source/BndCondInteriorPeriodic.cpp).  orient 1..4 is the reference's
 * CMesh::Trans code (source/ShadowOutIn.cpp:166-325); ranges as in
 * fv3d_patch_add.  A periodic patch rotates by the pitch 2*PI/NBlade of block1. */
int fv3d_connection_add(fv3d_ctx *ctx, int kind, int orient, int block1, int face1, const int *rng1, int block2,
                        int face2, const int *rng2);

/* Metrics (CGeometry::FaceVectorVolume, source/GeometryVariables.cpp:10-105),
 * device allocation and launch plans.  Must follow the last block/patch add.
 * With cfg.multigrid it decides the number of grid levels as
 * CUserInput::NumberOfMultigrid does (source/UserInput.cpp:799-1067: two when
 * every patch range starts and ends on an odd node, else one) and sets up the
 * coarse level (nodes 2i-1, 2j-1, 2k-1 and halved patch ranges,
 * source/UserInput.cpp:241-334 and 446-797). */
int fv3d_setup(fv3d_ctx *ctx);

/* Grid levels in use after fv3d_setup: 1, 2 for the two-grid cycle, or up to
 * cfg.mg_levels for the N-level V-cycle. */
int fv3d_levels(const fv3d_ctx *ctx);

/* Batching: independent passage instances in one context, advanced together by
 * every launch (no reference counterpart: the reference runs one case per
 * process, source/Structured_FV3D.cpp:19-214; each instance here behaves as
 * that run would, bit for bit).  Declare an instance before its blocks; the
 * blocks added next, and the patches and connections on them, belong to it.
 * p_out is the instance's outlet static pressure in config.dat units (the
 * config's p_out is ignored once instances are declared); every other config
 * value is shared.  Without a declaration all blocks form one instance.
 * Returns the instance number (1-based). */
int fv3d_instance_add(fv3d_ctx *ctx, double p_out);

/* Number of instances (1 unless declared). */
int fv3d_instances(const fv3d_ctx *ctx);

/* Residual norms (as fv3d_res_calculate) and performance integrals (as
 * fv3d_perf) of one instance after the last iteration; either pointer may be
 * NULL.  Synchronises. */
int fv3d_instance_results(fv3d_ctx *ctx, int inst, double *res6, double *perf5);

/* CMesh::Initial (source/ComputeAllBlockVars.cpp:38-48; Initial_Uniform and
 * Initial_Interpolate, source/Initialization.cpp:17-165): the start field when
 * there is no restart file (reuse_solution = N), on the level the reference
 * starts from: the coarse one with two levels, else the only one.  The
 * two-grid cycle must start here: from a restart file the reference's cycle
 * divides by coarse time steps it never computed. */
int fv3d_initial(fv3d_ctx *ctx);

/* Restart-file state (CUserInput::ReadSolution, source/UserInput.cpp:1118-1159):
 * pv (ni,nj,nk,6) = rho,u,v,w,p,T and ener (ni,nj,nk), i slowest. */
int fv3d_state_upload(fv3d_ctx *ctx, int block, const double *pv, const double *ener);

/* Restart-file nu~ of Spalart-Allmaras (the nut field CUserInput::ReadSolution
 * reads, source/UserInput.cpp:1134-1153): ev (ni,nj,nk), i slowest.  With
 * Spalart-Allmaras the iteration needs nu~ from here (every block) or from
 * fv3d_initial (ev_ratio everywhere). */
int fv3d_nut_upload(fv3d_ctx *ctx, int block, const double *ev);

/* Copies a field to the host (synchronises).  Names and shapes:
 *   cv (IM+1,JM+1,KM+1,5), pv (IM+1,JM+1,KM+1,6)  -- cell arrays incl. ghost cells
 *   res dq (ni,nj,nk,5); dt lamcI lamcJ lamcK vol (ni,nj,nk)
 *   fcI (IM,nj,nk,5) fcJ (ni,JM,nk,5) fcK (ni,nj,KM,5)  -- face fluxes
 *   Navier-Stokes: lamvI lamvJ lamvK omega (ni,nj,nk); turbulence models: mut
 *   (IM+1,JM+1,KM+1) incl. ghost cells; Spalart-Allmaras: ev (IM+1,JM+1,KM+1)
 *   incl. ghost cells, resev dev dwall (ni,nj,nk) -- nu~ residual, its LU-SGS
 *   increment and the wall distance; with debug bit 0 also
 *   fvI fvJ fvK (viscous face fluxes, shapes as fcI..), gradI gradJ gradK
 *   (face gradients, 12: d/dx, d/dy, d/dz of u v w T) and fv (ni,nj,nk,5)
 *   dIS_cv dIP_cv (2,nj,nk,5), dJS_cv dJP_cv (ni,2,nk,5), dKS_cv dKP_cv (ni,nj,2,5)
 *   and the same with _pv and 6 components -- Dummy layers 1..2.
 * Returns the number of doubles written. */
long fv3d_field_download(fv3d_ctx *ctx, int block, const char *name, double *out);

/* CUserInput::ReadWallDistance (source/UserInput.cpp:1094-1116, used with
 * ReuseWallDistance = Y): replaces one block's cell wall distances (ni*nj*nk,
 * k fastest, as fv3d_field_download's "dwall") with the caller's, after
 * fv3d_setup computed them (Spalart-Allmaras only). */
int fv3d_wall_distance_upload(fv3d_ctx *ctx, int block, const double *d);

/* The same for a grid level: 1 fine, 2 coarse (two-grid cycle). */
long fv3d_level_field_download(fv3d_ctx *ctx, int level, int block, const char *name, double *out);

/* Stages of CMesh::ResFlux (source/ComputeAllBlockVars.cpp:223-255). */
int fv3d_bc_apply(fv3d_ctx *ctx);          /* CMesh::BoundaryCondition */
/* CFluidProps::SpectralRadii (source/Spectral_curl_grad.cpp:9-42) and
 * CTimeDiscr::TimeStep (source/TimeStep.cpp:7-92) share one kernel: either
 * call forms lambda_c, lambda_v, dt and the LU-SGS diagonals; fv3d_timestep
 * also counts the low-dt damping down once, as TimeStep does. */
int fv3d_spectral(fv3d_ctx *ctx);
int fv3d_timestep(fv3d_ctx *ctx);
int fv3d_convective(fv3d_ctx *ctx);        /* CMesh::ConvectiveFlux incl. WallFlux, source term and Res */
/* CMesh::ViscousFlux (source/ComputeAllBlockVars.cpp:131-160) with
 * CFluidProps::Gradient (source/Spectral_curl_grad.cpp:85-493): subtracts the
 * viscous flux from the residual that fv3d_convective wrote (Navier-Stokes;
 * a no-op for Euler). */
int fv3d_viscous(fv3d_ctx *ctx);
/* CFluidProps::Curl (source/Spectral_curl_grad.cpp:44-83): the vorticity
 * magnitude "omega".  Only turbulence models read it, so fv3d_resflux does not
 * evaluate it for laminar flow. */
int fv3d_curl(fv3d_ctx *ctx);
/* The turbulence stage of ResFlux (source/ComputeAllBlockVars.cpp:228-232):
 * CBlocks::BL_Model (source/BL_Model.cpp:8-329) with its vorticity, or
 * CSpaceDiscr::SA_Model (source/SA_Model.cpp:9-258: vorticity, the nu~
 * residual, mue_t with its clipping, ghost mue_t), or the learned closure when
 * it is on; nothing for laminar flow.  fv3d_resflux runs it after the ghost
 * fill. */
int fv3d_turbulence(fv3d_ctx *ctx);
int fv3d_resflux(fv3d_ctx *ctx);           /* all of CMesh::ResFlux */

/* CMesh::LU_SGS_Marching (source/LU_SGS_Marching.cpp:10-291): save, ResFlux,
 * forward and backward hyperplane sweeps, update, PrimitiveVars, Limit_Vars. */
int fv3d_lusgs(fv3d_ctx *ctx);

/* CFluidProps::PrimitiveVars (source/PrimitiveVars.cpp:7-29) and
 * CFluidProps::Limit_Vars (source/Limit_FlowVars.cpp:8-119, every block
 * scanned; with the limiter off in the config it does nothing) as calls of
 * their own; fv3d_lusgs runs both at its end. */
int fv3d_primitive_vars(fv3d_ctx *ctx);
int fv3d_limit_vars(fv3d_ctx *ctx);

/* The two-grid cycle by stage (CSolver::Solver, source/Solver.cpp:61-103; two
 * levels in use).  A start-up iteration is fv3d_mg_coarse_lusgs(ctx, 0) (the
 * coarse level's own LU-SGS step), followed at Coarse_iter by
 * fv3d_mg_interpolate_state (CornerTreatment, Interpolate of the state,
 * PrimitiveVars, Limit_Vars); a cycle iteration is fv3d_lusgs, fv3d_resflux,
 * fv3d_mg_restrict (CSolver::Prolong with ComputeForceFunction /
 * AddForceFunction, source/Prolong_Interpolate.cpp:7-50 and
 * source/ComputeAllBlockVars.cpp:269-309, then the coarse PrimitiveVars and
 * Limit_Vars), fv3d_mg_coarse_lusgs(ctx, 1) (the forced coarse LU-SGS on the
 * coarse lambda and dt), fv3d_mg_correct (ComputeDeltConsVars, Interpolate of
 * the correction, PrimitiveVars, Limit_Vars, source/Prolong_Interpolate.cpp:
 * 52-176); each iteration ends with fv3d_iteration_end, which forms
 * Res_Calculate, the CFL ramp and the section integrals of the level that
 * reports and returns its history row (iter, Res.dens, mass in, mass out, PR,
 * efficiency, CFL; one row per instance).  The same kernels as fv3d_iterate,
 * so the same bits. */
int fv3d_mg_coarse_lusgs(fv3d_ctx *ctx, int forced);
int fv3d_mg_interpolate_state(fv3d_ctx *ctx);
int fv3d_mg_restrict(fv3d_ctx *ctx);
int fv3d_mg_correct(fv3d_ctx *ctx);
int fv3d_iteration_end(fv3d_ctx *ctx, double *row7);

/* CMesh::Res_Calculate (source/ComputeAllBlockVars.cpp:318-372) of the first
 * instance; synchronises. */
int fv3d_res_calculate(fv3d_ctx *ctx, double *res6);

/* COutput::OutputGeneralParameters (source/Ouput.cpp:282-300) of the first
 * instance; synchronises.
 * out5 = mass_in, mass_out, PR, efficiency, mass-flow ratio. */
int fv3d_perf(fv3d_ctx *ctx, double *out5);

/* n iterations of CSolver::Solver (source/Solver.cpp:47-135) with the CFL ramp
 * kept on the device, each followed by the performance integrals of the level
 * the reference reports (source/Structured_FV3D.cpp:148-158).  With two levels:
 * coarse LU-SGS up to Coarse_iter, the coarse-to-fine state interpolation at
 * Coarse_iter, then per iteration fine LU-SGS, fine residual, restriction,
 * forced coarse LU-SGS and the prolonged correction.  hist (may be NULL)
 * receives n rows of: iter res_dens mass_in mass_out PR eff CFL, per instance
 * (n x instances rows, the instances of an iteration consecutive).
 * Synchronises once at the end. */
int fv3d_iterate(fv3d_ctx *ctx, int n, double *hist);

/* Same as fv3d_iterate, but replays one captured HIP graph per iteration. */
int fv3d_iterate_graph(fv3d_ctx *ctx, int n, double *hist);

int fv3d_sync(fv3d_ctx *ctx);

/* Average device time in milliseconds of the last timed iterations, per
 * kernel class: 0 residual, 1 LU-SGS sweeps (with the row gather and dQ
 * scatter of the default sweep form), 2 update, 3 ghost fill + spectral + time
 * step, 4 norms/integrals, 5 forward sweep kernel, 6 backward sweep kernel
 * (5 and 6 for the default form only), 7 coarse-grid correction (second fine
 * residual, restriction, coarse LU-SGS, prolongation; multigrid only),
 * 8 convective residual kernel, 9 viscous and nu~ residual kernels, 10 ghost
 * fill and patch exchange, 11 turbulence model (vorticity, Baldwin-Lomax,
 * Spalart-Allmaras or the learned closure).
 * ms and counts hold FV3D_TIMING_SLOTS entries.  Filled by fv3d_iterate when
 * timing is on (fine-level iterations only). */
#define FV3D_TIMING_SLOTS 12
int fv3d_set_timing(fv3d_ctx *ctx, int on);
int fv3d_timing(fv3d_ctx *ctx, double *ms, long *counts);

/* Debug flags: bit 0 keeps the face fluxes of the last residual evaluation
 * so fv3d_field_download can return fcI/fcJ/fcK, and for Navier-Stokes the
 * viscous fvI/fvJ/fvK, gradI/gradJ/gradK and fv (single-block meshes).
 * Bit 1 runs the LU-SGS sweeps as one launch per hyperplane (the reference's
 * schedule, source/LU_SGS_Marching.cpp:59-260) instead of the persistent
 * wavefront, as a cross-check (plain FP64 only).  Other bits are refused. */
int fv3d_set_debug(fv3d_ctx *ctx, int flags);

/* Tile shapes of the residual kernels (no reference counterpart: the
 * reference loops over faces).  Each of the three residual kernels -- 0
 * convective, 1 viscous, 2 Spalart-Allmaras nu~ -- runs with one of the shapes
 * (cells along i x j x k per workgroup): 0 4x4x16 (default), 1 2x8x16,
 * 2 4x8x8, 3 8x4x8, 4 8x8x4.  The convective kernel also takes 5-9, a
 * column of (j x k) 16x16 or 8x32 cells streamed along i in chunks of 4, 2,
 * 4, 2 and 8 i-planes (5 16x16/4, 6 16x16/2, 7 8x32/4, 8 8x32/2, 9 16x16/8),
 * and 10, one thread per face (face fluxes to memory, then one thread per
 * cell sums them), which coarse multigrid levels of at most 1,000,000 cells
 * take by default (fv3d_autotune tries it on such grids only).  Results do
 * not depend on the shape.
 * fv3d_set_tiles sets the three shape numbers; fv3d_autotune times every shape
 * of every kernel in use on the current state (reps launches each, after one
 * warm-up), keeps the fastest and writes the chosen numbers (-1 for a kernel
 * not in use) to shape3 (may be NULL).  Both drop the captured graphs; call
 * them between iterations (the timing runs overwrite the residual, which the
 * next iteration forms again). */
int fv3d_set_tiles(fv3d_ctx *ctx, const int *shape3);
int fv3d_autotune(fv3d_ctx *ctx, int reps, int *shape3);

/* Learned eddy-viscosity closure (no reference counterpart; it takes the
 * place of the Spalart-Allmaras transport, source/SA_Model.cpp, in a context
 * with turbulence 200).  A 6 -> 64 -> 64 -> 64 -> 1 network, tanh hidden
 * layers, softplus output y = log(1 + mue_t / mue_lam), BF16 weights and
 * activations on the matrix cores with FP32 accumulation.  params: 8833 FP32
 * values W0[64][6] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64] b3.
 * Features per cell (fv3d_mlp_features, 6 floats, cells i-slowest):
 * log(1 + d|U|/nu), log(1 + S d^2/nu), log(1 + Omega d^2/nu),
 * (Omega^2 - S^2)/(Omega^2 + S^2), log(1 + |U| Vol^(1/3)/nu),
 * tanh(d U.grad p / (rho |U|^3)).
 * fv3d_mlp_set: parameters and the gate bounds (min and max of each feature,
 * then the largest output; NULL for none), and a fresh Adam state.
 * fv3d_mlp_forward: outputs y of n samples.  fv3d_mlp_grad: the gradient of
 * the mean squared error of y against targets t over n samples (and the loss).
 * fv3d_mlp_train: steps Adam steps (lr) over consecutive minibatches of batch
 * samples (a multiple of 64), loss_hist[steps] (may be NULL).
 * fv3d_mlp_closure(on): mue_t from the network each residual evaluation and no
 * nu~ equation; a cell outside the gate keeps the Spalart-Allmaras value of its
 * frozen nu~.  fv3d_mlp_relax(omega): mue_t moves by omega (0 < omega <= 1,
 * default 1) of the way from its previous value to the closure's each
 * evaluation (no reference counterpart: the closure is algebraic, and damping
 * its coupling to the flow leaves the converged state unchanged).
 * fv3d_mlp_tripped: cells that fell back since the last call. */
int fv3d_mlp_set(fv3d_ctx *ctx, const float *params, const float *bounds);
int fv3d_mlp_get(fv3d_ctx *ctx, float *params);
long fv3d_mlp_features(fv3d_ctx *ctx, int block, float *out);
/* Training targets log(1 + mue_t / mue_lam) of a block's cells from the
 * current eddy viscosity (ni,nj,nk floats). */
long fv3d_mlp_targets(fv3d_ctx *ctx, int block, float *out);
int fv3d_mlp_forward(fv3d_ctx *ctx, const float *x, long n, float *y);
int fv3d_mlp_grad(fv3d_ctx *ctx, const float *x, const float *t, long n, float *grad, float *loss);
/* No counterpart in the reference.  fv3d_mlp_grad for callers that cannot
 * allocate (code under HIP graph capture, training or serving loops that
 * forbid hipMalloc): x_dev (n x 6), t_dev (n), grad_dev (8833) and loss_dev
 * (1, the mean loss) are device buffers the caller owns; the partial
 * gradients go to a static device scratch array of the library (9 MB per
 * GPU, held until the process ends), so n is at most 16,384 samples and
 * larger n is rejected with FV3D_EINVAL.  Nothing is allocated or freed.
 * The call is asynchronous on the context's stream (fv3d_sync waits); calls
 * that overlap on one GPU, from any context, race on the shared scratch. */
int fv3d_mlp_grad_noalloc(fv3d_ctx *ctx, const float *x_dev, const float *t_dev, long n, float *grad_dev,
                          float *loss_dev);
int fv3d_mlp_train(fv3d_ctx *ctx, const float *x, const float *t, long n, int steps, int batch, float lr,
                   float *loss_hist);
int fv3d_mlp_closure(fv3d_ctx *ctx, int on);
int fv3d_mlp_relax(fv3d_ctx *ctx, double omega);
long fv3d_mlp_tripped(fv3d_ctx *ctx);

/* Number of interior cells over all blocks. */
long fv3d_cell_count(const fv3d_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
