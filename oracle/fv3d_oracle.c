/* CPU restatement of the reference solver's per-iteration path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never linked by
 * it.  Pinned against the compiled reference (oracle/_ref) by
 * tests/test_oracle_vs_ref.py and against closed forms by tests/test_oracle_laws.py.
 *
 * Storage mirrors the reference's extents (source/FluidProps.cpp:650-760,
 * source/Geometry.cpp:109-155): cell arrays [0..IM][0..JM][0..KM], face arrays
 * over the same box, Dummy slabs [1..ND] per face.  Arithmetic follows the
 * reference's expression order so that, built with -ffp-contract=off against
 * the same libm, results are bit-identical to the reference at one thread.
 */
#include "fv3d_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ND 2
/* Cell and face loops run on OpenMP threads for blocks of more than this many
 * cells; every iteration writes only its own cell or face, so the results do
 * not depend on the thread count. */
#define PAR_CELLS 4096L
#define PAR(b) ((long)(b)->IM * (b)->JM * (b)->KM > PAR_CELLS)
#define F32(x) ((double)(x##f))

typedef struct { double r, ru, rv, rw, re; } cv_t;
typedef struct { double d, u, v, w, p, t; } pv_t;
typedef struct { double x, y, z; } v3_t;
/* g: the nine coordinate-transformation coefficients of JacobiConvertCoeff in
 * the reference's order gradx_kesi grady_kesi gradz_kesi, gradx_yita ...,
 * gradx_zita ... (include/defs.h TFace / TCell) */
typedef struct { v3_t S; double s; v3_t c; double r, th; double g[9]; } face_t;
typedef struct { double vol; v3_t c; double r, th; double g[9]; } cell_t;

typedef struct {
    int kind, face, is, ie, js, je, ks, ke, rotate;
    double rot;
} patch_t;

/* two-sided patch of the boundary table: 0 interior, 1 periodic
 * (source/UserInput.cpp:386-417); side 0 is block 1 of the row */
typedef struct {
    int kind, orient;
    int blk[2], face[2], rng[2][6];
    int flip;   /* periodic: theta of side 2 below theta of side 1 (source/BndCondInteriorPeriodic.cpp:54-59) */
} conn_t;

typedef struct {
    int IM, JM, KM, nblade;
    double w_rot, cellvol;
    double delta_th;             /* blade pitch 2*PI/NBlade, sign flipped by the periodic exchange */
    v3_t *node;                  /* [1..IM][1..JM][1..KM] in a (IM+1)(JM+1)(KM+1) box */
    face_t *fi, *fj, *fk;        /* face boxes (IM+1)(JM+1)(KM+1) */
    cell_t *cell;
    cv_t *cv, *cv_old, *dq, *res, *fc, *face_fc[3];
    pv_t *pv;
    double *lamc[3], *lamv[3], *dt, *mut;
    /* Navier-Stokes: vorticity magnitude per cell, face gradients (12 per face:
     * d/dx of u v w T, then d/dy, then d/dz), face and cell viscous fluxes */
    double *omega, *grad[3];
    cv_t *fv, *face_fv[3];
    cv_t *dcv[6];                /* IS IP JS JP KS KP Dummy slabs */
    pv_t *dpv[6];
    int flag_low_dt;
    patch_t *patch;
    int npatch;
    /* coarse level of the two-grid cycle (source/Mesh.cpp:109-152): restricted
     * state cv_2h, restricted residual Res_2h and the correction delt_cv */
    cv_t *cv2h, *res2h, *delt;
    /* Spalart-Allmaras: nu~ over the cell box, its old value, increment and
     * residual, the wall distance, the face gradients of nu~ (x, y, z per face),
     * the face fluxes of SA_Model (convective, viscous, and the plain gradient
     * flux used by the c_b2 term) and the cell sums of ConvectiveFlux/SA_Model */
    double *ev, *ev_old, *dev, *res_ev, *dwall;
    double *gev[3], *fc_tur[3], *fv_tur[3], *fv1_tur[3];
    double *cfc_tur, *cfv_tur;
} blk_t;

struct orc_ctx {
    orc_config cfg;
    orc_nondim nd;
    double cfl, cfl_res;
    double total_vol;
    double res[6];
    int iter;
    int nb;
    blk_t blk[64];
    conn_t conn[256];
    int nconn;
    /* mixing planes (source/UserInput.cpp:420-432): block 1's K-high face
     * against block 2's K-low face, cell ranges as the table gives them */
    struct { int blk[2], rng[2][6]; } mix[16];
    int nmix;
    int first_in[2], first_out[2];   /* block, patch index of BndGeom.Inlet[1] / Outlet[1] */
    int ng;                          /* grid levels in use */
    int force;                       /* coarse level: UseForceFunction (source/LU_SGS_Marching.cpp:49-57) */
    struct orc_ctx *coarse;
};

static void curl(blk_t *b);
static void gradients(blk_t *b);
static void wall_distance(orc_ctx *c);

/* ------------------------------------------------------------------ index */
static inline long C3(const blk_t *b, int i, int j, int k) { return ((long)i * (b->JM + 1) + j) * (b->KM + 1) + k; }
static inline long CI(const blk_t *b, int i, int j, int k) { return (((long)(i - 1)) * (b->JM - 1) + (j - 1)) * (b->KM - 1) + (k - 1); }
/* Dummy slab index: d 0..5 = IS IP JS JP KS KP; (a,b) tangential, n layer */
static inline long DI(const blk_t *b, int d, int i, int j, int k)
{
    /* IS/IP: [nd][JM][KM], JS/JP: [IM][nd][KM], KS/KP: [IM][JM][nd] (nd 0..ND) */
    if (d < 2) return ((long)i * b->JM + j) * b->KM + k;
    if (d < 4) return ((long)i * (ND + 1) + j) * b->KM + k;
    return ((long)i * b->JM + j) * (ND + 1) + k;
}
static long dsize(const blk_t *b, int d)
{
    if (d < 2) return (long)(ND + 1) * b->JM * b->KM;
    if (d < 4) return (long)b->IM * (ND + 1) * b->KM;
    return (long)b->IM * b->JM * (ND + 1);
}

/* ------------------------------------------------------------ small algebra */
static inline v3_t vsub(v3_t a, v3_t b) { v3_t c = {a.x - b.x, a.y - b.y, a.z - b.z}; return c; }
static inline v3_t vadd(v3_t a, v3_t b) { v3_t c = {a.x + b.x, a.y + b.y, a.z + b.z}; return c; }
static inline v3_t vscale(v3_t a, double s) { v3_t c = {a.x * s, a.y * s, a.z * s}; return c; }
static inline v3_t vdiv(v3_t a, double s) { v3_t c = {a.x / s, a.y / s, a.z / s}; return c; }
static inline double vdot(v3_t a, v3_t b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

static inline cv_t cadd(cv_t a, cv_t b) { cv_t c = {a.r + b.r, a.ru + b.ru, a.rv + b.rv, a.rw + b.rw, a.re + b.re}; return c; }
static inline cv_t csub(cv_t a, cv_t b) { cv_t c = {a.r - b.r, a.ru - b.ru, a.rv - b.rv, a.rw - b.rw, a.re - b.re}; return c; }
static inline cv_t cmul(cv_t a, cv_t b) { cv_t c = {a.r * b.r, a.ru * b.ru, a.rv * b.rv, a.rw * b.rw, a.re * b.re}; return c; }
static inline cv_t cdivv(cv_t a, cv_t b) { cv_t c = {a.r / b.r, a.ru / b.ru, a.rv / b.rv, a.rw / b.rw, a.re / b.re}; return c; }
static inline cv_t cscale(cv_t a, double s) { cv_t c = {a.r * s, a.ru * s, a.rv * s, a.rw * s, a.re * s}; return c; }
static inline cv_t cdiv(cv_t a, double s) { cv_t c = {a.r / s, a.ru / s, a.rv / s, a.rw / s, a.re / s}; return c; }
static inline cv_t caddk(cv_t a, double s) { cv_t c = {a.r + s, a.ru + s, a.rv + s, a.rw + s, a.re + s}; return c; }
static inline cv_t ksubc(double s, cv_t a) { cv_t c = {s - a.r, s - a.ru, s - a.rv, s - a.rw, s - a.re}; return c; }
/* POW(x, 2.0): the reference's component-wise POW (source/defs.cpp:163-172,
 * 295-305) receives the exponent at run time and calls libm pow, whose result
 * differs from x*x in the last bit for some x.  The exponent is kept opaque here
 * so the compiler cannot fold the call into x*x. */
static volatile double pow_two = 2.0;
static inline cv_t cvsq(cv_t a) { double e = pow_two; cv_t c = {pow(a.r, e), pow(a.ru, e), pow(a.rv, e), pow(a.rw, e), pow(a.re, e)}; return c; }

static inline pv_t padd(pv_t a, pv_t b) { pv_t c = {a.d + b.d, a.u + b.u, a.v + b.v, a.w + b.w, a.p + b.p, a.t + b.t}; return c; }
static inline pv_t psub(pv_t a, pv_t b) { pv_t c = {a.d - b.d, a.u - b.u, a.v - b.v, a.w - b.w, a.p - b.p, a.t - b.t}; return c; }
static inline pv_t pmul(pv_t a, pv_t b) { pv_t c = {a.d * b.d, a.u * b.u, a.v * b.v, a.w * b.w, a.p * b.p, a.t * b.t}; return c; }
static inline pv_t pdivv(pv_t a, pv_t b) { pv_t c = {a.d / b.d, a.u / b.u, a.v / b.v, a.w / b.w, a.p / b.p, a.t / b.t}; return c; }
static inline pv_t kmulp(double s, pv_t a) { pv_t c = {s * a.d, s * a.u, s * a.v, s * a.w, s * a.p, s * a.t}; return c; }
static inline pv_t pdiv(pv_t a, double s) { pv_t c = {a.d / s, a.u / s, a.v / s, a.w / s, a.p / s, a.t / s}; return c; }
static inline pv_t paddk(pv_t a, double s) { pv_t c = {a.d + s, a.u + s, a.v + s, a.w + s, a.p + s, a.t + s}; return c; }
static inline pv_t ksubp(double s, pv_t a) { pv_t c = {s - a.d, s - a.u, s - a.v, s - a.w, s - a.p, s - a.t}; return c; }
static inline pv_t pvsq(pv_t a) { double e = pow_two; pv_t c = {pow(a.d, e), pow(a.u, e), pow(a.v, e), pow(a.w, e), pow(a.p, e), pow(a.t, e)}; return c; }

static inline double MAXd(double a, double b) { return a > b ? a : b; }
static inline double MINd(double a, double b) { return a < b ? a : b; }
static inline double MIN3d(double a1, double a2, double a3) { double m = a1; if (m > a2) m = a2; if (m > a3) m = a3; return m; }

/* source/Reconstruction.cpp:472-559 */
static inline double mm1(double a, double b)
{
    if (a * b < 0.0) return 0.0;
    if (fabs(a) <= fabs(b)) return a;
    return b;
}
static cv_t cminmod(cv_t a, cv_t b) { cv_t c = {mm1(a.r, b.r), mm1(a.ru, b.ru), mm1(a.rv, b.rv), mm1(a.rw, b.rw), mm1(a.re, b.re)}; return c; }
static pv_t pminmod(pv_t a, pv_t b) { pv_t c = {mm1(a.d, b.d), mm1(a.u, b.u), mm1(a.v, b.v), mm1(a.w, b.w), mm1(a.p, b.p), mm1(a.t, b.t)}; return c; }

/* ------------------------------------------------------------- set-up */
void orc_nondimensionalize(const orc_config *cfg, orc_nondim *nd)
{
    /* config values are parsed with stof (source/UserInput.cpp:60-76) */
    double gamma = (float)cfg->gamma, rgas = (float)cfg->rgas;
    double ttin = (float)cfg->tt_in, ptin = (float)cfg->pt_in, pout = (float)cfg->p_out;
    double len_ref = (float)cfg->len_ref;
    /* source/UserInput.cpp:1073-1084 */
    nd->temp_ref = ttin;
    nd->press_ref = ptin;
    nd->vel_ref = sqrt(gamma * rgas * nd->temp_ref);
    nd->mue_ref = F32(1.716e-5) * pow(nd->temp_ref / F32(288.15), F32(3.0) / F32(2.0)) * (double)(288.15f + 110.4f) /
                  (nd->temp_ref + F32(110.4));
    nd->dens_ref = nd->press_ref / rgas / nd->temp_ref;
    nd->re_ref = nd->dens_ref * nd->vel_ref * len_ref / nd->mue_ref;
    nd->p_out = pout / (nd->dens_ref * nd->vel_ref * nd->vel_ref);
    nd->tt_in = F32(1.0);
    nd->pt_in = F32(1.0) / gamma;
    nd->rgas = F32(1.0) / (F32(1.0) * F32(1.0) * gamma);
    nd->gamma = gamma;
    nd->pr_lam = (float)cfg->pr_lam;
    nd->pr_tur = (float)cfg->pr_tur;
    nd->ev_ratio = (float)cfg->ev_ratio;
    for (int c = 0; c < 3; ++c) nd->inlet_dir[c] = cfg->inlet_dir[c];
}

orc_ctx *orc_create(const orc_config *cfg)
{
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->cfg = *cfg;
    orc_nondimensionalize(cfg, &c->nd);
    c->cfl = (float)cfg->cfl;
    c->cfl_res = F32(1.0E-3);     /* source/Solver.cpp:21 */
    c->first_in[0] = c->first_out[0] = -1;
    c->ng = 1;
    return c;
}

static void blk_free(blk_t *b)
{
    free(b->node); free(b->fi); free(b->fj); free(b->fk); free(b->cell);
    free(b->cv); free(b->cv_old); free(b->dq); free(b->res); free(b->fc); free(b->pv);
    for (int d = 0; d < 3; ++d) {
        free(b->face_fc[d]); free(b->lamc[d]); free(b->lamv[d]); free(b->grad[d]); free(b->face_fv[d]);
    }
    free(b->dt); free(b->mut); free(b->omega); free(b->fv);
    for (int d = 0; d < 6; ++d) { free(b->dcv[d]); free(b->dpv[d]); }
    free(b->patch);
    free(b->cv2h); free(b->res2h); free(b->delt);
    free(b->ev); free(b->ev_old); free(b->dev); free(b->res_ev); free(b->dwall); free(b->cfc_tur); free(b->cfv_tur);
    for (int d = 0; d < 3; ++d) { free(b->gev[d]); free(b->fc_tur[d]); free(b->fv_tur[d]); free(b->fv1_tur[d]); }
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    orc_destroy(c->coarse);
    for (int n = 0; n < c->nb; ++n) blk_free(&c->blk[n]);
    free(c);
}

int orc_block_add(orc_ctx *c, int IM, int JM, int KM, const double *nodes, int nblade, double rpm)
{
    if (c->nb >= 64) return -1;
    blk_t *b = &c->blk[c->nb++];
    memset(b, 0, sizeof *b);
    b->IM = IM; b->JM = JM; b->KM = KM; b->nblade = nblade;
    /* source/UserInput.cpp:444 and 1088 */
    b->w_rot = (double)(2.0f * 3.1415926535897932f) * rpm / F32(60.0);
    b->w_rot = b->w_rot / (c->nd.vel_ref / (double)(float)c->cfg.len_ref);
    b->delta_th = (double)(2.0f * 3.1415926535897932f / (float)nblade);   /* source/UserInput.cpp:443 */
    long box = (long)(IM + 1) * (JM + 1) * (KM + 1);
    b->node = (v3_t *)calloc(box, sizeof(v3_t));
    for (int i = 1; i <= IM; ++i)
        for (int j = 1; j <= JM; ++j)
            for (int k = 1; k <= KM; ++k) {
                const double *p = nodes + (((long)(i - 1) * JM + (j - 1)) * KM + (k - 1)) * 3;
                v3_t v = {p[0], p[1], p[2]};
                b->node[C3(b, i, j, k)] = v;
            }
    b->fi = (face_t *)calloc(box, sizeof(face_t));
    b->fj = (face_t *)calloc(box, sizeof(face_t));
    b->fk = (face_t *)calloc(box, sizeof(face_t));
    b->cell = (cell_t *)calloc(box, sizeof(cell_t));
    b->cv = (cv_t *)calloc(box, sizeof(cv_t));
    b->cv_old = (cv_t *)calloc(box, sizeof(cv_t));
    b->dq = (cv_t *)calloc(box, sizeof(cv_t));
    b->res = (cv_t *)calloc(box, sizeof(cv_t));
    b->fc = (cv_t *)calloc(box, sizeof(cv_t));
    b->pv = (pv_t *)calloc(box, sizeof(pv_t));
    for (int d = 0; d < 3; ++d) {
        b->face_fc[d] = (cv_t *)calloc(box, sizeof(cv_t));
        b->lamc[d] = (double *)calloc(box, sizeof(double));
        b->lamv[d] = (double *)calloc(box, sizeof(double));
        b->grad[d] = (double *)calloc(box * 12, sizeof(double));
        b->face_fv[d] = (cv_t *)calloc(box, sizeof(cv_t));
    }
    b->dt = (double *)calloc(box, sizeof(double));
    b->mut = (double *)calloc(box, sizeof(double));
    b->omega = (double *)calloc(box, sizeof(double));
    b->fv = (cv_t *)calloc(box, sizeof(cv_t));
    for (int d = 0; d < 6; ++d) {
        b->dcv[d] = (cv_t *)calloc(dsize(b, d), sizeof(cv_t));
        b->dpv[d] = (pv_t *)calloc(dsize(b, d), sizeof(pv_t));
    }
    double **sa[] = {&b->ev, &b->ev_old, &b->dev, &b->res_ev, &b->dwall, &b->cfc_tur, &b->cfv_tur};
    for (size_t q = 0; q < sizeof sa / sizeof sa[0]; ++q) *sa[q] = (double *)calloc(box, sizeof(double));
    for (int d = 0; d < 3; ++d) {
        b->gev[d] = (double *)calloc(box * 3, sizeof(double));
        b->fc_tur[d] = (double *)calloc(box, sizeof(double));
        b->fv_tur[d] = (double *)calloc(box, sizeof(double));
        b->fv1_tur[d] = (double *)calloc(box, sizeof(double));
    }
    return c->nb - 1;
}

int orc_mixing_plane_add(orc_ctx *c, int block1, const int *rng1, int block2, const int *rng2)
{
    if (c->nmix >= 16 || block1 < 1 || block1 > c->nb || block2 < 1 || block2 > c->nb) return -1;
    const blk_t *A = &c->blk[block1 - 1], *B = &c->blk[block2 - 1];
    if (rng1[4] != A->KM || rng1[5] != A->KM || rng2[4] != 1 || rng2[5] != 1) return -1;
    if (rng1[0] < 1 || rng1[1] > A->IM - 1 || rng1[1] > B->IM - 1 || rng1[0] > rng1[1]) return -1;
    if (rng1[2] < 1 || rng1[3] > A->JM - 1 || rng2[2] < 1 || rng2[3] > B->JM - 1) return -1;
    c->mix[c->nmix].blk[0] = block1 - 1;
    c->mix[c->nmix].blk[1] = block2 - 1;
    memcpy(c->mix[c->nmix].rng[0], rng1, 6 * sizeof(int));
    memcpy(c->mix[c->nmix].rng[1], rng2, 6 * sizeof(int));
    ++c->nmix;
    return 0;
}

int orc_connection_add(orc_ctx *c, int kind, int orient, int block1, int face1, const int *rng1, int block2, int face2,
                       const int *rng2)
{
    if (c->nconn >= 256 || block1 < 1 || block1 > c->nb || block2 < 1 || block2 > c->nb) return -1;
    if (orient < 1 || orient > 4 || face1 < 1 || face1 > 6 || face2 < 1 || face2 > 6) return -1;
    for (int q = 0; q < 3; ++q)
        if (rng1[2 * q] > rng1[2 * q + 1] || rng2[2 * q] > rng2[2 * q + 1]) return -1;
    conn_t *t = &c->conn[c->nconn++];
    t->kind = kind; t->orient = orient;
    t->blk[0] = block1 - 1; t->blk[1] = block2 - 1;
    t->face[0] = face1; t->face[1] = face2;
    memcpy(t->rng[0], rng1, 6 * sizeof(int));
    memcpy(t->rng[1], rng2, 6 * sizeof(int));
    t->flip = 0;
    if (kind == 1) {
        const blk_t *A = &c->blk[block1 - 1], *B = &c->blk[block2 - 1];
        v3_t na = A->node[C3(A, rng1[0], rng1[2], rng1[4])], nb = B->node[C3(B, rng2[0], rng2[2], rng2[4])];
        double th_1 = atan(na.y / na.x), th_2 = atan(nb.y / nb.x);
        t->flip = (th_2 - th_1) < 0.0;
    }
    return 0;
}

int orc_patch_add(orc_ctx *c, int kind, int block, int face, const int *rng, int rotate)
{
    if (block < 1 || block > c->nb) return -1;
    blk_t *b = &c->blk[block - 1];
    b->patch = (patch_t *)realloc(b->patch, sizeof(patch_t) * (b->npatch + 1));
    patch_t *p = &b->patch[b->npatch++];
    p->kind = kind; p->face = face; p->rotate = rotate;
    p->is = rng[0]; p->ie = rng[1]; p->js = rng[2]; p->je = rng[3]; p->ks = rng[4]; p->ke = rng[5];
    p->rot = 0.0;
    if (kind == 0 && c->first_in[0] < 0) { c->first_in[0] = block - 1; c->first_in[1] = b->npatch - 1; }
    if (kind == 1 && c->first_out[0] < 0) { c->first_out[0] = block - 1; c->first_out[1] = b->npatch - 1; }
    return 0;
}

/* source/GeometryVariables.cpp:10-105 */
static void face_vector_volume(blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    v3_t *N = b->node;
#define ND3(i, j, k) N[C3(b, i, j, k)]
    for (int i = 1; i <= IM; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                face_t *f = &b->fi[C3(b, i, j, k)];
                v3_t A = vsub(ND3(i, j + 1, k), ND3(i, j, k + 1)), B = vsub(ND3(i, j + 1, k + 1), ND3(i, j, k));
                f->S.x = F32(0.5) * (A.y * B.z - B.y * A.z);
                f->S.y = F32(0.5) * (A.z * B.x - A.x * B.z);
                f->S.z = F32(0.5) * (A.x * B.y - B.x * A.y);
                f->c = vscale(vadd(vadd(vadd(ND3(i, j, k), ND3(i, j + 1, k)), ND3(i, j + 1, k + 1)), ND3(i, j, k + 1)), 0.25);
                f->r = sqrt(f->c.x * f->c.x + f->c.y * f->c.y);
                f->th = atan(f->c.y / f->c.x);
                f->s = sqrt(f->S.x * f->S.x + f->S.y * f->S.y + f->S.z * f->S.z);
            }
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                face_t *f = &b->fj[C3(b, i, j, k)];
                v3_t A = vsub(ND3(i, j, k + 1), ND3(i + 1, j, k)), B = vsub(ND3(i + 1, j, k + 1), ND3(i, j, k));
                f->S.x = F32(0.5) * (A.y * B.z - B.y * A.z);
                f->S.y = F32(0.5) * (A.z * B.x - A.x * B.z);
                f->S.z = F32(0.5) * (A.x * B.y - B.x * A.y);
                f->c = vscale(vadd(vadd(vadd(ND3(i, j, k), ND3(i + 1, j, k)), ND3(i + 1, j, k + 1)), ND3(i, j, k + 1)), 0.25);
                f->r = sqrt(f->c.x * f->c.x + f->c.y * f->c.y);
                f->th = atan(f->c.y / f->c.x);
                f->s = sqrt(f->S.x * f->S.x + f->S.y * f->S.y + f->S.z * f->S.z);
            }
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM; ++k) {
                face_t *f = &b->fk[C3(b, i, j, k)];
                v3_t A = vsub(ND3(i + 1, j + 1, k), ND3(i, j, k)), B = vsub(ND3(i, j + 1, k), ND3(i + 1, j, k));
                f->S.x = F32(0.5) * (A.y * B.z - B.y * A.z);
                f->S.y = F32(0.5) * (A.z * B.x - A.x * B.z);
                f->S.z = F32(0.5) * (A.x * B.y - B.x * A.y);
                f->c = vscale(vadd(vadd(vadd(ND3(i, j, k), ND3(i + 1, j, k)), ND3(i + 1, j + 1, k)), ND3(i, j + 1, k)), 0.25);
                f->r = sqrt(f->c.x * f->c.x + f->c.y * f->c.y);
                f->th = atan(f->c.y / f->c.x);
                f->s = sqrt(f->S.x * f->S.x + f->S.y * f->S.y + f->S.z * f->S.z);
            }
    b->cellvol = 0.0;
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                cell_t *cl = &b->cell[C3(b, i, j, k)];
                const face_t *a0 = &b->fi[C3(b, i, j, k)], *a1 = &b->fi[C3(b, i + 1, j, k)];
                const face_t *b0 = &b->fj[C3(b, i, j, k)], *b1 = &b->fj[C3(b, i, j + 1, k)];
                const face_t *c0 = &b->fk[C3(b, i, j, k)], *c1 = &b->fk[C3(b, i, j, k + 1)];
                double vol = vdot(a1->c, a1->S) - vdot(a0->c, a0->S) + vdot(b1->c, b1->S) - vdot(b0->c, b0->S) +
                             vdot(c1->c, c1->S) - vdot(c0->c, c0->S);
                cl->vol = vol / F32(3.0);
                v3_t s = vadd(ND3(i, j, k), ND3(i + 1, j, k));
                s = vadd(s, ND3(i, j + 1, k)); s = vadd(s, ND3(i + 1, j + 1, k));
                s = vadd(s, ND3(i, j, k + 1)); s = vadd(s, ND3(i + 1, j, k + 1));
                s = vadd(s, ND3(i, j + 1, k + 1)); s = vadd(s, ND3(i + 1, j + 1, k + 1));
                cl->c = vscale(s, 0.125);
                cl->r = sqrt(cl->c.x * cl->c.x + cl->c.y * cl->c.y);
                cl->th = atan(cl->c.y / cl->c.x);
                b->cellvol = b->cellvol + cl->vol;
            }
#undef ND3
}

/* The nine coefficients from the three coordinate derivatives along kesi,
 * yita, zita (source/GeometryVariables.cpp:191-210, repeated at 252-267,
 * 310-326, 367-387).  The reference's gradz_kesi repeats the gradx_kesi
 * expression; that is kept. */
static void metric9(v3_t e1, v3_t e2, v3_t e3, double g[9])
{
    const double lim_zero = F32(1E-20);
    double partA = (e1.x * (e2.y * e3.z - e3.y * e2.z));
    double partB = (e2.x * (e1.y * e3.z - e3.y * e1.z));
    double partC = (e3.x * (e1.y * e2.z - e2.y * e1.z));
    double J1 = (partA - partB + partC);
    if (fabs(J1) <= lim_zero) J1 = lim_zero;
    double J = F32(1.0) / J1;
    g[0] = J * (e2.y * e3.z - e3.y * e2.z);
    g[1] = J * (e2.z * e3.x - e3.z * e2.x);
    g[2] = J * (e2.y * e3.z - e3.y * e2.z);
    g[3] = J * (e3.y * e1.z - e1.y * e3.z);
    g[4] = J * (e3.z * e1.x - e1.z * e3.x);
    g[5] = J * (e3.x * e1.y - e1.x * e3.y);
    g[6] = J * (e1.y * e2.z - e2.y * e1.z);
    g[7] = J * (e1.z * e2.x - e2.z * e1.x);
    g[8] = J * (e1.x * e2.y - e2.x * e1.y);
}

/* one-sided at the first and last index, central between */
static v3_t diff3(v3_t lo, v3_t mid, v3_t hi, int first, int last)
{
    if (first) return vsub(hi, mid);
    if (last) return vsub(mid, lo);
    return vscale(vsub(hi, lo), F32(0.5));
}

/* 0.5f*(a + b) - 0.5f*(c + d) */
static v3_t half_pair_diff(v3_t a, v3_t b, v3_t c, v3_t d)
{
    return vsub(vscale(vadd(a, b), F32(0.5)), vscale(vadd(c, d), F32(0.5)));
}

/* source/GeometryVariables.cpp:108-391 */
static void jacobi_convert_coeff(blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    v3_t *N = b->node;
#define ND3(i, j, k) N[C3(b, i, j, k)]
#define CC(i, j, k) b->cell[C3(b, i, j, k)].c
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                v3_t e1 = diff3(CC(i - 1, j, k), CC(i, j, k), CC(i + 1, j, k), i == 1, i == IM - 1);
                v3_t e2 = diff3(CC(i, j - 1, k), CC(i, j, k), CC(i, j + 1, k), j == 1, j == JM - 1);
                v3_t e3 = diff3(CC(i, j, k - 1), CC(i, j, k), CC(i, j, k + 1), k == 1, k == KM - 1);
                metric9(e1, e2, e3, b->cell[C3(b, i, j, k)].g);
            }
#define FC(F, i, j, k) b->F[C3(b, i, j, k)].c
    for (int i = 1; i <= IM; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                v3_t e1 = diff3(FC(fi, i - 1, j, k), FC(fi, i, j, k), FC(fi, i < IM ? i + 1 : i, j, k), i == 1, i == IM);
                v3_t e2 = half_pair_diff(ND3(i, j + 1, k), ND3(i, j + 1, k + 1), ND3(i, j, k), ND3(i, j, k + 1));
                /* the reference averages Node[i][j][k] with itself here (line 247) */
                v3_t e3 = half_pair_diff(ND3(i, j, k + 1), ND3(i, j + 1, k + 1), ND3(i, j, k), ND3(i, j, k));
                metric9(e1, e2, e3, b->fi[C3(b, i, j, k)].g);
            }
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                v3_t e1 = half_pair_diff(ND3(i + 1, j, k), ND3(i + 1, j, k + 1), ND3(i, j, k), ND3(i, j, k + 1));
                v3_t e2 = diff3(FC(fj, i, j - 1, k), FC(fj, i, j, k), FC(fj, i, j < JM ? j + 1 : j, k), j == 1, j == JM);
                v3_t e3 = half_pair_diff(ND3(i, j, k + 1), ND3(i + 1, j, k + 1), ND3(i, j, k), ND3(i + 1, j, k));
                metric9(e1, e2, e3, b->fj[C3(b, i, j, k)].g);
            }
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM; ++k) {
                v3_t e1 = half_pair_diff(ND3(i + 1, j + 1, k), ND3(i + 1, j, k), ND3(i, j, k), ND3(i, j + 1, k));
                v3_t e2 = half_pair_diff(ND3(i + 1, j + 1, k), ND3(i, j + 1, k), ND3(i + 1, j, k), ND3(i, j, k));
                v3_t e3 = diff3(FC(fk, i, j, k - 1), FC(fk, i, j, k), FC(fk, i, j, k < KM ? k + 1 : k), k == 1, k == KM);
                metric9(e1, e2, e3, b->fk[C3(b, i, j, k)].g);
            }
#undef FC
#undef CC
#undef ND3
}

/* ------------------------------------------------------- grid levels */
/* One patch range of CUserInput::NumberOfMultigrid (source/UserInput.cpp:
 * 816-844): tangential cell ranges become node ranges (end + 1, unless the range
 * is a single index), and the level halves only if every start and end node is
 * odd.  The block loop of the same function (1020-1040) never lowers the count
 * below two, so with the two-level cap only the patches decide. */
static int range_coarsens(const int *r)
{
    int q[6];
    memcpy(q, r, sizeof q);
    for (int a = 0; a < 3; ++a)
        if (q[2 * a] != q[2 * a + 1]) q[2 * a + 1] = q[2 * a + 1] + 1;
    for (int e = 0; e < 6; ++e)
        if ((q[e] - 1) % 2 != 0) return 0;
    return 1;
}

/* Patch ranges of the coarse level (source/UserInput.cpp:446-797): a single
 * index n becomes (n+1)/2, a cell range s..e becomes (s+1)/2 .. e/2. */
static void coarse_range(const int *r, int *o)
{
    for (int a = 0; a < 3; ++a) {
        int s = r[2 * a], e = r[2 * a + 1];
        o[2 * a] = (s + 1) / 2;
        o[2 * a + 1] = s == e ? (e + 1) / 2 : e / 2;
    }
}

static int build_coarse(orc_ctx *c)
{
    orc_config cc = c->cfg;
    cc.multigrid = 0;
    cc.turbulence = 0;   /* the turbulence model runs on the fine level only (source/ComputeAllBlockVars.cpp:230) */
    orc_ctx *m = orc_create(&cc);
    for (int n = 0; n < c->nb; ++n) {
        const blk_t *b = &c->blk[n];
        /* source/UserInput.cpp:296-331 */
        int IM = (b->IM + 1) / 2, JM = (b->JM + 1) / 2, KM = (b->KM + 1) / 2;
        double *nodes = (double *)malloc(sizeof(double) * 3 * (size_t)IM * JM * KM);
        for (int i = 1; i <= IM; ++i)
            for (int j = 1; j <= JM; ++j)
                for (int k = 1; k <= KM; ++k) {
                    v3_t v = b->node[C3(b, 2 * i - 1, 2 * j - 1, 2 * k - 1)];
                    double *p = nodes + (((long)(i - 1) * JM + (j - 1)) * KM + (k - 1)) * 3;
                    p[0] = v.x; p[1] = v.y; p[2] = v.z;
                }
        orc_block_add(m, IM, JM, KM, nodes, b->nblade, 0.0);
        free(nodes);
        blk_t *mb = &m->blk[n];
        /* source/UserInput.cpp:449-454 copies the pitch and the rotation speed */
        mb->w_rot = b->w_rot;
        mb->delta_th = b->delta_th;
        long box = (long)(IM + 1) * (JM + 1) * (KM + 1);
        mb->cv2h = (cv_t *)calloc(box, sizeof(cv_t));
        mb->res2h = (cv_t *)calloc(box, sizeof(cv_t));
        mb->delt = (cv_t *)calloc(box, sizeof(cv_t));
    }
    for (int n = 0; n < c->nb; ++n)
        for (int q = 0; q < c->blk[n].npatch; ++q) {
            const patch_t *P = &c->blk[n].patch[q];
            int r[6] = {P->is, P->ie, P->js, P->je, P->ks, P->ke}, o[6];
            coarse_range(r, o);
            orc_patch_add(m, P->kind, n + 1, P->face, o, P->rotate);
        }
    m->first_in[0] = c->first_in[0]; m->first_in[1] = c->first_in[1];
    m->first_out[0] = c->first_out[0]; m->first_out[1] = c->first_out[1];
    for (int t = 0; t < c->nconn; ++t) {
        const conn_t *T = &c->conn[t];
        int o1[6], o2[6];
        coarse_range(T->rng[0], o1);
        coarse_range(T->rng[1], o2);
        if (orc_connection_add(m, T->kind, T->orient, T->blk[0] + 1, T->face[0], o1, T->blk[1] + 1, T->face[1], o2))
            return -1;
    }
    for (int q = 0; q < c->nmix; ++q) {
        int o1[6], o2[6];
        coarse_range(c->mix[q].rng[0], o1);
        coarse_range(c->mix[q].rng[1], o2);
        if (orc_mixing_plane_add(m, c->mix[q].blk[0] + 1, o1, c->mix[q].blk[1] + 1, o2)) return -1;
    }
    orc_geometry(m);
    c->coarse = m;
    return 0;
}

int orc_geometry(orc_ctx *c)
{
    c->total_vol = 0.0;
    for (int n = 0; n < c->nb; ++n) {
        face_vector_volume(&c->blk[n]);
        c->total_vol = c->total_vol + c->blk[n].cellvol;   /* source/ComputeAllBlockVars.cpp:25-36 */
    }
    for (int n = 0; n < c->nb; ++n) jacobi_convert_coeff(&c->blk[n]);
    if (c->cfg.turbulence == 200) wall_distance(c);
    c->ng = 1;
    if (c->cfg.multigrid) {
        int ok = 1;
        for (int n = 0; n < c->nb; ++n)
            for (int q = 0; q < c->blk[n].npatch; ++q) {
                const patch_t *P = &c->blk[n].patch[q];
                int r[6] = {P->is, P->ie, P->js, P->je, P->ks, P->ke};
                ok = ok && range_coarsens(r);
            }
        /* the two-sided patches are read with their first side only (source/UserInput.cpp:932, 962) */
        for (int t = 0; t < c->nconn; ++t) ok = ok && range_coarsens(c->conn[t].rng[0]);
        if (ok) {
            if (build_coarse(c)) return -1;
            c->ng = 2;
        }
    }
    return 0;
}

int orc_levels(orc_ctx *c) { return c->ng; }
orc_ctx *orc_coarse(orc_ctx *c) { return c->coarse; }

/* ------------------------------------------------------- start field */
/* CBlocks::Initial_Uniform (source/Initialization.cpp:17-72) */
static void initial_uniform(orc_ctx *c, blk_t *b)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas, Ptin = c->nd.pt_in, Ttin = c->nd.tt_in;
    double Ma0 = (float)c->cfg.ma0;
    const double *dir = c->cfg.inlet_dir;
    double m2 = F32(1.0) + (Gamma - F32(1.0)) / F32(2.0) * Ma0 * Ma0;
    double press = Ptin * pow(m2, F32(0.0) - Gamma / (Gamma - F32(1.0)));
    double temp = Ttin * pow(m2, (double)(0.0f - 1.0f));
    double dens = press / (Rgas * temp);
    double Vel = pow(Gamma * Rgas * temp, F32(0.5)) * Ma0;
    double n_m = pow(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2], 0.5);
    for (int i = 1; i <= b->IM - 1; ++i)
        for (int j = 1; j <= b->JM - 1; ++j)
            for (int k = 1; k <= b->KM - 1; ++k) {
                long x = C3(b, i, j, k);
                double u = Vel * dir[0] / n_m, v = Vel * dir[1] / n_m, w = Vel * dir[2] / n_m;
                pv_t p = {dens, u, v, w, press, temp};
                b->pv[x] = p;
                double V2 = u * u + v * v + w * w;
                double ener = press / (dens * (Gamma - F32(1.0))) + V2 / F32(2.0);
                cv_t q = {dens, dens * u, dens * v, dens * w, dens * ener};
                b->cv[x] = q;
                if (c->cfg.turbulence == 200) b->ev[x] = c->nd.ev_ratio * F32(1.0);
            }
}

/* CMesh::Initial_Interpolate (source/Initialization.cpp:74-165): pressure and
 * flow direction linear in z between the first inlet and the first outlet. */
static void initial_interpolate(orc_ctx *c)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas, Ptin = c->nd.pt_in, Ttin = c->nd.tt_in, Pout = c->nd.p_out;
    double P21ratio = (float)c->cfg.p21ratio;
    const double *din = c->cfg.inlet_dir, *dout = c->cfg.outlet_dir;
    const blk_t *bi = &c->blk[c->first_in[0]], *bo = &c->blk[c->first_out[0]];
    int ks = bi->patch[c->first_in[1]].ks, ke = bo->patch[c->first_out[1]].ke;
    double zmin = F32(1E+32), zmax = F32(-1E+32);
    for (int i = 1; i <= bi->IM - 1; ++i)
        for (int j = 1; j <= bi->JM - 1; ++j) {
            double z = bi->fk[C3(bi, i, j, ks)].c.z;
            zmin = z < zmin ? z : zmin;
        }
    for (int i = 1; i <= bo->IM - 1; ++i)
        for (int j = 1; j <= bo->JM - 1; ++j) {
            double z = bo->fk[C3(bo, i, j, ke)].c.z;
            zmax = z > zmax ? z : zmax;
        }
    double dz = zmax - zmin;
    double Pin = Pout / P21ratio;
    if (Pin >= Ptin) Pin = F32(0.99) * Ptin;
    double dp = Pout - Pin;
    double dB[3] = {dout[0] - din[0], dout[1] - din[1], dout[2] - din[2]};
    double temp = Ttin * pow(Pin / Ptin, (Gamma - F32(1.0)) / Gamma);
    double dens0 = Pin / (Rgas * temp);
    double cs = sqrt(Gamma * Pin / dens0);
    double Ma = sqrt(F32(2.0) * ((Ttin / temp) - F32(1.0)) / (Gamma - F32(1.0)));
    double Vel = Ma * cs;
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    double z = b->cell[x].c.z;
                    double B[3];
                    for (int a = 0; a < 3; ++a) B[a] = din[a] + dB[a] * (z - zmin) / dz;
                    double press = Pin + dp * (z - zmin) / dz;
                    double dens = press / Rgas / temp;
                    double n_m = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
                    double u = Vel * B[0] / n_m, v = Vel * B[1] / n_m, w = Vel * B[2] / n_m;
                    pv_t p = {dens, u, v, w, press, temp};
                    b->pv[x] = p;
                    double V2 = u * u + v * v + w * w;
                    double ener = press / (Gamma - F32(1.0)) + dens * V2 / F32(2.0);
                    cv_t q = {dens, dens * u, dens * v, dens * w, ener};
                    b->cv[x] = q;
                    if (c->cfg.turbulence == 200) b->ev[x] = c->nd.ev_ratio * F32(1.0);
                }
    }
}

static void initial_level(orc_ctx *m, int type)
{
    if (type == 0)
        for (int n = 0; n < m->nb; ++n) initial_uniform(m, &m->blk[n]);
    else
        initial_interpolate(m);
}

int orc_initial(orc_ctx *c)
{
    orc_ctx *m = c->ng == 2 ? c->coarse : c;
    if (m->first_in[0] < 0 || m->first_out[0] < 0) return -1;
    initial_level(m, c->cfg.initial_type);
    /* with Spalart-Allmaras the fine level starts from its own field too
     * (source/Structured_FV3D.cpp:89-95); its state is replaced at Coarse_iter,
     * its nu~ is not */
    if (m != c && c->cfg.turbulence == 200) initial_level(c, c->cfg.initial_type);
    return 0;
}

/* nu~ of one block from the restart file (source/UserInput.cpp:1134-1153) */
int orc_set_nut(orc_ctx *c, int block, const double *ev)
{
    if (block < 1 || block > c->nb) return -1;
    blk_t *b = &c->blk[block - 1];
    int ni = b->IM - 1, nj = b->JM - 1, nk = b->KM - 1;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            for (int k = 1; k <= nk; ++k) b->ev[C3(b, i, j, k)] = ev[((long)(i - 1) * nj + (j - 1)) * nk + (k - 1)];
    return 0;
}

int orc_set_state(orc_ctx *c, int block, const double *pv, const double *ener)
{
    if (block < 1 || block > c->nb) return -1;
    blk_t *b = &c->blk[block - 1];
    int ni = b->IM - 1, nj = b->JM - 1, nk = b->KM - 1;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            for (int k = 1; k <= nk; ++k) {
                long s = ((long)(i - 1) * nj + (j - 1)) * nk + (k - 1);
                const double *w = pv + s * 6;
                long x = C3(b, i, j, k);
                /* source/UserInput.cpp:1139-1150 */
                b->cv[x].r = w[0];
                b->cv[x].ru = w[0] * w[1];
                b->cv[x].rv = w[0] * w[2];
                b->cv[x].rw = w[0] * w[3];
                b->cv[x].re = ener[s];
                pv_t p = {w[0], w[1], w[2], w[3], w[4], w[5]};
                b->pv[x] = p;
            }
    return 0;
}

/* ---------------------------------------------------------- boundary fill */
static void set_dummies(blk_t *b, int d, int a, int bb, cv_t cv, pv_t pv)
{
    for (int nd = 1; nd <= ND; ++nd) {
        long x;
        if (d < 2) x = DI(b, d, nd, a, bb);
        else if (d < 4) x = DI(b, d, a, nd, bb);
        else x = DI(b, d, a, bb, nd);
        b->dcv[d][x] = cv;
        b->dpv[d][x] = pv;
    }
}

/* source/BndCondInlet.cpp:7-127 */
static void bc_inlet(orc_ctx *c, blk_t *b, const patch_t *P)
{
    const orc_nondim *nd = &c->nd;
    double Gamma = nd->gamma, Rgas = nd->rgas, Ttin = nd->tt_in, Ptin = nd->pt_in;
    const double *A = nd->inlet_dir;
    int KM = b->KM;
    double Vn_m = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
    if (P->ks != P->ke) return;
    for (int i = P->is; i <= P->ie; ++i)
        for (int j = P->js; j <= P->je; ++j) {
            int k, kv;
            double uvel, vvel, wvel, vdn, n_m;
            v3_t S;
            if (P->ks == 1) {
                k = 1; kv = 0;
                S = b->fk[C3(b, i, j, k)].S;
                n_m = sqrt(S.x * S.x + S.y * S.y + S.z * S.z);
                pv_t *w = &b->pv[C3(b, i, j, k)];
                uvel = w->u; vvel = w->v; wvel = w->w;
                vdn = F32(-1.0) * (uvel * S.x / n_m + vvel * S.y / n_m + wvel * S.z / n_m);
            } else if (P->ks == KM) {
                k = KM - 1; kv = KM;
                S = b->fk[C3(b, i, j, kv)].S;
                n_m = sqrt(S.x * S.x + S.y * S.y + S.z * S.z);
                pv_t *w = &b->pv[C3(b, i, j, k)];
                uvel = w->u; vvel = w->v; wvel = w->w;
                vdn = (uvel * S.x / n_m + vvel * S.y / n_m + wvel * S.z / n_m);
            } else {
                continue;
            }
            double c_d = sqrt(Gamma * Rgas * b->pv[C3(b, i, j, k)].t);
            double V2_d = uvel * uvel + vvel * vvel + wvel * wvel;
            double costh = F32(-1.0) * vdn / sqrt(V2_d);
            double c_02 = c_d * c_d + (Gamma - F32(1.0)) * V2_d / F32(2.0);
            double R_minus = vdn - F32(2.0) * c_d / (Gamma - F32(1.0));
            double meduA = (-R_minus * (Gamma - F32(1.0))) / ((Gamma - F32(1.0)) * costh * costh + F32(2.0));
            double meduB = sqrt(c_02 * ((Gamma - F32(1.0)) * costh * costh + F32(2.0)) / (Gamma - F32(1.0)) / (R_minus * R_minus) -
                                (Gamma - F32(1.0)) / F32(2.0));
            double c_b = meduA * (F32(1.0) + costh * meduB);
            double TempBC = Ttin * (c_b * c_b / c_02);
            if (TempBC > Ttin) TempBC = F32(0.99) * Ttin;
            double PressBC = Ptin * pow(TempBC / Ttin, Gamma / (Gamma - F32(1.0)));
            double DensBC = PressBC / Rgas / TempBC;
            double VBC = sqrt(F32(2.0) * Gamma / (Gamma - 1) * Rgas * (Ttin - TempBC));
            long g = C3(b, i, j, kv);
            pv_t *pg = &b->pv[g];
            pg->u = VBC * A[0] / Vn_m;
            pg->v = VBC * A[1] / Vn_m;
            pg->w = VBC * A[2] / Vn_m;
            uvel = pg->u; vvel = pg->v; wvel = pg->w;
            V2_d = uvel * uvel + vvel * vvel + wvel * wvel;
            double ener = PressBC / DensBC / (Gamma - F32(1.0)) + V2_d / F32(2.0);
            pg->p = PressBC; pg->t = TempBC; pg->d = DensBC;
            cv_t *cg = &b->cv[g];
            cg->r = DensBC; cg->ru = DensBC * uvel; cg->rv = DensBC * vvel; cg->rw = DensBC * wvel; cg->re = DensBC * ener;
            if (c->cfg.turbulence == 200) b->ev[g] = c->nd.ev_ratio * F32(1.0);   /* source/BndCondInlet.cpp:104-105 */
            set_dummies(b, P->ks == 1 ? 4 : 5, i, j, *cg, *pg);
        }
}

/* source/BndCondOutlet.cpp:116-305 (centrifugal = general without the supersonic branch) */
static void bc_outlet_general(orc_ctx *c, blk_t *b, const patch_t *P, int centrifugal)
{
    const orc_nondim *nd = &c->nd;
    double Gamma = nd->gamma, Rgas = nd->rgas, Pout = nd->p_out;
    int KM = b->KM;
    if (P->ks != P->ke) return;
    for (int i = P->is; i <= P->ie; ++i)
        for (int j = P->js; j <= P->je; ++j) {
            int k, kv;
            v3_t norm;
            if (P->ks == 1) { k = 1; kv = 0; const face_t *f = &b->fk[C3(b, i, j, k)]; norm = vdiv(f->S, f->s); }
            else if (P->ks == KM) { k = KM - 1; kv = KM; const face_t *f = &b->fk[C3(b, i, j, kv)]; norm = vdiv(f->S, f->s); }
            else continue;
            pv_t *pi = &b->pv[C3(b, i, j, k)];
            pv_t *pg = &b->pv[C3(b, i, j, kv)];
            double dens0 = pi->d;
            double c0 = sqrt(Gamma * Rgas * pi->t);
            double Vn = pi->u * norm.x + pi->v * norm.y + pi->w * norm.z;
            double Ma_n = Vn / c0;
            if (centrifugal || Ma_n <= 1.0) {
                pg->p = Pout;
                pg->d = pi->d + (Pout - pi->p) / c0 / c0;
                pg->u = pi->u + norm.x * (pi->p - Pout) / dens0 / c0;
                pg->v = pi->v + norm.y * (pi->p - Pout) / dens0 / c0;
                pg->w = pi->w + norm.z * (pi->p - Pout) / dens0 / c0;
            } else {
                pg->p = pi->p; pg->d = pi->d; pg->u = pi->u; pg->v = pi->v; pg->w = pi->w;
            }
            pg->t = pg->p / pg->d / Rgas;
            double V2_d = pg->u * pg->u + pg->v * pg->v + pg->w * pg->w;
            double ener = pg->p / pg->d / (Gamma - F32(1.0)) + V2_d / F32(2.0);
            cv_t *cg = &b->cv[C3(b, i, j, kv)];
            cg->r = pg->d; cg->ru = pg->d * pg->u; cg->rv = pg->d * pg->v; cg->rw = pg->d * pg->w; cg->re = pg->d * ener;
            /* source/BndCondOutlet.cpp:180-181, 282-283 */
            if (c->cfg.turbulence == 200) b->ev[C3(b, i, j, kv)] = b->ev[C3(b, i, j, k)];
            set_dummies(b, P->ks == 1 ? 4 : 5, i, j, *cg, *pg);
        }
}

/* source/BndCondOutlet.cpp:13-113 (radial equilibrium, K-max faces only) */
static void bc_outlet_axial(orc_ctx *c, blk_t *b, const patch_t *P)
{
    const orc_nondim *nd = &c->nd;
    double Gamma = nd->gamma, Rgas = nd->rgas, Pout = nd->p_out;
    int IM = b->IM, JM = b->JM, KM = b->KM;
    if (P->ks != P->ke || P->ke != KM) return;
    int k = KM - 1;
    double cu[1024];
    if (IM > 1000) return;
    for (int i = 1; i <= IM - 1; ++i) {
        cu[i] = 0.0;
        double S_temp = 0.0;
        for (int j = 1; j <= JM - 1; ++j) {
            double uvel = b->pv[C3(b, i, j, k)].u, vvel = b->pv[C3(b, i, j, k)].v;
            const face_t *f = &b->fk[C3(b, i, j, KM)];
            double cu_temp = F32(-1.0) * uvel * sin(f->th) + vvel * cos(f->th);
            cu[i] = cu[i] + cu_temp * f->s;
            S_temp = S_temp + f->s;
        }
        cu[i] = cu[i] / S_temp;
    }
    int iref = (int)(0.5 * IM);
    for (int j = 1; j <= JM - 1; ++j) {
        b->pv[C3(b, iref, j, KM)].p = Pout;
        for (int i = iref - 1; i >= 1; --i) {
            double rn = b->fk[C3(b, i, j, KM)].r, rn_1 = b->fk[C3(b, i + 1, j, KM)].r;
            double dp_dr = b->cv[C3(b, i, j, KM)].r * cu[i] * cu[i] / rn;
            b->pv[C3(b, i, j, KM)].p = b->pv[C3(b, i + 1, j, KM)].p - dp_dr * (rn_1 - rn);
        }
        for (int i = iref + 1; i <= IM - 1; ++i) {
            double rn_1 = b->fk[C3(b, i, j, k)].r, rn = b->fk[C3(b, i - 1, j, k)].r;
            double dp_dr = b->cv[C3(b, i, j, KM)].r * cu[i] * cu[i] / rn;
            b->pv[C3(b, i, j, KM)].p = b->pv[C3(b, i - 1, j, KM)].p + dp_dr * (rn_1 - rn);
        }
    }
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j) {
            cv_t *cg = &b->cv[C3(b, i, j, KM)];
            const cv_t *ci = &b->cv[C3(b, i, j, KM - 1)];
            pv_t *pg = &b->pv[C3(b, i, j, KM)];
            const pv_t *pi = &b->pv[C3(b, i, j, KM - 1)];
            cg->r = ci->r; cg->ru = ci->ru; cg->rv = ci->rv; cg->rw = ci->rw;
            pg->d = ci->r; pg->u = pi->u; pg->v = pi->v; pg->w = pi->w;
            pg->t = pg->p / Rgas / cg->r;
            double V2 = pg->u * pg->u + pg->v * pg->v + pg->w * pg->w;
            cg->re = pg->p / (Gamma - F32(1.0)) + pg->d * V2 / F32(2.0);
            /* source/BndCondOutlet.cpp:102-103 */
            if (c->cfg.turbulence == 200) b->ev[C3(b, i, j, KM)] = b->ev[C3(b, i, j, KM - 1)];
            set_dummies(b, 5, i, j, *cg, *pg);
        }
}

/* source/BndCondWall.cpp:20-284 */
static void bc_wall(blk_t *b, const patch_t *P, int sa)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    double rot = P->rot;
    int dir, lo;
    if (P->is == P->ie) { dir = 0; lo = (P->is == 1); if (!lo && P->is != IM) return; }
    else if (P->js == P->je) { dir = 1; lo = (P->js == 1); if (!lo && P->js != JM) return; }
    else if (P->ks == P->ke) { dir = 2; lo = (P->ks == 1); if (!lo && P->ks != KM) return; }
    else return;
    int aS = dir == 0 ? P->js : P->is, aE = dir == 0 ? P->je : P->ie;
    int bS = dir == 2 ? P->js : P->ks, bE = dir == 2 ? P->je : P->ke;
    for (int a = aS; a <= aE; ++a)
        for (int bb = bS; bb <= bE; ++bb) {
            int i, j, k, iv, jv, kv;
            double x, y;
            const face_t *f;
            if (dir == 0) { j = a; k = bb; i = lo ? 1 : IM - 1; iv = lo ? 0 : IM; jv = j; kv = k; f = &b->fi[C3(b, lo ? 1 : IM, j, k)]; }
            else if (dir == 1) { i = a; k = bb; j = lo ? 1 : JM - 1; jv = lo ? 0 : JM; iv = i; kv = k; f = &b->fj[C3(b, i, lo ? 1 : JM, k)]; }
            else { i = a; j = bb; k = lo ? 1 : KM - 1; kv = lo ? 0 : KM; iv = i; jv = j; f = &b->fk[C3(b, i, j, lo ? 1 : KM)]; }
            x = f->c.x; y = f->c.y;
            pv_t *pi = &b->pv[C3(b, i, j, k)];
            pv_t *pg = &b->pv[C3(b, iv, jv, kv)];
            cv_t *ci = &b->cv[C3(b, i, j, k)];
            cv_t *cg = &b->cv[C3(b, iv, jv, kv)];
            pg->p = pi->p; pg->t = pi->t; pg->d = pi->d;
            double Vwall_x = -rot * y, Vwall_y = rot * x;
            v3_t Vin = {pi->u - Vwall_x, pi->v - Vwall_y, pi->w - F32(0.0)};
            v3_t Vout = {F32(0.0) - Vin.x, F32(0.0) - Vin.y, F32(0.0) - Vin.z};
            pg->u = Vout.x + Vwall_x;
            pg->v = Vout.y + Vwall_y;
            pg->w = Vout.z + F32(0.0);
            cg->r = ci->r;
            cg->ru = pg->u * cg->r;
            cg->rv = pg->v * cg->r;
            cg->rw = pg->w * cg->r;
            cg->re = ci->re;
            if (sa) b->ev[C3(b, iv, jv, kv)] = F32(0.0);   /* source/BndCondWall.cpp:104-107 */
            int d = 2 * dir + (lo ? 0 : 1);
            pv_t pd = kmulp(0.5, padd(*pg, *pi));
            cv_t cd = cscale(cadd(*cg, *ci), 0.5);
            set_dummies(b, d, a, bb, cd, pd);
        }
}

/* source/BoundaryCondition.cpp:9-63 */
/* ------------------------------------------------ interior / periodic */
/* Position (l, m) of the 2-D shadow array of face fc: the first and second
 * tangential index it runs over (source/ShadowOutIn.cpp:12-164): i- (k, j),
 * i+ (j, k), j- (i, k), j+ (k, i), k- (j, i), k+ (i, j). */
static void shadow_axes(int fc, int *al, int *am)
{
    static const int L[6] = {2, 1, 0, 2, 1, 0}, Mx[6] = {1, 2, 2, 0, 0, 1};
    *al = L[fc - 1];
    *am = Mx[fc - 1];
}

/* cell index of layer nd next to face fc at shadow position (l, m) (1-based);
 * nd = 0 is the ghost cell */
static long shadow_cell(const blk_t *b, int fc, int nd, const int *r, int l, int m)
{
    int al, am, x[3];
    shadow_axes(fc, &al, &am);
    x[al] = r[2 * al] + l - 1;
    x[am] = r[2 * am] + m - 1;
    int d = (fc - 1) / 2, N = d == 0 ? b->IM : d == 1 ? b->JM : b->KM;
    x[d] = fc % 2 == 1 ? nd : N - nd;
    return C3(b, x[0], x[1], x[2]);
}

static void shadow_extent(int fc, const int *r, int *L, int *M)
{
    int al, am;
    shadow_axes(fc, &al, &am);
    *L = r[2 * al + 1] - r[2 * al] + 1;
    *M = r[2 * am + 1] - r[2 * am] + 1;
}

/* CMesh::ShadowOut, Trans, ShadowIn (source/ShadowOutIn.cpp): layer nd of the
 * cells of side s next to its face goes, re-indexed by the orientation and
 * rotated by dr about z, into layer nd of the other side's Dummy slab (and its
 * ghost cell for nd = 1). */
static void shadow_copy(orc_ctx *c, const conn_t *t, int s, int nd, double dr)
{
    const blk_t *A = &c->blk[t->blk[s]];
    blk_t *B = &c->blk[t->blk[1 - s]];
    int fa = t->face[s], fb = t->face[1 - s];
    int L1, M1;
    shadow_extent(fa, t->rng[s], &L1, &M1);
    int L2 = (t->orient == 2 || t->orient == 4) ? M1 : L1, M2 = (t->orient == 2 || t->orient == 4) ? L1 : M1;
    for (int l1 = 1; l1 <= L1; ++l1)
        for (int m1 = 1; m1 <= M1; ++m1) {
            int l2, m2;
            switch (t->orient) {
            case 1: l2 = l1; m2 = M1 + 1 - m1; break;
            case 2: l2 = m1; m2 = l1; break;
            case 3: l2 = L1 + 1 - l1; m2 = m1; break;
            default: l2 = M1 + 1 - m1; m2 = L1 + 1 - l1; break;
            }
            (void)L2; (void)M2;
            long xs = shadow_cell(A, fa, nd, t->rng[s], l1, m1);
            cv_t q = A->cv[xs];
            pv_t w = A->pv[xs];
            if (dr != 0.0) {
                cv_t q0 = q;
                pv_t w0 = w;
                q.ru = q0.ru * cos(dr) - q0.rv * sin(dr);
                q.rv = q0.rv * cos(dr) + q0.ru * sin(dr);
                w.u = w0.u * cos(dr) - w0.v * sin(dr);
                w.v = w0.v * cos(dr) + w0.u * sin(dr);
            }
            /* the target's own (l, m) axes */
            int al, am, x[3];
            shadow_axes(fb, &al, &am);
            x[al] = t->rng[1 - s][2 * al] + l2 - 1;
            x[am] = t->rng[1 - s][2 * am] + m2 - 1;
            int d = (fb - 1) / 2;
            x[d] = nd;
            long xd = DI(B, fb - 1, x[0], x[1], x[2]);
            B->dcv[fb - 1][xd] = q;
            B->dpv[fb - 1][xd] = w;
            if (nd == 1) {
                long xg = shadow_cell(B, fb, 0, t->rng[1 - s], l2, m2);
                B->cv[xg] = q;
                B->pv[xg] = w;
                if (c->cfg.turbulence == 200) B->ev[xg] = A->ev[xs];   /* source/ShadowOutIn.cpp:339-345 */
            }
        }
}

/* CMesh::BndCondInterior / BndCondPeriodic (source/BndCondInteriorPeriodic.cpp) */
static void bc_connections(orc_ctx *c)
{
    for (int kind = 0; kind < 2; ++kind)
        for (int q = 0; q < c->nconn; ++q) {
            const conn_t *t = &c->conn[q];
            if (t->kind != kind) continue;
            blk_t *A = &c->blk[t->blk[0]];
            if (kind == 1 && t->flip) A->delta_th = F32(0.0) - A->delta_th;
            double d1 = kind == 1 ? A->delta_th : 0.0, d2 = kind == 1 ? F32(0.0) - A->delta_th : 0.0;
            for (int nd = 1; nd <= ND; ++nd) {
                shadow_copy(c, t, 0, nd, d1);
                shadow_copy(c, t, 1, nd, d2);
            }
        }
}

/* CMesh::BndCondMixingPlane (source/BndCondMixingPlane.cpp:287-522): per
 * radial line i, the area-weighted pitch-wise averages of the mass, momentum,
 * energy (and rho nu~) fluxes through the K faces on each side, the mixed-out
 * state that carries them (pressure from the quadratic of lines 370 / 432),
 * then characteristic ghost states: block 1's K-high ghosts from the
 * downstream average plus its outgoing waves, block 2's K-low ghosts from the
 * upstream average plus its outgoing wave. */
typedef struct { pv_t pv; double c, ev; } mixavg_t;

static mixavg_t mix_average(const orc_ctx *c, const blk_t *b, int i, int js, int je, int kf, int kc)
{
    const double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    double I[7] = {0, 0, 0, 0, 0, 0, 0}, s_sum = 0.0;
    v3_t n_sum = {0.0, 0.0, 0.0};
    for (int j = js; j <= je; ++j) {
        const face_t *f = &b->fk[C3(b, i, j, kf)];
        v3_t n = {f->S.x / f->s, f->S.y / f->s, f->S.z / f->s};
        double ts = f->s;
        n_sum.x = n_sum.x + n.x * ts;
        n_sum.y = n_sum.y + n.y * ts;
        n_sum.z = n_sum.z + n.z * ts;
        s_sum = s_sum + ts;
        const pv_t *p = &b->pv[C3(b, i, j, kc)];
        const cv_t *q = &b->cv[C3(b, i, j, kc)];
        double Vn = p->u * n.x + p->v * n.y + p->w * n.z;
        I[1] = I[1] + q->r * Vn * ts;
        I[2] = I[2] + (q->ru * Vn + n.x * p->p) * ts;
        I[3] = I[3] + (q->rv * Vn + n.y * p->p) * ts;
        I[4] = I[4] + (q->rw * Vn + n.z * p->p) * ts;
        I[5] = I[5] + (q->re + p->p) * Vn * ts;
        if (c->cfg.turbulence == 200) I[6] = I[6] + q->r * b->ev[C3(b, i, j, kc)] * Vn * ts;
    }
    mixavg_t a;
    for (int m = 1; m <= 5; ++m) I[m] = I[m] / s_sum;
    a.ev = 0.0;
    if (c->cfg.turbulence == 200) {
        I[6] = I[6] / s_sum;
        a.ev = I[6] / I[1];
    }
    v3_t nu = {n_sum.x / s_sum, n_sum.y / s_sum, n_sum.z / s_sum};
    double tA = I[2] * nu.x + I[3] * nu.y + I[4] * nu.z;
    double tB = F32(2.0) * I[1] * I[5] - (I[2] * I[2] + I[3] * I[3] + I[4] * I[4]);
    double tC = nu.x * nu.x + nu.y * nu.y + nu.z * nu.z;
    a.pv.p = (tA / tC + sqrt(tA * tA / tC / tC - (Gamma * Gamma - F32(1.0)) * tB / tC)) / (Gamma + F32(1.0));
    a.pv.u = (I[2] - a.pv.p * nu.x) / I[1];
    a.pv.v = (I[3] - a.pv.p * nu.y) / I[1];
    a.pv.w = (I[4] - a.pv.p * nu.z) / I[1];
    a.pv.d = I[1] / (a.pv.u * nu.x + a.pv.v * nu.y + a.pv.w * nu.z);
    a.pv.t = a.pv.p / Rgas / a.pv.d;
    a.c = sqrt(Gamma * Rgas * a.pv.t);
    return a;
}

static void mix_ghost(const orc_ctx *c, blk_t *b, int side, int i, int j, int kg, pv_t g, double ev)
{
    const double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    g.t = g.p / Rgas / g.d;
    double V2 = g.u * g.u + g.v * g.v + g.w * g.w;
    cv_t q = {g.d, g.d * g.u, g.d * g.v, g.d * g.w, g.p / (Gamma - F32(1.0)) + g.d * V2 / F32(2.0)};
    long x = C3(b, i, j, kg);
    b->pv[x] = g;
    b->cv[x] = q;
    if (c->cfg.turbulence == 200) b->ev[x] = ev;
    set_dummies(b, side, i, j, q, g);
}

static void bc_mixing(orc_ctx *c)
{
    for (int q = 0; q < c->nmix; ++q) {
        blk_t *U = &c->blk[c->mix[q].blk[0]], *D = &c->blk[c->mix[q].blk[1]];
        const int *r1 = c->mix[q].rng[0], *r2 = c->mix[q].rng[1];
        const int KM = U->KM;
        for (int i = r1[0]; i <= r1[1]; ++i) {
            mixavg_t u = mix_average(c, U, i, r1[2], r1[3], KM, KM - 1);
            mixavg_t d = mix_average(c, D, i, r2[2], r2[3], 1, 1);
            for (int j = r1[2]; j <= r1[3]; ++j) {   /* lines 460-491 */
                const pv_t *p = &U->pv[C3(U, i, j, KM - 1)];
                double C1 = -u.c * u.c * (p->d - u.pv.d) + (p->p - u.pv.p);
                double C2 = u.pv.d * u.c * (p->w - u.pv.w) + (p->p - u.pv.p);
                double C5 = 0.0;
                pv_t g;
                g.d = d.pv.d + (-C1 + (C2 + C5) / F32(2.0)) / (d.c * d.c);
                g.w = d.pv.w + (C2 - C5) / (F32(2.0) * d.pv.d * d.c);
                g.v = p->v;
                g.u = p->u;
                g.p = d.pv.p + (C2 + C5) / F32(2.0);
                mix_ghost(c, U, 5, i, j, KM, g, u.ev);
            }
            for (int j = r2[2]; j <= r2[3]; ++j) {   /* lines 494-522 */
                const pv_t *p = &D->pv[C3(D, i, j, 1)];
                double C1 = 0.0, C2 = 0.0, C3 = 0.0, C4 = 0.0;
                double C5 = (F32(0.0) - d.pv.d * d.c * (p->w - d.pv.w) + (p->p - d.pv.p));
                pv_t g;
                g.d = u.pv.d + (-C1 + (C2 + C5) / F32(2.0)) / (u.c * u.c);
                g.w = u.pv.w - C5 / (F32(2.0) * u.pv.d * u.c);
                g.v = u.pv.v + C3 / (u.pv.d * u.c);
                g.u = u.pv.u + C4 / (u.pv.d * u.c);
                g.p = u.pv.p + (C2 + C5) / F32(2.0);
                mix_ghost(c, D, 4, i, j, 0, g, u.ev);
            }
        }
    }
}

int orc_bc(orc_ctx *c)
{
    for (int pass = 0; pass < 3; ++pass)
        for (int n = 0; n < c->nb; ++n) {
            blk_t *b = &c->blk[n];
            for (int q = 0; q < b->npatch; ++q) {
                patch_t *P = &b->patch[q];
                if (P->kind != pass) continue;
                if (pass == 0) bc_inlet(c, b, P);
                else if (pass == 1) {
                    if (c->cfg.turbo_type == 0) bc_outlet_axial(c, b, P);
                    else bc_outlet_general(c, b, P, c->cfg.turbo_type == 1);
                } else {
                    P->rot = P->rotate ? b->w_rot : 0.0;
                    bc_wall(b, P, c->cfg.turbulence == 200);
                }
            }
        }
    bc_connections(c);
    bc_mixing(c);
    return 0;
}

/* ------------------------------------------------------- spectral radii */
/* source/Spectral_curl_grad.cpp:9-42 */
int orc_spectral(orc_ctx *c)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        double w_rot = b->w_rot;
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    const pv_t *w = &b->pv[x];
                    const cell_t *cl = &b->cell[x];
                    double uvelr = w->u + w_rot * cl->c.y;
                    double vvelr = w->v - w_rot * cl->c.x;
                    double wvelr = w->w;
                    double csoun = sqrt(Gamma * Rgas * w->t);
                    const face_t *fa[3][2] = {{&b->fi[x], &b->fi[C3(b, i + 1, j, k)]},
                                              {&b->fj[x], &b->fj[C3(b, i, j + 1, k)]},
                                              {&b->fk[x], &b->fk[C3(b, i, j, k + 1)]}};
                    for (int d = 0; d < 3; ++d) {
                        v3_t S = vscale(vadd(fa[d][0]->S, fa[d][1]->S), F32(0.5));
                        double CS = csoun * sqrt(S.x * S.x + S.y * S.y + S.z * S.z);
                        b->lamc[d][x] = fabs(uvelr * S.x + vvelr * S.y + wvelr * S.z) + CS;
                    }
                }
    }
    /* source/ComputeAllBlockVars.cpp:56-63 */
    if (c->cfg.viscous)
        for (int n = 0; n < c->nb; ++n) {
            curl(&c->blk[n]);
            gradients(&c->blk[n]);
        }
    return 0;
}

/* ------------------------------------------------- viscous quantities */
/* CFluidProps::SutherLand (include/FluidProps.h:107-112) */
static double sutherland(const orc_nondim *nd, double temp)
{
    return F32(1.0) / nd->re_ref * pow(temp, F32(3.0) / F32(2.0)) * (double)(1.0f + 110.4f / 288.15f) /
           (temp + (double)(110.4f / 288.15f));
}

/* u v w T and nu~ of a cell (nu~ is zero without Spalart-Allmaras) */
#define NGV 5
static inline void uvwt(const blk_t *b, int i, int j, int k, double o[NGV])
{
    const long x = C3(b, i, j, k);
    const pv_t *w = &b->pv[x];
    o[0] = w->u; o[1] = w->v; o[2] = w->w; o[3] = w->t; o[4] = b->ev[x];
}

/* Difference of u v w T along tangential direction t at the face of
 * direction D whose higher-index cell is x (source/Spectral_curl_grad.cpp:
 * 85-493).  P averages the cells x, x+t, x+t-D, x-D and M the cells x, x-f,
 * x-f-s, x-s, where f and s are the lower and the higher of the directions D
 * and t, summed in that order.  A cell outside the block along both D and t
 * (a ghost-box corner, never filled) is left out and the remaining three are
 * averaged.  At the J face j == JM, k == 1 the reference scales M's
 * three-cell sum by 0.25 as well (line 331); that is kept. */
static void tangential_diff(const blk_t *b, int D, int t, const int x[3], double out[NGV])
{
    const int N[3] = {b->IM, b->JM, b->KM};
    const int f = D < t ? D : t, s = D < t ? t : D;
    int off[2][4][3];
    memset(off, 0, sizeof off);
    off[0][1][t] = 1;
    off[0][2][t] = 1; off[0][2][D] = -1;
    off[0][3][D] = -1;
    off[1][1][f] = -1;
    off[1][2][f] = -1; off[1][2][s] = -1;
    off[1][3][s] = -1;
    double avg[2][NGV];
    for (int h = 0; h < 2; ++h) {
        double sum[NGV] = {0, 0, 0, 0, 0};
        int n = 0;
        for (int m = 0; m < 4; ++m) {
            int y[3] = {x[0] + off[h][m][0], x[1] + off[h][m][1], x[2] + off[h][m][2]};
            int outD = y[D] < 1 || y[D] > N[D] - 1, outT = y[t] < 1 || y[t] > N[t] - 1;
            if (outD && outT) continue;
            double v[NGV];
            uvwt(b, y[0], y[1], y[2], v);
            for (int q = 0; q < NGV; ++q) sum[q] = n == 0 ? v[q] : sum[q] + v[q];
            ++n;
        }
        int quirk = h == 1 && D == 1 && t == 2 && x[1] == b->JM && x[2] == 1;
        for (int q = 0; q < NGV; ++q)
            avg[h][q] = n == 4 ? sum[q] / F32(4.0) : quirk ? F32(0.25) * sum[q] / F32(3.0) : sum[q] / F32(3.0);
    }
    for (int q = 0; q < NGV; ++q) out[q] = avg[0][q] - avg[1][q];
}

/* CFluidProps::Gradient (source/Spectral_curl_grad.cpp:85-493): gradients of
 * u v w T at the faces from index-space differences and the face coefficients */
static void gradients(blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    const int par = PAR(b);
    for (int D = 0; D < 3; ++D) {
        const face_t *F = D == 0 ? b->fi : D == 1 ? b->fj : b->fk;
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - (D != 0); ++i)
            for (int j = 1; j <= JM - (D != 1); ++j)
                for (int k = 1; k <= KM - (D != 2); ++k) {
                    int x[3] = {i, j, k};
                    double dv[3][NGV];
                    for (int t = 0; t < 3; ++t) {
                        if (t == D) {
                            double a[NGV], m[NGV];
                            uvwt(b, i, j, k, a);
                            uvwt(b, i - (D == 0), j - (D == 1), k - (D == 2), m);
                            for (int q = 0; q < NGV; ++q) dv[t][q] = a[q] - m[q];
                        } else {
                            tangential_diff(b, D, t, x, dv[t]);
                        }
                    }
                    long xx = C3(b, i, j, k);
                    const double *g = F[xx].g;
                    double *o = b->grad[D] + xx * 12;
                    for (int c = 0; c < 3; ++c)
                        for (int q = 0; q < 4; ++q)
                            o[c * 4 + q] = dv[0][q] * g[c] + dv[1][q] * g[3 + c] + dv[2][q] * g[6 + c];
                    /* nu~ (source/Spectral_curl_grad.cpp:221-223, 352-354, 483-485) */
                    for (int c = 0; c < 3; ++c)
                        b->gev[D][xx * 3 + c] = dv[0][4] * g[c] + dv[1][4] * g[3 + c] + dv[2][4] * g[6 + c];
                }
    }
}

/* CFluidProps::Curl (source/Spectral_curl_grad.cpp:44-83) */
static void curl(blk_t *b)
{
    const int par = PAR(b);
    #pragma omp parallel for schedule(static) if (par)
    for (int k = 1; k <= b->KM - 1; ++k)
        for (int j = 1; j <= b->JM - 1; ++j)
            for (int i = 1; i <= b->IM - 1; ++i) {
                double d[3][4], p[4], m[4];
                for (int t = 0; t < 3; ++t) {
                    uvwt(b, i + (t == 0), j + (t == 1), k + (t == 2), p);
                    uvwt(b, i - (t == 0), j - (t == 1), k - (t == 2), m);
                    for (int q = 0; q < 3; ++q) d[t][q] = F32(0.5) * (p[q] - m[q]);
                }
                long x = C3(b, i, j, k);
                const double *g = b->cell[x].g;
                double v_x = d[0][1] * g[0] + d[1][1] * g[3] + d[2][1] * g[6];
                double w_x = d[0][2] * g[0] + d[1][2] * g[3] + d[2][2] * g[6];
                double u_y = d[0][0] * g[1] + d[1][0] * g[4] + d[2][0] * g[7];
                double w_y = d[0][2] * g[1] + d[1][2] * g[4] + d[2][2] * g[7];
                double u_z = d[0][0] * g[2] + d[1][0] * g[5] + d[2][0] * g[8];
                double v_z = d[0][1] * g[2] + d[1][1] * g[5] + d[2][1] * g[8];
                b->omega[x] = sqrt(F32(0.5) * ((w_y - v_z) * (w_y - v_z) + (u_z - w_x) * (u_z - w_x) +
                                               (v_x - u_y) * (v_x - u_y)));
            }
}

/* CSpaceDiscr::ViscousFlux (source/ViscousFlux.cpp:9-283); without a
 * turbulence model mue_tur = cond_tur = 0 (lines 68-72) */
static void viscous_flux(orc_ctx *c, blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas, Pr_lam = c->nd.pr_lam, Pr_tur = c->nd.pr_tur;
    const double c43 = (double)(4.0f / 3.0f), c23 = (double)(2.0f / 3.0f);
    const int par = PAR(b);
    for (int D = 0; D < 3; ++D) {
        const face_t *F = D == 0 ? b->fi : D == 1 ? b->fj : b->fk;
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - (D != 0); ++i)
            for (int j = 1; j <= JM - (D != 1); ++j)
                for (int k = 1; k <= KM - (D != 2); ++k) {
                    long x = C3(b, i, j, k), xm = C3(b, i - (D == 0), j - (D == 1), k - (D == 2));
                    const pv_t *a = &b->pv[x], *m = &b->pv[xm];
                    double uf = F32(0.5) * (a->u + m->u), vf = F32(0.5) * (a->v + m->v);
                    double wf = F32(0.5) * (a->w + m->w), tf = F32(0.5) * (a->t + m->t);
                    double mue_lam = sutherland(&c->nd, tf);
                    double cond_lam = mue_lam * Gamma * Rgas / (Gamma - F32(1.0)) / Pr_lam;
                    double mue_tur = 0.0, cond_tur = 0.0;
                    if (c->cfg.turbulence) {
                        mue_tur = (b->mut[xm] + b->mut[x]) / F32(2.0);
                        cond_tur = mue_tur * Gamma * Rgas / (Gamma - F32(1.0)) / Pr_tur;
                    }
                    double mue = mue_lam + mue_tur, cond = cond_lam + cond_tur;
                    const double *gx = b->grad[D] + x * 12, *gy = gx + 4, *gz = gx + 8;
                    double txx = mue * (c43 * gx[0] - c23 * (gy[1] + gz[2]));
                    double txy = mue * (gy[0] + gx[1]);
                    double txz = mue * (gz[0] + gx[2]);
                    double tyy = mue * (c43 * gy[1] - c23 * (gx[0] + gz[2]));
                    double tyz = mue * (gz[1] + gy[2]);
                    double tzz = mue * (c43 * gz[2] - c23 * (gx[0] + gy[1]));
                    v3_t F1 = {0.0, 0.0, 0.0}, F2 = {txx, txy, txz}, F3 = {txy, tyy, tyz}, F4 = {txz, tyz, tzz};
                    v3_t F5 = {uf * txx + vf * txy + wf * txz + cond * gx[3], uf * txy + vf * tyy + wf * tyz + cond * gy[3],
                               uf * txz + vf * tyz + wf * tzz + cond * gz[3]};
                    v3_t S = F[x].S;
                    cv_t fv = {vdot(F1, S), vdot(F2, S), vdot(F3, S), vdot(F4, S), vdot(F5, S)};
                    b->face_fv[D][x] = fv;
                }
    }
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                long x = C3(b, i, j, k);
                cv_t s = csub(b->face_fv[0][C3(b, i + 1, j, k)], b->face_fv[0][x]);
                s = cadd(s, b->face_fv[1][C3(b, i, j + 1, k)]); s = csub(s, b->face_fv[1][x]);
                s = cadd(s, b->face_fv[2][C3(b, i, j, k + 1)]); s = csub(s, b->face_fv[2][x]);
                b->fv[x] = s;
            }
}

/* -------------------------------------------------------- Baldwin-Lomax */
/* CBlocks::BL_model_1d (source/BL_Model.cpp:283-329): arrays indexed 1..M-1
 * from the wall outwards */
static void bl_model_1d(int M, const double *dens, const double *Vvel, const double *dis, const double *Omega,
                        const double *mue_l, double *mue_t)
{
    const double A_plus = F32(26.0), Cwk = F32(1.0), k = F32(0.4), Ckleb = F32(0.3), alpha = F32(0.0168), Ccp = F32(1.6);
    double Tw = fabs(Omega[1] * mue_l[1]);
    for (int i = 1; i < M; ++i)
        if (fabs(Omega[i] * mue_l[i]) > Tw) Tw = fabs(Omega[i] * mue_l[i]);
    double Ret = sqrt(dens[1] * Tw) / mue_l[1];
    double Fmax = 0.0, y_max = 0.0, Udif = 0.0;
    for (int i = 1; i < M; ++i) {
        if (Vvel[i] > Udif) Udif = Vvel[i];
        double y_plus = dis[i] * Ret;
        double F = dis[i] * fabs(Omega[i]) * (F32(1.0) - exp(-y_plus / A_plus));
        if (F > Fmax) { Fmax = F; y_max = dis[i]; }
    }
    double wa = y_max * Fmax, wb = Cwk * y_max * Udif * Udif / Fmax;
    double Fwake = wa < wb ? wa : wb;
    for (int i = 1; i < M; ++i) {
        double y_plus = Ret * dis[i];
        double bL = k * dis[i] * (F32(1.0) - exp(-y_plus / A_plus));
        double mue_in = dens[i] * bL * bL * Omega[i];
        double Fkleb = F32(1.0) / (F32(1.0) + F32(5.5) * pow((Ckleb * dis[i] / y_max), 6.0));
        double mue_out = alpha * Ccp * dens[i] * Fwake * Fkleb;
        mue_t[i] = fabs(mue_in) > fabs(mue_out) ? mue_out : mue_in;
    }
}

/* CBlocks::BL_Model (source/BL_Model.cpp:8-279): for every wall patch of the
 * block, in table order, the 1-D model along each grid line normal to it; a
 * cell's first visit sets mue_t, later visits keep the minimum; the first cell
 * and the ghost cell of each line are then set to zero.  The reference passes
 * JM as the line length for K-face walls (line 225), which is only meaningful
 * when JM == KM; orc_bl rejects K-face walls otherwise. */
/* cell (i,j,k) at position n along a line normal to a wall of direction D
 * whose tangential indices are a (j for I walls, i otherwise) and cc (j for K
 * walls, k otherwise) */
static inline long bl_cell(const blk_t *b, int D, int n, int a, int cc)
{
    return D == 0 ? C3(b, n, a, cc) : D == 1 ? C3(b, a, n, cc) : C3(b, a, cc, n);
}

static int bl_model(orc_ctx *c, blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    int MM = IM > JM ? IM : JM;
    if (KM > MM) MM = KM;
    double *dens = calloc(MM, sizeof(double)), *Vvel = calloc(MM, sizeof(double)), *dis = calloc(MM, sizeof(double));
    double *Om = calloc(MM, sizeof(double)), *mue_l = calloc(MM, sizeof(double)), *mue_t = calloc(MM, sizeof(double));
    long box = (long)(IM + 1) * (JM + 1) * (KM + 1);
    unsigned char *flag = calloc(box, 1);
    int rc = 0;
    for (int q = 0; q < b->npatch && !rc; ++q) {
        const patch_t *P = &b->patch[q];
        if (P->kind != 2) continue;
        int D = (P->face - 1) / 2, hi = (P->face - 1) % 2;
        if (D == 2 && JM != KM) { rc = -1; break; }
        int N = D == 0 ? IM : D == 1 ? JM : KM;
        int M = D == 0 ? IM : JM;
        int a0 = D == 0 ? P->js : P->is, a1 = D == 0 ? P->je : P->ie;
        int c0 = D == 2 ? P->js : P->ks, c1 = D == 2 ? P->je : P->ke;
        for (int a = a0; a <= a1; ++a)
            for (int cc = c0; cc <= c1; ++cc) {
                const face_t *F = D == 0 ? b->fi : D == 1 ? b->fj : b->fk;
                const v3_t cw = F[bl_cell(b, D, hi ? N : 1, a, cc)].c;   /* wall face of the line */
                for (int n = 1; n <= N - 1; ++n) {
                    int n0 = hi ? N - n : n;
                    long x = bl_cell(b, D, n, a, cc);
                    v3_t cl = b->cell[x].c;
                    dis[n0] = sqrt(pow((cl.x - cw.x), 2.0) + pow((cl.y - cw.y), 2.0) + pow((cl.z - cw.z), 2.0));
                    const pv_t *w = &b->pv[x];
                    dens[n0] = w->d;
                    Vvel[n0] = sqrt(w->u * w->u + w->v * w->v + w->w * w->w);
                    Om[n0] = b->omega[x];
                    mue_l[n0] = sutherland(&c->nd, w->t);
                }
                bl_model_1d(M, dens, Vvel, dis, Om, mue_l, mue_t);
                for (int n = 1; n <= N - 1; ++n) {
                    int n0 = hi ? N - n : n;
                    long x = bl_cell(b, D, n, a, cc);
                    if (!flag[x]) { flag[x] = 1; b->mut[x] = mue_t[n0]; }
                    else b->mut[x] = b->mut[x] < mue_t[n0] ? b->mut[x] : mue_t[n0];
                }
            }
        /* first cell and ghost cell of every line of this wall */
        for (int a = a0; a <= a1; ++a)
            for (int cc = c0; cc <= c1; ++cc) {
                b->mut[bl_cell(b, D, hi ? N - 1 : 1, a, cc)] = 0.0;
                b->mut[bl_cell(b, D, hi ? N : 0, a, cc)] = 0.0;
            }
    }
    free(dens); free(Vvel); free(dis); free(Om); free(mue_l); free(mue_t); free(flag);
    return rc;
}

/* ------------------------------------------------------------ time step */
/* source/TimeStep.cpp:7-92 */
/* ------------------------------------------------------ Spalart-Allmaras */
/* fv1 of the model for nu~ and the laminar kinematic viscosity ev_lam */
static inline double sa_fv1(double X, double Cv1)
{
    return pow(X, F32(3.0)) / (pow(X, F32(3.0)) + Cv1 * Cv1 * Cv1);
}

/* CSpaceDiscr::SA_Model (source/SA_Model.cpp:9-258): mue_t = fv1 rho nu~ / Re
 * with the clipping of lines 48-75, the face fluxes of nu~ (first-order upwind
 * convection, diffusion (1 + c_b2)/sigma (nu + nu~) grad nu~, and the plain
 * gradient flux of the c_b2 term), the cell source and the viscous sum, and
 * mue_t of the ghost cells.  The reference reads its face gradients through
 * aliased pointers (lines 21-32): on I faces all three components are the x
 * derivative, on K faces the y and z derivatives are swapped; that is kept. */
static void sa_model(orc_ctx *c, blk_t *b)
{
    const double Re = c->nd.re_ref;
    const double Cb1 = F32(0.1355), Cb2 = F32(0.622), Cv1 = F32(7.1), sigma = (double)(2.0f / 3.0f), kavf = F32(0.41);
    const double Cw1 = Cb1 / kavf / kavf + (F32(1.0) + Cb2) / sigma;
    const double Cw2 = F32(0.3), Cw3 = F32(2.0), Ct3 = F32(1.2), Ct4 = F32(0.5);
    int IM = b->IM, JM = b->JM, KM = b->KM;
    const int par = PAR(b);
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                long x = C3(b, i, j, k);
                const pv_t *w = &b->pv[x];
                double ev_lam = Re * sutherland(&c->nd, w->t) / w->d;
                double X = b->ev[x] / ev_lam;
                b->mut[x] = sa_fv1(X, Cv1) * w->d * b->ev[x] / Re;
            }
    const double Max = 200.0;
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                long x = C3(b, i, j, k);
                const pv_t *w = &b->pv[x];
                if (b->mut[x] < 0.0) b->mut[x] = 0.0;
                if (b->ev[x] < 0.0) b->ev[x] = 0.0;
                double mue_lam = sutherland(&c->nd, w->t);
                double ev_lam = Re * mue_lam / w->d;
                if (mue_lam > 0 && b->mut[x] > (Max * mue_lam)) b->mut[x] = Max * mue_lam;
                if (ev_lam > 0 && b->ev[x] > (Max * ev_lam)) b->ev[x] = Max * ev_lam;
            }
    /* faces of the three directions (lines 82-181) */
    for (int D = 0; D < 3; ++D) {
        const face_t *Fa = D == 0 ? b->fi : D == 1 ? b->fj : b->fk;
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - (D != 0); ++i)
            for (int j = 1; j <= JM - (D != 1); ++j)
                for (int k = 1; k <= KM - (D != 2); ++k) {
                    long x = C3(b, i, j, k), xm = C3(b, i - (D == 0), j - (D == 1), k - (D == 2));
                    const pv_t *wp = &b->pv[x], *wm = &b->pv[xm];
                    const face_t *f = &Fa[x];
                    double densf = F32(0.5) * (wp->d + wm->d);
                    double tempf = F32(0.5) * (wp->t + wm->t);
                    double evf = F32(0.5) * (b->ev[x] + b->ev[xm]);
                    double ev_lam = Re * sutherland(&c->nd, tempf) / densf;
                    const double *g = b->gev[D] + x * 3;
                    double gx = g[0], gy = D == 0 ? g[0] : D == 1 ? g[1] : g[2], gz = D == 0 ? g[0] : D == 1 ? g[2] : g[1];
                    double tx = (F32(1.0) + Cb2) / sigma * (ev_lam + evf) * gx;
                    double ty = (F32(1.0) + Cb2) / sigma * (ev_lam + evf) * gy;
                    double tz = (F32(1.0) + Cb2) / sigma * (ev_lam + evf) * gz;
                    v3_t n = {f->S.x / f->s, f->S.y / f->s, f->S.z / f->s};
                    double Vn1 = wm->u * n.x + wm->v * n.y + wm->w * n.z;
                    double Vn2 = wp->u * n.x + wp->v * n.y + wp->w * n.z;
                    b->fc_tur[D][x] = (F32(0.5) * (Vn1 + fabs(Vn1)) * b->ev[xm] + F32(0.5) * (Vn2 - fabs(Vn2)) * b->ev[x]) * f->s;
                    b->fv_tur[D][x] = (tx * f->S.x + ty * f->S.y + tz * f->S.z) / Re;
                    b->fv1_tur[D][x] = gx * f->S.x + gy * f->S.y + gz * f->S.z;
                }
    }
    /* source and viscous sum (lines 184-213) */
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                long x = C3(b, i, j, k);
                long xi = C3(b, i + 1, j, k), xj = C3(b, i, j + 1, k), xk = C3(b, i, j, k + 1);
                const pv_t *w = &b->pv[x];
                double ev = b->ev[x], dis = b->dwall[x], Om = b->omega[x];
                double ev_lam = Re * sutherland(&c->nd, w->t) / w->d;
                double X = ev / ev_lam;
                double fv1 = sa_fv1(X, Cv1);
                double fv2 = F32(1.0) - X / (F32(1.0) + X * fv1);
                double ft2 = Ct3 * exp(F32(0.0) - Ct4 * X * X);
                double kd = kavf * dis;
                double sa = Om + ev * fv2 / ((kd * kd) * Re), sb = F32(0.3) * Om;
                double Sbar = sa > sb ? sa : sb;
                double ra = ev / (Sbar * kavf * kavf * dis * dis * Re), rb = F32(10.0);
                double r = ra < rb ? ra : rb;
                double gg = r + Cw2 * (pow(r, F32(6.0)) - r);
                double fw = gg * pow((1 + pow(Cw3, F32(6.0))) / (pow(gg, F32(6.0)) + pow(Cw3, F32(6.0))), (double)(1.0f / 6.0f));
                double ed = ev / dis;
                double Qc = (Cb1 * (F32(1.0) - ft2) * Sbar * ev -
                             (Cw1 * fw - Cb1 * ((1 - ft2) * fv2 + ft2) / (kavf * kavf)) * (ed * ed) / Re) * b->cell[x].vol;
                double sv = Qc + b->fv_tur[0][xi] - b->fv_tur[0][x] + b->fv_tur[1][xj] - b->fv_tur[1][x] + b->fv_tur[2][xk] -
                            b->fv_tur[2][x];
                double s1 = b->fv1_tur[0][xi] - b->fv1_tur[0][x] + b->fv1_tur[1][xj] - b->fv1_tur[1][x] + b->fv1_tur[2][xk] -
                            b->fv1_tur[2][x];
                b->cfv_tur[x] = sv - Cb2 * (ev_lam + ev) / sigma / Re * s1;
            }
    /* mue_t of the ghost cells (lines 215-257) */
    for (int side = 0; side < 6; ++side) {
        int d = side / 2, hi = side % 2;
        int N[3] = {IM, JM, KM};
        int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        for (int a = 1; a <= N[ta] - 1; ++a)
            for (int bb = 1; bb <= N[tb] - 1; ++bb) {
                int y[3];
                y[d] = hi ? N[d] : 0; y[ta] = a; y[tb] = bb;
                long x = C3(b, y[0], y[1], y[2]);
                const pv_t *w = &b->pv[x];
                double ev_lam = Re * sutherland(&c->nd, w->t) / w->d;
                double X = b->ev[x] / ev_lam;
                b->mut[x] = sa_fv1(X, Cv1) * w->d * b->ev[x] / Re;
            }
    }
}

/* CMesh::WallDistanceSA (source/WallDistanceSA.cpp:9-132): distance from each
 * cell centroid to the nearest wall-face centroid over all wall patches */
static void wall_distance(orc_ctx *c)
{
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    v3_t cc = b->cell[x].c;
                    double dmin = 10000.0;
                    for (int m = 0; m < c->nb; ++m) {
                        const blk_t *W = &c->blk[m];
                        for (int q = 0; q < W->npatch; ++q) {
                            const patch_t *P = &W->patch[q];
                            if (P->kind != 2) continue;
                            int dir = (P->face - 1) / 2;
                            const face_t *F = dir == 0 ? W->fi : dir == 1 ? W->fj : W->fk;
                            for (int ib = P->is; ib <= P->ie; ++ib)
                                for (int jb = P->js; jb <= P->je; ++jb)
                                    for (int kb = P->ks; kb <= P->ke; ++kb) {
                                        v3_t fcn = F[C3(W, ib, jb, kb)].c;
                                        double dx = cc.x - fcn.x, dy = cc.y - fcn.y, dz = cc.z - fcn.z;
                                        double dt = pow(dx * dx + dy * dy + dz * dz, 0.5);
                                        if (dmin >= dt) dmin = dt;
                                    }
                        }
                    }
                    b->dwall[x] = dmin;
                }
    }
}

int orc_timestep(orc_ctx *c)
{
    double Gamma = c->nd.gamma, Pr_lam = c->nd.pr_lam, Pr_tur = c->nd.pr_tur;
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        double dt_factor = b->flag_low_dt == 0 ? F32(1.0) : (double)(1.0f / (1.0f + b->flag_low_dt));
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    double lc = b->lamc[0][x] + b->lamc[1][x] + b->lamc[2][x];
                    double lv = 0.0;
                    if (c->cfg.viscous) {
                        const face_t *fa[3][2] = {{&b->fi[x], &b->fi[C3(b, i + 1, j, k)]},
                                                  {&b->fj[x], &b->fj[C3(b, i, j + 1, k)]},
                                                  {&b->fk[x], &b->fk[C3(b, i, j, k + 1)]}};
                        double mue_lam = sutherland(&c->nd, b->pv[x].t), mue_tur = b->mut[x], dens = b->pv[x].d;
                        double a = (double)(4.0f / 3.0f) / dens, g = Gamma / dens;
                        for (int d = 0; d < 3; ++d) {
                            v3_t S = vscale(vadd(fa[d][0]->S, fa[d][1]->S), 0.5);
                            double eq = S.x * S.x + S.y * S.y + S.z * S.z;
                            b->lamv[d][x] = (a > g ? a : g) * (mue_lam / Pr_lam + mue_tur / Pr_tur) * eq / b->cell[x].vol;
                        }
                        lv = b->lamv[0][x] + b->lamv[1][x] + b->lamv[2][x];
                    }
                    b->dt[x] = dt_factor * c->cfl * b->cell[x].vol / (lc + (double)(float)c->cfg.c_timestep * lv);
                }
        if (c->cfg.global_dt) {
            double m = 1E10;
            for (int i = 1; i <= b->IM - 1; ++i)
                for (int j = 1; j <= b->JM - 1; ++j)
                    for (int k = 1; k <= b->KM - 1; ++k) m = MINd(m, b->dt[C3(b, i, j, k)]);
            for (int i = 1; i <= b->IM - 1; ++i)
                for (int j = 1; j <= b->JM - 1; ++j)
                    for (int k = 1; k <= b->KM - 1; ++k) b->dt[C3(b, i, j, k)] = m;
        }
        if (b->flag_low_dt > 0) b->flag_low_dt = b->flag_low_dt - 1;
    }
    return 0;
}

/* Branch counters of the repair paths (reconstruction positivity fallback,
 * Limit_Vars, Limit_SA), read and cleared by orc_stats; tests use them to show
 * that a case really takes the branch it is meant to pin. */
static long g_stats[ORC_NSTATS];

static void stat_add(int which)
{
#pragma omp atomic
    g_stats[which]++;
}

int orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void orc_stats(long *out)
{
    for (int n = 0; n < ORC_NSTATS; ++n) { out[n] = g_stats[n]; g_stats[n] = 0; }
}

/* ------------------------------------------------------- reconstruction */
/* source/Reconstruction.cpp:70-265 */
static void recon_prim(int scheme, double Gamma, pv_t FF, pv_t F, pv_t B, pv_t BB, cv_t *cvL, cv_t *cvR)
{
    const double eps = F32(1E-6), eta = F32(0.2), k2c = 1.0, k2u = -1.0, lim_zero = F32(1E-20);
    pv_t L = {0}, R = {0};
    if (scheme == 100) {
        L = padd(F, kmulp(0.5, pminmod(psub(B, F), psub(F, FF))));
        R = psub(B, kmulp(0.5, pminmod(psub(B, F), psub(BB, B))));
    } else if (scheme == 200) {
        L = padd(F, kmulp(F32(0.25), padd(kmulp(F32(1.0) - k2u, pminmod(kmulp(F32(2.0), psub(B, F)), psub(F, FF))),
                                         kmulp(F32(1.0) + k2u, pminmod(kmulp(F32(2.0), psub(F, FF)), psub(B, F))))));
        R = psub(B, kmulp(F32(0.25), padd(kmulp(F32(1.0) - k2u, pminmod(kmulp(F32(2.0), psub(B, F)), psub(BB, B))),
                                         kmulp(F32(1.0) + k2u, pminmod(kmulp(F32(2.0), psub(BB, B)), psub(B, F))))));
    } else if (scheme == 300) {
        L = padd(F, kmulp(F32(0.25), padd(kmulp(F32(1.0) - k2c, pminmod(kmulp(F32(2.0), psub(B, F)), psub(F, FF))),
                                         kmulp(F32(1.0) + k2c, pminmod(psub(B, F), kmulp(2.0, psub(F, FF)))))));
        R = psub(B, kmulp(F32(0.25), padd(kmulp(F32(1.0) - k2c, pminmod(kmulp(F32(2.0), psub(B, F)), psub(BB, B))),
                                         kmulp(F32(1.0) + k2c, pminmod(psub(B, F), kmulp(2.0, psub(BB, B)))))));
    } else if (scheme == 400) {
        pv_t r1 = pdivv(paddk(psub(F, FF), eps), paddk(psub(B, F), eps));
        pv_t r2 = pdivv(paddk(psub(B, F), eps), paddk(psub(BB, B), eps));
        /* f1 = 1.0f-eta + (eta-0.55f)/2.0f/r2 + (eta+0.55f)/2.0f*r1 */
        double a0 = F32(1.0) - eta, a1 = (eta - F32(0.55)) / F32(2.0), a2 = (eta + F32(0.55)) / F32(2.0);
        pv_t f1 = padd(paddk(kmulp(1.0, pdivv((pv_t){a1, a1, a1, a1, a1, a1}, r2)), a0), kmulp(a2, r1));
        pv_t f;
        f.d = MAXd(F32(0.0), MIN3d(F32(2.0), f1.d, F32(2.0) * r1.d));
        f.u = MAXd(F32(0.0), MIN3d(F32(2.0), f1.u, F32(2.0) * r1.u));
        f.v = MAXd(F32(0.0), MIN3d(F32(2.0), f1.v, F32(2.0) * r1.v));
        f.w = MAXd(F32(0.0), MIN3d(F32(2.0), f1.w, F32(2.0) * r1.w));
        f.p = MAXd(F32(0.0), MIN3d(F32(2.0), f1.p, F32(2.0) * r1.p));
        f.t = 0.0;
        L = padd(F, pmul(kmulp(F32(0.5), f), psub(B, F)));
        r1 = pdivv(paddk(psub(BB, B), eps), paddk(psub(B, F), eps));
        r2 = pdivv(paddk(psub(B, F), eps), paddk(psub(F, FF), eps));
        f1 = padd(paddk(kmulp(1.0, pdivv((pv_t){a1, a1, a1, a1, a1, a1}, r2)), a0), kmulp(a2, r1));
        f.d = MAXd(F32(0.0), MIN3d(F32(2.0), f1.d, F32(2.0) * r1.d));
        f.u = MAXd(F32(0.0), MIN3d(F32(2.0), f1.u, F32(2.0) * r1.u));
        f.v = MAXd(F32(0.0), MIN3d(F32(2.0), f1.v, F32(2.0) * r1.v));
        f.w = MAXd(F32(0.0), MIN3d(F32(2.0), f1.w, F32(2.0) * r1.w));
        f.p = MAXd(F32(0.0), MIN3d(F32(2.0), f1.p, F32(2.0) * r1.p));
        R = psub(B, pmul(kmulp(F32(0.5), f), psub(B, F)));
    } else if (scheme == 500) {
        pv_t dF = psub(F, FF), dB = psub(B, F), dR = psub(BB, B);
        pv_t SL = pdivv(paddk(pmul(kmulp(2.0, dF), dB), eps), paddk(padd(pvsq(dF), pvsq(dB)), eps));
        pv_t SR = pdivv(paddk(pmul(kmulp(2.0, dR), dB), eps), paddk(padd(pvsq(dR), pvsq(dB)), eps));
        L = padd(F, pmul(pdiv(SL, 4.0), padd(pmul(ksubp(1.0, pdiv(SL, 3.0)), dF), pmul(paddk(pdiv(SL, 3.0), 1.0), dB))));
        R = psub(B, pmul(pdiv(SR, 4.0), padd(pmul(ksubp(1.0, pdiv(SR, 3.0)), dR), pmul(paddk(pdiv(SR, 3.0), 1.0), dB))));
    }
    if (L.d <= lim_zero || L.p <= lim_zero) { L = F; stat_add(ORC_STAT_RECON_FALLBACK); }
    if (R.d <= lim_zero || R.p <= lim_zero) { R = B; stat_add(ORC_STAT_RECON_FALLBACK); }
    cvL->r = L.d; cvL->ru = L.d * L.u; cvL->rv = L.d * L.v; cvL->rw = L.d * L.w;
    cvL->re = L.p / (Gamma - F32(1.0)) + L.d * (L.u * L.u + L.v * L.v + L.w * L.w) / F32(2.0);
    cvR->r = R.d; cvR->ru = R.d * R.u; cvR->rv = R.d * R.v; cvR->rw = R.d * R.w;
    cvR->re = R.p / (Gamma - F32(1.0)) + R.d * (R.u * R.u + R.v * R.v + R.w * R.w) / F32(2.0);
}

/* The limited left/right states of one five-component variable set, before
 * the positivity test: the scheme part of ConservationReconstruct
 * (source/Reconstruction.cpp:15-59), which CharacterReconstruct repeats on
 * characteristic variables (source/Reconstruction.cpp:407-459). */
static void recon_cons_core(int scheme, cv_t FF, cv_t F, cv_t B, cv_t BB, cv_t *cvL, cv_t *cvR)
{
    const double eps = F32(1E-6), eta = F32(0.05), k2c = F32(1.0), k2u = F32(-1.0);
    cv_t L = {0}, R = {0};
    if (scheme == 100) {
        L = cadd(F, cscale(cminmod(csub(B, F), csub(F, FF)), F32(0.5)));
        R = csub(B, cscale(cminmod(csub(B, F), csub(BB, B)), F32(0.5)));
    } else if (scheme == 200) {
        L = cadd(F, cscale(cadd(cscale(cminmod(cscale(csub(B, F), F32(2.0)), csub(F, FF)), F32(1.0) - k2u),
                                cscale(cminmod(cscale(csub(F, FF), F32(2.0)), csub(B, F)), F32(1.0) + k2u)), 0.25));
        R = csub(B, cscale(cadd(cscale(cminmod(cscale(csub(B, F), F32(2.0)), csub(BB, B)), F32(1.0) - k2u),
                                cscale(cminmod(cscale(csub(BB, B), F32(2.0)), csub(B, F)), F32(1.0) + k2u)), 0.25));
    } else if (scheme == 300) {
        L = cadd(F, cscale(cadd(cscale(cminmod(cscale(csub(B, F), F32(2.0)), csub(F, FF)), F32(1.0) - k2c),
                                cscale(cminmod(csub(B, F), cscale(csub(F, FF), F32(2.0))), F32(1.0) + k2c)), 0.25));
        R = csub(B, cscale(cadd(cscale(cminmod(cscale(csub(B, F), F32(2.0)), csub(BB, B)), F32(1.0) - k2c),
                                cscale(cminmod(csub(B, F), cscale(csub(BB, B), F32(2.0))), F32(1.0) + k2c)), 0.25));
    } else if (scheme == 400) {
        double a0 = F32(1.0) - eta, a1 = (eta - F32(0.55)) / F32(2.0), a2 = (eta + F32(0.55)) / F32(2.0);
        cv_t r1 = cdivv(caddk(csub(F, FF), eps), caddk(csub(B, F), eps));
        cv_t r2 = cdivv(caddk(csub(B, F), eps), caddk(csub(BB, B), eps));
        cv_t f1 = cadd(caddk(cdivv((cv_t){a1, a1, a1, a1, a1}, r2), a0), cscale(r1, a2));
        cv_t f;
        f.r = MAXd(F32(0.0), MIN3d(F32(2.0), f1.r, F32(2.0) * r1.r));
        f.ru = MAXd(F32(0.0), MIN3d(F32(2.0), f1.ru, F32(2.0) * r1.ru));
        f.rv = MAXd(F32(0.0), MIN3d(F32(2.0), f1.rv, F32(2.0) * r1.rv));
        f.rw = MAXd(F32(0.0), MIN3d(F32(2.0), f1.rw, F32(2.0) * r1.rw));
        f.re = MAXd(F32(0.0), MIN3d(F32(2.0), f1.re, F32(2.0) * r1.re));
        L = cadd(F, cmul(cscale(f, F32(0.5)), csub(B, F)));
        r1 = cdivv(caddk(csub(BB, B), eps), caddk(csub(B, F), eps));
        r2 = cdivv(caddk(csub(B, F), eps), caddk(csub(F, FF), eps));
        f1 = cadd(caddk(cdivv((cv_t){a1, a1, a1, a1, a1}, r2), a0), cscale(r1, a2));
        f.r = MAXd(F32(0.0), MIN3d(F32(2.0), f1.r, F32(2.0) * r1.r));
        f.ru = MAXd(F32(0.0), MIN3d(F32(2.0), f1.ru, F32(2.0) * r1.ru));
        f.rv = MAXd(F32(0.0), MIN3d(F32(2.0), f1.rv, F32(2.0) * r1.rv));
        f.rw = MAXd(F32(0.0), MIN3d(F32(2.0), f1.rw, F32(2.0) * r1.rw));
        f.re = MAXd(F32(0.0), MIN3d(F32(2.0), f1.re, F32(2.0) * r1.re));
        R = csub(B, cmul(cscale(f, F32(0.5)), csub(B, F)));
    } else if (scheme == 500) {
        cv_t dF = csub(F, FF), dB = csub(B, F), dR = csub(BB, B);
        cv_t SL = cdivv(caddk(cmul(cscale(dF, F32(2.0)), dB), eps), caddk(cadd(cvsq(dF), cvsq(dB)), eps));
        cv_t SR = cdivv(caddk(cmul(cscale(dR, F32(2.0)), dB), eps), caddk(cadd(cvsq(dR), cvsq(dB)), eps));
        L = cadd(F, cmul(cdiv(SL, F32(4.0)), cadd(cmul(ksubc(F32(1.0), cdiv(SL, F32(3.0))), dF), cmul(caddk(cdiv(SL, F32(3.0)), F32(1.0)), dB))));
        R = csub(B, cmul(cdiv(SR, F32(4.0)), cadd(cmul(ksubc(F32(1.0), cdiv(SR, F32(3.0))), dR), cmul(caddk(cdiv(SR, F32(3.0)), F32(1.0)), dB))));
    }
    *cvL = L; *cvR = R;
}

/* First-order fallback of a conservative face state whose density or
 * pressure is not positive (source/Reconstruction.cpp:61-67, 461-467). */
static void cons_positivity(double Gamma, cv_t F, cv_t B, cv_t *L, cv_t *R)
{
    const double lim_zero = F32(1E-20);
    double pL = (Gamma - F32(1.0)) * (L->re - F32(0.5) * (L->ru * L->ru + L->rv * L->rv + L->rw * L->rw) / L->r);
    double pR = (Gamma - F32(1.0)) * (R->re - F32(0.5) * (R->ru * R->ru + R->rv * R->rv + R->rw * R->rw) / R->r);
    if (L->r <= lim_zero || pL <= lim_zero) { *L = F; stat_add(ORC_STAT_RECON_FALLBACK); }
    if (R->r <= lim_zero || pR <= lim_zero) { *R = B; stat_add(ORC_STAT_RECON_FALLBACK); }
}

/* source/Reconstruction.cpp:7-68 */
static void recon_cons(int scheme, double Gamma, cv_t FF, cv_t F, cv_t B, cv_t BB, cv_t *cvL, cv_t *cvR)
{
    recon_cons_core(scheme, FF, F, B, BB, cvL, cvR);
    cons_positivity(Gamma, F, B, cvL, cvR);
}

/* REAL[6][6] * TConsVars of source/defs.cpp:359-370, rows 1..5 summed left
 * to right */
static cv_t mat5(const double A[6][6], cv_t b)
{
    cv_t c;
    c.r = A[1][1] * b.r + A[1][2] * b.ru + A[1][3] * b.rv + A[1][4] * b.rw + A[1][5] * b.re;
    c.ru = A[2][1] * b.r + A[2][2] * b.ru + A[2][3] * b.rv + A[2][4] * b.rw + A[2][5] * b.re;
    c.rv = A[3][1] * b.r + A[3][2] * b.ru + A[3][3] * b.rv + A[3][4] * b.rw + A[3][5] * b.re;
    c.rw = A[4][1] * b.r + A[4][2] * b.ru + A[4][3] * b.rv + A[4][4] * b.rw + A[4][5] * b.re;
    c.re = A[5][1] * b.r + A[5][2] * b.ru + A[5][3] * b.rv + A[5][4] * b.rw + A[5][5] * b.re;
    return c;
}

/* CSpaceDiscr::CharacterReconstruct (source/Reconstruction.cpp:363-468,
 * reconstruction variable code 300, and any code other than 100 and 200 per
 * source/UserInput.cpp:117-130): the left and right eigenvector matrices S, S1
 * of the face state averaged from the primitive variables of F and B, the
 * four conservative states taken to characteristic variables by S, limited
 * as ConservationReconstruct limits conservative ones, and back through S1.
 * The -1/3, 2/3 and 1/6 entries are float quotients in the reference, and
 * its SQRT(3.0f) is the float root. */
static void recon_char(int scheme, double Gamma, double Rgas, pv_t wF, pv_t wB, cv_t FF, cv_t F, cv_t B, cv_t BB,
                       cv_t *cvL, cv_t *cvR)
{
    double S[6][6] = {{0}}, S1[6][6] = {{0}};
    double fd = F32(0.5) * (wF.d + wB.d), uvel = F32(0.5) * (wF.u + wB.u), vvel = F32(0.5) * (wF.v + wB.v);
    double wvel = F32(0.5) * (wF.w + wB.w), press = F32(0.5) * (wF.p + wB.p);
    double V2 = F32(0.5) * (uvel * uvel + vvel * vvel + wvel * wvel);
    double temp = press / fd / Rgas;
    double C = sqrt(Gamma * Rgas * temp);
    double C2 = C * C;
    double Med0 = C2 / (Gamma - F32(1.0));
    /* SQRT(3.0f) is C++'s float overload of sqrt: the float root of 3 */
    const double sqrt3f = (double)sqrtf(3.0f);
    double Med1 = (Gamma - F32(1.0)) / C * sqrt3f;
    double Med3 = F32(1.0) / Med0;
    double Med4 = F32(0.5) / C / sqrt3f;
    double V_bar = (uvel + vvel + wvel);
    const double m_third = (double)(-1.0f / 3.0f), two_thirds = (double)(2.0f / 3.0f);
    const double sixth = (double)(1.0f / 6.0f);
    S[1][1] = V2 - Med0; S[1][2] = -uvel; S[1][3] = -vvel; S[1][4] = -wvel; S[1][5] = F32(1.0);
    S[2][1] = uvel - vvel; S[2][2] = F32(-1.0); S[2][3] = F32(1.0); S[2][4] = 0.0; S[2][5] = 0.0;
    S[3][1] = uvel - wvel; S[3][2] = F32(-1.0); S[3][3] = 0.0; S[3][4] = F32(1.0); S[3][5] = 0.0;
    S[4][1] = -V_bar - V2 * Med1; S[4][2] = F32(1.0) + Med1 * uvel; S[4][3] = F32(1.0) + Med1 * vvel;
    S[4][4] = F32(1.0) + Med1 * wvel; S[4][5] = -Med1;
    S[5][1] = -V_bar + V2 * Med1; S[5][2] = F32(1.0) - Med1 * uvel; S[5][3] = F32(1.0) - Med1 * vvel;
    S[5][4] = F32(1.0) - Med1 * wvel; S[5][5] = Med1;
    S1[1][1] = -Med3; S1[1][2] = 0.0; S1[1][3] = 0.0; S1[1][4] = -Med4; S1[1][5] = Med4;
    S1[2][1] = -Med3 * uvel; S1[2][2] = m_third; S1[2][3] = m_third;
    S1[2][4] = sixth - Med4 * uvel; S1[2][5] = sixth + Med4 * uvel;
    S1[3][1] = -Med3 * vvel; S1[3][2] = two_thirds; S1[3][3] = m_third;
    S1[3][4] = sixth - Med4 * vvel; S1[3][5] = sixth + Med4 * vvel;
    S1[4][1] = -Med3 * wvel; S1[4][2] = m_third; S1[4][3] = two_thirds;
    S1[4][4] = sixth - Med4 * wvel; S1[4][5] = sixth + Med4 * wvel;
    S1[5][1] = -Med3 * V2; S1[5][2] = vvel - V_bar / F32(3.0); S1[5][3] = wvel - V_bar / F32(3.0);
    S1[5][4] = V_bar / F32(6.0) - Med4 * (V2 + Med0); S1[5][5] = V_bar / F32(6.0) + Med4 * (V2 + Med0);
    cv_t chL, chR;
    recon_cons_core(scheme, mat5(S, FF), mat5(S, F), mat5(S, B), mat5(S, BB), &chL, &chR);
    *cvL = mat5(S1, chL);
    *cvR = mat5(S1, chR);
    cons_positivity(Gamma, F, B, cvL, cvR);
}

/* --------------------------------------------------------- face fluxes */
typedef struct { v3_t n, t1, t2; double que; } frame_t;

/* source/Flux_Split_Scheme.cpp:728-776 */
static frame_t coord_rotate(const face_t *f, v3_t diag, double w_rot)
{
    frame_t F;
    F.n = vdiv(f->S, f->s);
    double fue = F32(-1.0) * f->c.y * w_rot, fve = f->c.x * w_rot;
    F.t1 = vdiv(diag, sqrt(diag.x * diag.x + diag.y * diag.y + diag.z * diag.z));
    v3_t r = {F.n.y * F.t1.z - F.n.z * F.t1.y, F.n.z * F.t1.x - F.n.x * F.t1.z, F.n.x * F.t1.y - F.n.y * F.t1.x};
    F.t2 = vdiv(r, sqrt(r.x * r.x + r.y * r.y + r.z * r.z));
    F.que = fue * F.n.x + fve * F.n.y + F32(0.0) * F.n.z;
    return F;
}

static cv_t rotate_in(cv_t q, const frame_t *F)
{
    cv_t Q;
    Q.r = q.r;
    Q.ru = q.ru * F->n.x + q.rv * F->n.y + q.rw * F->n.z;
    Q.rv = q.ru * F->t1.x + q.rv * F->t1.y + q.rw * F->t1.z;
    Q.rw = q.ru * F->t2.x + q.rv * F->t2.y + q.rw * F->t2.z;
    Q.re = q.re;
    return Q;
}

static inline double entropy_corr(double z, double d) { return z > d ? z : F32(0.5) * (z * z + d * d) / d; }

/* source/Flux_Split_Scheme.cpp:568-635 */
static cv_t roe(double Gamma, cv_t QL, cv_t QR, double que)
{
    double densL = QL.r, uvelL = QL.ru / QL.r, vvelL = QL.rv / QL.r, wvelL = QL.rw / QL.r, EL = QL.re / QL.r;
    double pL = (Gamma - F32(1.0)) * densL * (EL - F32(0.5) * (uvelL * uvelL + vvelL * vvelL + wvelL * wvelL));
    double HL = EL + pL / densL;
    double densR = QR.r, uvelR = QR.ru / QR.r, vvelR = QR.rv / QR.r, wvelR = QR.rw / QR.r, ER = QR.re / QR.r;
    double pR = (Gamma - F32(1.0)) * densR * (ER - F32(0.5) * (uvelR * uvelR + vvelR * vvelR + wvelR * wvelR));
    double HR = ER + pR / densR;
    cv_t FL, FR;
    FL.r = QL.r * (uvelL - que); FL.ru = QL.ru * (uvelL - que) + pL; FL.rv = QL.rv * (uvelL - que);
    FL.rw = QL.rw * (uvelL - que); FL.re = (QL.re + pL) * (uvelL - que) + que * pL;
    FR.r = QR.r * (uvelR - que); FR.ru = QR.ru * (uvelR - que) + pR; FR.rv = QR.rv * (uvelR - que);
    FR.rw = QR.rw * (uvelR - que); FR.re = (QR.re + pR) * (uvelR - que) + que * pR;
    double tmpp = sqrt(densR / densL);
    double uav = (uvelL + tmpp * uvelR) / (F32(1.0) + tmpp);
    double vav = (vvelL + tmpp * vvelR) / (F32(1.0) + tmpp);
    double wav = (wvelL + tmpp * wvelR) / (F32(1.0) + tmpp);
    double Hav = (HL + tmpp * HR) / (F32(1.0) + tmpp);
    double cav = sqrt((Gamma - F32(1.0)) * (Hav - F32(0.5) * (uav * uav + vav * vav + wav * wav)));
    double l1 = fabs((uav - que) - cav), l2 = fabs((uav - que)), l5 = fabs((uav - que) + cav);
    double delta = F32(0.5);
    l1 = entropy_corr(l1, delta); l2 = entropy_corr(l2, delta); l5 = entropy_corr(l5, delta);
    double a2 = (QR.rv - QL.rv) / (vav + F32(1E-8)) - (densR - densL);
    double a3 = (QR.rw - QL.rw) / (wav + F32(1E-8)) - (densR - densL);
    double a4 = (Gamma - F32(1.0)) * ((densR - densL) * (Hav - uav * uav - vav * vav - wav * wav) + uav * (QR.ru - QL.ru) +
                                      vav * (QR.rv - QL.rv) + wav * (QR.rw - QL.rw) - (QR.re - QL.re)) / (cav * cav);
    double a1 = ((uav + cav) * (densR - densL) - (QR.ru - QL.ru) - cav * a4) / (F32(2.0) * cav);
    double a5 = densR - densL - (a1 + a4);
    cv_t Fc;
    Fc.r = F32(0.5) * (FL.r + FR.r) - F32(0.5) * (l1 * a1 + l2 * a4 + l5 * a5);
    Fc.ru = F32(0.5) * (FL.ru + FR.ru) - F32(0.5) * (l1 * a1 * (uav - cav) + l2 * a4 * uav + l5 * a5 * (uav + cav));
    Fc.rv = F32(0.5) * (FL.rv + FR.rv) - F32(0.5) * (l1 * a1 * vav + l2 * a2 * vav + l2 * a4 * vav + l5 * a5 * vav);
    Fc.rw = F32(0.5) * (FL.rw + FR.rw) - F32(0.5) * (l1 * a1 * wav + l2 * a3 * wav + l2 * a4 * wav + l5 * a5 * wav);
    Fc.re = F32(0.5) * (FL.re + FR.re) - F32(0.5) * (l1 * a1 * (Hav - uav * cav) + l2 * a2 * vav * vav + l2 * a3 * wav * wav +
                                                     l2 * a4 * F32(0.5) * (uav * uav + vav * vav + wav * wav) + l5 * a5 * (Hav + uav * cav));
    return Fc;
}

/* source/Flux_Split_Scheme.cpp:82-149 */
static cv_t steger_warming(double Gamma, cv_t QL, cv_t QR, double que)
{
    const double eps = F32(1E-6);
    double lp[6], lm[6], fp[6], fm[6];
    double densL = QL.r, uL = QL.ru / QL.r, vL = QL.rv / QL.r, wL = QL.rw / QL.r, eL = QL.re / QL.r;
    double t = Gamma * (Gamma - F32(1.0)) * (eL - F32(0.5) * (uL * uL + vL * vL + wL * wL));
    t = sqrt(t * t);
    double cL = sqrt(t);
    lp[1] = F32(0.5) * ((uL - que) + sqrt((uL - que) * (uL - que) + eps * eps));
    lp[2] = lp[1]; lp[3] = lp[1];
    lp[4] = F32(0.5) * (((uL - que) - cL) + sqrt(((uL - que) - cL) * ((uL - que) - cL) + eps * eps));
    lp[5] = F32(0.5) * (((uL - que) + cL) + sqrt(((uL - que) + cL) * ((uL - que) + cL) + eps * eps));
    double W2 = (F32(3.0) - Gamma) * (lp[4] + lp[5]) * cL * cL / F32(2.0) / (Gamma - F32(1.0));
    double A = (uL - cL) * (uL - cL) + vL * vL + wL * wL;
    double B = (uL + cL) * (uL + cL) + vL * vL + wL * wL;
    fp[1] = densL / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lp[1] + lp[4] + lp[5]);
    fp[2] = densL / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lp[1] * uL + lp[4] * (uL - cL) + lp[5] * (uL + cL));
    fp[3] = densL / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lp[1] * vL + lp[4] * vL + lp[5] * vL);
    fp[4] = densL / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lp[1] * wL + lp[4] * wL + lp[5] * wL);
    fp[5] = densL / F32(2.0) / Gamma * ((Gamma - F32(1.0)) * lp[1] * (uL * uL + vL * vL + wL * wL) + F32(0.5) * lp[4] * A +
                                        F32(0.5) * lp[5] * B + W2);
    double densR = QR.r, uR = QR.ru / QR.r, vR = QR.rv / QR.r, wR = QR.rw / QR.r, eR = QR.re / QR.r;
    t = Gamma * (Gamma - F32(1.0)) * (eR - F32(0.5) * (uR * uR + vR * vR + wR * wR));
    t = sqrt(t * t);
    double cR = sqrt(t);
    lm[1] = F32(0.5) * ((uR - que) - sqrt((uR - que) * (uR - que) + eps * eps));
    lm[2] = lm[1]; lm[3] = lm[1];
    lm[4] = F32(0.5) * (((uR - que) - cR) - sqrt(((uR - que) - cR) * ((uR - que) - cR) + eps * eps));
    lm[5] = F32(0.5) * (((uR - que) + cR) - sqrt(((uR - que) + cR) * ((uR - que) + cR) + eps * eps));
    W2 = (F32(3.0) - Gamma) * (lm[4] + lm[5]) * cR * cR / F32(2.0) / (Gamma - F32(1.0));
    A = (uR - cR) * (uR - cR) + vR * vR + wR * wR;
    B = (uR + cR) * (uR + cR) + vR * vR + wR * wR;
    fm[1] = densR / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lm[1] + lm[4] + lm[5]);
    fm[2] = densR / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lm[1] * uR + lm[4] * (uR - cR) + lm[5] * (uR + cR));
    fm[3] = densR / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lm[1] * vR + lm[4] * vR + lm[5] * vR);
    fm[4] = densR / F32(2.0) / Gamma * (F32(2.0) * (Gamma - F32(1.0)) * lm[1] * wR + lm[4] * wR + lm[5] * wR);
    fm[5] = densR / F32(2.0) / Gamma * ((Gamma - F32(1.0)) * lm[1] * (uR * uR + vR * vR + wR * wR) + F32(0.5) * lm[4] * A +
                                        F32(0.5) * lm[5] * B + W2);
    cv_t Fc = {fp[1] + fm[1], fp[2] + fm[2], fp[3] + fm[3], fp[4] + fm[4], fp[5] + fm[5]};
    return Fc;
}

/* stencil element: cell array inside [1, n-1], Dummy layers outside */
static void stencil(const blk_t *b, int dir, int i, int j, int k, int off, cv_t *q, pv_t *w)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    int ii = i, jj = j, kk = k, *m = dir == 0 ? &ii : dir == 1 ? &jj : &kk;
    int N = dir == 0 ? IM : dir == 1 ? JM : KM;
    *m += off;
    int p = *m;
    if (p >= 1 && p <= N - 1) { long x = C3(b, ii, jj, kk); *q = b->cv[x]; *w = b->pv[x]; return; }
    int d, layer;
    if (p < 1) { d = 2 * dir; layer = 1 - p; }
    else { d = 2 * dir + 1; layer = p - (N - 1); }
    long x;
    if (dir == 0) x = DI(b, d, layer, jj, kk);
    else if (dir == 1) x = DI(b, d, ii, layer, kk);
    else x = DI(b, d, ii, jj, layer);
    *q = b->dcv[d][x];
    *w = b->dpv[d][x];
}

/* source/Upwind_Convective.cpp:5-360 */
static void face_flux(orc_ctx *c, blk_t *b, int dir, int i, int j, int k)
{
    double Gamma = c->nd.gamma;
    cv_t qFF, qF, qB, qBB, cvL, cvR;
    pv_t wFF, wF, wB, wBB;
    stencil(b, dir, i, j, k, -2, &qFF, &wFF);
    stencil(b, dir, i, j, k, -1, &qF, &wF);
    stencil(b, dir, i, j, k, 0, &qB, &wB);
    stencil(b, dir, i, j, k, 1, &qBB, &wBB);
    if (c->cfg.recon_var == 100) recon_cons(c->cfg.recon_scheme, Gamma, qFF, qF, qB, qBB, &cvL, &cvR);
    else if (c->cfg.recon_var == 200) recon_prim(c->cfg.recon_scheme, Gamma, wFF, wF, wB, wBB, &cvL, &cvR);
    else recon_char(c->cfg.recon_scheme, Gamma, c->nd.rgas, wF, wB, qFF, qF, qB, qBB, &cvL, &cvR);
    const face_t *f;
    v3_t diag;
    v3_t *N = b->node;
    if (dir == 0) { f = &b->fi[C3(b, i, j, k)]; diag = vsub(N[C3(b, i, j + 1, k + 1)], N[C3(b, i, j, k)]); }
    else if (dir == 1) { f = &b->fj[C3(b, i, j, k)]; diag = vsub(N[C3(b, i + 1, j, k + 1)], N[C3(b, i, j, k)]); }
    else { f = &b->fk[C3(b, i, j, k)]; diag = vsub(N[C3(b, i + 1, j + 1, k)], N[C3(b, i, j, k)]); }
    frame_t F = coord_rotate(f, diag, b->w_rot);
    cv_t QL = rotate_in(cvL, &F), QR = rotate_in(cvR, &F);
    cv_t Fc = c->cfg.flux_scheme == 200 ? steger_warming(Gamma, QL, QR, F.que) : roe(Gamma, QL, QR, F.que);
    cv_t *o = &b->face_fc[dir][C3(b, i, j, k)];
    o->r = Fc.r * f->s;
    o->ru = (Fc.ru * F.n.x + Fc.rv * F.t1.x + Fc.rw * F.t2.x) * f->s;
    o->rv = (Fc.ru * F.n.y + Fc.rv * F.t1.y + Fc.rw * F.t2.y) * f->s;
    o->rw = (Fc.ru * F.n.z + Fc.rv * F.t1.z + Fc.rw * F.t2.z) * f->s;
    o->re = Fc.re * f->s;
}

/* source/WallFlux.cpp:70-175 */
static void wall_flux(orc_ctx *c, blk_t *b, const patch_t *P)
{
    double Gamma = c->nd.gamma;
    for (int i = P->is; i <= P->ie; ++i)
        for (int j = P->js; j <= P->je; ++j)
            for (int k = P->ks; k <= P->ke; ++k) {
                int dir = (P->face - 1) / 2;
                const face_t *f;
                long xa = C3(b, i, j, k), xb;
                if (dir == 0) { f = &b->fi[xa]; xb = C3(b, i - 1, j, k); }
                else if (dir == 1) { f = &b->fj[xa]; xb = C3(b, i, j - 1, k); }
                else { f = &b->fk[xa]; xb = C3(b, i, j, k - 1); }
                v3_t S = f->S;
                pv_t fp = kmulp(F32(0.5), padd(b->pv[xa], b->pv[xb]));
                double fue = F32(-1.0) * f->c.y * P->rot, fve = f->c.x * P->rot;
                cv_t fc;
                fc.r = fp.d; fc.ru = fp.d * fp.u; fc.rv = fp.d * fp.v; fc.rw = fp.d * fp.w;
                double Vn_a = fp.u * S.x + fp.v * S.y + fp.w * S.z;
                double Vn_e = fue * S.x + fve * S.y + F32(0.0) * S.z;
                fc.re = fp.p / (Gamma - F32(1.0)) + fp.d * (fp.u * fp.u + fp.v * fp.v + fp.w * fp.w) / F32(2.0);
                cv_t *o = &b->face_fc[dir][xa];
                o->r = fc.r * (Vn_a - Vn_e);
                o->ru = fc.ru * (Vn_a - Vn_e) + fp.p * S.x;
                o->rv = fc.rv * (Vn_a - Vn_e) + fp.p * S.y;
                o->rw = fc.rw * (Vn_a - Vn_e) + fp.p * S.z;
                o->re = (fc.re + fp.p) * (Vn_a - Vn_e) + fp.p * Vn_e;
                if (c->cfg.turbulence == 200) b->fc_tur[dir][xa] = 0;   /* no nu~ convection through a wall */
                if (P->face % 2 == 1) b->mut[xb] = -b->mut[xa];
                else b->mut[xa] = -b->mut[xb];
            }
}

/* source/ComputeAllBlockVars.cpp:73-119 */
int orc_convective(orc_ctx *c)
{
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        int IM = b->IM, JM = b->JM, KM = b->KM;
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM; ++i)
            for (int j = 1; j <= JM - 1; ++j)
                for (int k = 1; k <= KM - 1; ++k) face_flux(c, b, 0, i, j, k);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - 1; ++i)
            for (int j = 1; j <= JM; ++j)
                for (int k = 1; k <= KM - 1; ++k) face_flux(c, b, 1, i, j, k);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - 1; ++i)
            for (int j = 1; j <= JM - 1; ++j)
                for (int k = 1; k <= KM; ++k) face_flux(c, b, 2, i, j, k);
    }
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        for (int q = 0; q < b->npatch; ++q)
            if (b->patch[q].kind == 2) wall_flux(c, b, &b->patch[q]);
    }
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        double w_rot = b->w_rot;
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    const pv_t *w = &b->pv[x];
                    double vol = b->cell[x].vol;
                    cv_t Qc = {0.0, w->d * (w_rot * w->v) * vol, F32(-1.0) * w->d * (w_rot * w->u) * vol, 0.0, 0.0};
                    cv_t s = csub(b->face_fc[0][C3(b, i + 1, j, k)], b->face_fc[0][x]);
                    s = cadd(s, b->face_fc[1][C3(b, i, j + 1, k)]);
                    s = csub(s, b->face_fc[1][x]);
                    s = cadd(s, b->face_fc[2][C3(b, i, j, k + 1)]);
                    s = csub(s, b->face_fc[2][x]);
                    b->fc[x] = csub(s, Qc);
                    if (c->cfg.turbulence == 200)   /* source/ComputeAllBlockVars.cpp:109-114 */
                        b->cfc_tur[x] = b->fc_tur[0][C3(b, i + 1, j, k)] - b->fc_tur[0][x] + b->fc_tur[1][C3(b, i, j + 1, k)] -
                                        b->fc_tur[1][x] + b->fc_tur[2][C3(b, i, j, k + 1)] - b->fc_tur[2][x];
                }
    }
    return 0;
}

/* source/ComputeAllBlockVars.cpp:223-255; Fd = 0 for the upwind schemes and
 * Fv = 0 for Euler (131-160) */
int orc_resflux(orc_ctx *c)
{
    orc_bc(c);
    orc_spectral(c);
    /* source/ComputeAllBlockVars.cpp:200-221 (fine grid) */
    if (c->cfg.viscous && c->cfg.turbulence == 100)
        for (int n = 0; n < c->nb; ++n)
            if (bl_model(c, &c->blk[n])) return -1;
    if (c->cfg.viscous && c->cfg.turbulence == 200)
        for (int n = 0; n < c->nb; ++n) sa_model(c, &c->blk[n]);
    orc_timestep(c);
    orc_convective(c);
    if (c->cfg.viscous)
        for (int n = 0; n < c->nb; ++n) viscous_flux(c, &c->blk[n]);
    const cv_t zero = {0, 0, 0, 0, 0};
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    b->res[x] = csub(csub(b->fc[x], c->cfg.viscous ? b->fv[x] : zero), zero);
                    if (c->cfg.turbulence == 200) b->res_ev[x] = b->cfc_tur[x] - b->cfv_tur[x];   /* line 249 */
                }
    }
    return 0;
}

/* ------------------------------------------------------------- LU-SGS */
/* source/LU_SGS_Marching.cpp:293-419; nb = neighbour cell, fa/fb its two faces
 * (higher index first), as delt_Fn_up/down read them. */
static cv_t delt_fn(orc_ctx *c, blk_t *b, long nb, const face_t *fa, const face_t *fb, double *dfe)
{
    double Gamma = c->nd.gamma;
    v3_t norm;
    norm.x = (fa->S.x / fa->s + fb->S.x / fb->s) / F32(2.0);
    norm.y = (fa->S.y / fa->s + fb->S.y / fb->s) / F32(2.0);
    norm.z = (fa->S.z / fa->s + fb->S.z / fb->s) / F32(2.0);
    double uvel_e = F32(-1.0) * b->cell[nb].c.y * b->w_rot;
    double vvel_e = b->cell[nb].c.x * b->w_rot;
    cv_t old = b->cv[nb];
    cv_t nw = cadd(old, b->dq[nb]);
    double dens = nw.r, uvel = nw.ru / nw.r, vvel = nw.rv / nw.r, wvel = nw.rw / nw.r, E = nw.re / nw.r;
    double press_new = (Gamma - F32(1.0)) * dens * (E - F32(0.5) * (uvel * uvel + vvel * vvel + wvel * wvel));
    double press_old = b->pv[nb].p;
    double Vn_old = old.ru / old.r * norm.x + old.rv / old.r * norm.y + old.rw / old.r * norm.z;
    double Vn_new = uvel * norm.x + vvel * norm.y + wvel * norm.z;
    double Vn_e = uvel_e * norm.x + vvel_e * norm.y + F32(0.0) * norm.z;
    cv_t d;
    d.r = nw.r * (Vn_new - Vn_e) - old.r * (Vn_old - Vn_e);
    d.ru = (nw.ru * (Vn_new - Vn_e) + press_new * norm.x) - (old.ru * (Vn_old - Vn_e) + press_old * norm.x);
    d.rv = (nw.rv * (Vn_new - Vn_e) + press_new * norm.y) - (old.rv * (Vn_old - Vn_e) + press_old * norm.y);
    d.rw = (nw.rw * (Vn_new - Vn_e) + press_new * norm.z) - (old.rw * (Vn_old - Vn_e) + press_old * norm.z);
    d.re = ((nw.re + press_new) * (Vn_new - Vn_e) + press_new * Vn_e) - ((old.re + press_old) * (Vn_old - Vn_e) + press_old * Vn_e);
    /* the nu~ flux difference, without the grid velocity (lines 352-357) */
    if (dfe) {
        double ev_old = b->ev[nb], ev_new = ev_old + b->dev[nb];
        *dfe = ev_new * Vn_new - ev_old * Vn_old;
    }
    return d;
}

static void lusgs_sweeps(orc_ctx *c, blk_t *b)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
    const int sa = c->cfg.turbulence == 200;
    double relax = (float)c->cfg.relax;
    const cv_t zero = {0, 0, 0, 0, 0};
    /* source/LU_SGS_Marching.cpp:61-160; the cells of one hyperplane depend
     * only on the previous one */
    const int par = PAR(b);
    for (int plane = 3; plane <= IM + JM + KM - 3; ++plane)
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= IM - 1; ++i)
            for (int j = 1; j <= JM - 1; ++j) {
                int k = plane - i - j;
                if (k < 1 || k > KM - 1) continue;
                long x = C3(b, i, j, k);
                double alpha = b->cell[x].vol / b->dt[x] + relax * (b->lamc[0][x] + b->lamc[1][x] + b->lamc[2][x]) +
                               F32(2.0) * (b->lamv[0][x] + b->lamv[1][x] + b->lamv[2][x]);
                double alpha2 = b->cell[x].vol / b->dt[x] + relax * (b->lamc[0][x] + b->lamc[1][x] + b->lamc[2][x]) +
                                F32(3.0) * (b->lamv[0][x] + b->lamv[1][x] + b->lamv[2][x]);
                cv_t A[3], V[3];
                double Ae[3], Ve[3];
                for (int d = 0; d < 3; ++d) {
                    int ok = d == 0 ? i != 1 : d == 1 ? j != 1 : k != 1;
                    if (!ok) { A[d] = zero; V[d] = zero; Ae[d] = 0.0; Ve[d] = 0.0; continue; }
                    long nb = d == 0 ? C3(b, i - 1, j, k) : d == 1 ? C3(b, i, j - 1, k) : C3(b, i, j, k - 1);
                    const face_t *fa = d == 0 ? &b->fi[x] : d == 1 ? &b->fj[x] : &b->fk[x];
                    const face_t *fb = d == 0 ? &b->fi[nb] : d == 1 ? &b->fj[nb] : &b->fk[nb];
                    double dfe = 0.0;
                    cv_t dF = delt_fn(c, b, nb, fa, fb, sa ? &dfe : 0);
                    double s = F32(0.5) * (fb->s + fa->s);
                    A[d] = cscale(cadd(cscale(dF, s), cscale(b->dq[nb], relax * b->lamc[d][nb])), F32(0.5));
                    V[d] = cscale(b->dq[nb], b->lamv[d][nb]);
                    if (sa) {
                        Ae[d] = F32(0.5) * (dfe * s + relax * b->lamc[d][nb] * b->dev[nb]);
                        Ve[d] = b->lamv[d][nb] * b->dev[nb];
                    }
                }
                cv_t s = cadd(cscale(b->res[x], F32(-1.0)), A[0]);
                s = cadd(s, A[1]); s = cadd(s, A[2]); s = cadd(s, V[0]); s = cadd(s, V[1]); s = cadd(s, V[2]);
                b->dq[x] = cdiv(s, alpha);
                if (sa)   /* source/LU_SGS_Marching.cpp:156-157 */
                    b->dev[x] = (F32(-1.0) * b->res_ev[x] + Ae[0] + Ae[1] + Ae[2] + Ve[0] + Ve[1] + Ve[2]) / alpha2;
            }
    /* source/LU_SGS_Marching.cpp:162-260 */
    for (int plane = IM + JM + KM - 3; plane >= 3; --plane)
        #pragma omp parallel for schedule(static) if (par)
        for (int i = IM - 1; i >= 1; --i)
            for (int j = JM - 1; j >= 1; --j) {
                int k = plane - i - j;
                if (k < 1 || k > KM - 1) continue;
                long x = C3(b, i, j, k);
                double alpha = b->cell[x].vol / b->dt[x] + relax * (b->lamc[0][x] + b->lamc[1][x] + b->lamc[2][x]) +
                               F32(2.0) * (b->lamv[0][x] + b->lamv[1][x] + b->lamv[2][x]);
                double alpha2 = b->cell[x].vol / b->dt[x] + relax * (b->lamc[0][x] + b->lamc[1][x] + b->lamc[2][x]) +
                                F32(3.0) * (b->lamv[0][x] + b->lamv[1][x] + b->lamv[2][x]);
                cv_t A[3], V[3];
                double Ae[3], Ve[3];
                for (int d = 0; d < 3; ++d) {
                    int ok = d == 0 ? i != IM - 1 : d == 1 ? j != JM - 1 : k != KM - 1;
                    if (!ok) { A[d] = zero; V[d] = zero; Ae[d] = 0.0; Ve[d] = 0.0; continue; }
                    long nb = d == 0 ? C3(b, i + 1, j, k) : d == 1 ? C3(b, i, j + 1, k) : C3(b, i, j, k + 1);
                    long n2 = d == 0 ? C3(b, i + 2, j, k) : d == 1 ? C3(b, i, j + 2, k) : C3(b, i, j, k + 2);
                    const face_t *fa = d == 0 ? &b->fi[n2] : d == 1 ? &b->fj[n2] : &b->fk[n2];
                    const face_t *fb = d == 0 ? &b->fi[nb] : d == 1 ? &b->fj[nb] : &b->fk[nb];
                    double dfe = 0.0;
                    cv_t dF = delt_fn(c, b, nb, fa, fb, sa ? &dfe : 0);
                    double s = F32(0.5) * (fb->s + fa->s);
                    A[d] = cscale(csub(cscale(dF, s), cscale(b->dq[nb], relax * b->lamc[d][nb])), F32(0.5));
                    V[d] = cscale(b->dq[nb], b->lamv[d][nb]);
                    if (sa) {
                        Ae[d] = F32(0.5) * (dfe * s - relax * b->lamc[d][nb] * b->dev[nb]);
                        Ve[d] = b->lamv[d][nb] * b->dev[nb];
                    }
                }
                cv_t s = csub(cscale(b->dq[x], alpha), A[0]);
                s = csub(s, A[1]); s = csub(s, A[2]); s = cadd(s, V[0]); s = cadd(s, V[1]); s = cadd(s, V[2]);
                b->dq[x] = cdiv(s, alpha);
                if (sa)   /* source/LU_SGS_Marching.cpp:256-257 */
                    b->dev[x] = (alpha2 * b->dev[x] - Ae[0] - Ae[1] - Ae[2] + Ve[0] + Ve[1] + Ve[2]) / alpha2;
            }
}

/* source/PrimitiveVars.cpp:7-29 */
static void primitive_vars(orc_ctx *c, blk_t *b)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    const int par = PAR(b);
    #pragma omp parallel for schedule(static) if (par)
    for (int i = 1; i < b->IM; ++i)
        for (int j = 1; j < b->JM; ++j)
            for (int k = 1; k < b->KM; ++k) {
                long x = C3(b, i, j, k);
                cv_t *q = &b->cv[x];
                pv_t *w = &b->pv[x];
                w->d = q->r;
                w->u = q->ru / q->r;
                w->v = q->rv / q->r;
                w->w = q->rw / q->r;
                double V2 = w->u * w->u + w->v * w->v + w->w * w->w;
                w->p = (Gamma - F32(1.0)) * (q->re - q->r * V2 / F32(2.0));
                w->t = w->p / w->d / Rgas;
            }
}

static inline int bad_cell(const pv_t *w)
{
    const double dmin = F32(1.0E-6), dmax = F32(1000.0), pmin = F32(1.0E-6), pmax = F32(1000.0), vmax = F32(1000.0);
    return w->d < dmin || w->d > dmax || w->p < pmin || w->p > pmax || w->u > vmax || w->v > vmax || w->w > vmax;
}

/* source/Limit_FlowVars.cpp:8-119 */
static void limit_vars(orc_ctx *c, blk_t *b)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    const double dmin = F32(1.0E-6), dmax = F32(1000.0), pmin = F32(1.0E-6), pmax = F32(1000.0), vmax = F32(1000.0);
    int IM = b->IM, JM = b->JM, KM = b->KM, nneg = 0;
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                pv_t *w = &b->pv[C3(b, i, j, k)];
                if (!bad_cell(w)) continue;
                nneg++;
                int kn = 0, i1 = 0, j1 = 0, k1 = 0;
                double d = 0.0, u = 0.0, v = 0.0, ww = 0.0, p = 0.0;
                for (int ia = -1; ia <= 1; ++ia)
                    for (int ja = -1; ja <= 1; ++ja)
                        for (int ka = -1; ka <= 1; ++ka) {
                            k1 = k + ka; j1 = j + ja; i1 = i + ia;
                            if (i1 > 0 && i1 < IM && j1 > 0 && j1 < JM && k1 > 0 && k1 < KM) {
                                const pv_t *q = &b->pv[C3(b, i1, j1, k1)];
                                if (!bad_cell(q)) {
                                    d = d + q->d; u = u + q->u; v = v + q->v; ww = ww + q->w; p = p + q->p;
                                    kn = kn + 1;
                                }
                            }
                        }
                stat_add(ORC_STAT_LIMIT_VARS);
                if (kn == 0) {
                    stat_add(ORC_STAT_LIMIT_VARS_KN0);
                    const pv_t *q = &b->pv[C3(b, i1, j1, k1)];
                    d = q->d; u = q->u; v = q->v; ww = q->w; p = q->p;
                    if (d < dmin) d = dmin;
                    if (d > dmax) d = dmax;
                    if (p < pmin) p = pmin;
                    if (p > pmax) p = pmax;
                    if (fabs(u) > vmax) u = u < 0 ? -fabs(vmax) : fabs(vmax);
                    if (fabs(v) > vmax) v = v < 0 ? -fabs(vmax) : fabs(vmax);
                    if (fabs(ww) > vmax) ww = ww < 0 ? -fabs(vmax) : fabs(vmax);
                } else {
                    double sn1 = (double)(1.0f / kn);
                    d = d * sn1; u = u * sn1; v = v * sn1; ww = ww * sn1; p = p * sn1;
                }
                w->d = d; w->u = u; w->v = v; w->w = ww; w->p = p;
                w->t = p / d / Rgas;
                cv_t *q = &b->cv[C3(b, i, j, k)];
                q->r = d; q->ru = d * u; q->rv = d * v; q->rw = d * ww;
                q->re = p / (Gamma - F32(1.0)) + d * (u * u + v * v + ww * ww) / F32(2.0);
            }
    if (nneg > 0) { b->flag_low_dt = 10; stat_add(ORC_STAT_LOW_DT); }
}

/* CFluidProps::Limit_SA (source/Limit_FlowVars.cpp:122-168): nu~ below 1e-8
 * is raised to it; above 1000 it becomes the mean of the admissible
 * neighbours (k outermost), or 1000 when there are none */
static void limit_sa(blk_t *b)
{
    const double SAmin = F32(1.0E-8), LSAmax = F32(1000.0);
    int IM = b->IM, JM = b->JM, KM = b->KM;
    for (int i = 1; i <= IM - 1; ++i)
        for (int j = 1; j <= JM - 1; ++j)
            for (int k = 1; k <= KM - 1; ++k) {
                double *e = &b->ev[C3(b, i, j, k)];
                if (*e < SAmin) *e = SAmin;
                if (*e > LSAmax) {
                    double s0 = 0.0;
                    int kn = 0;
                    for (int ka = -1; ka <= 1; ++ka)
                        for (int ja = -1; ja <= 1; ++ja)
                            for (int ia = -1; ia <= 1; ++ia) {
                                int k1 = k + ka, j1 = j + ja, i1 = i + ia;
                                if (i1 > 0 && i1 < IM && j1 > 0 && j1 < JM && k1 > 0 && k1 < KM) {
                                    double v = b->ev[C3(b, i1, j1, k1)];
                                    if (v >= SAmin && v <= LSAmax) { s0 = s0 + v; kn = kn + 1; }
                                }
                            }
                    stat_add(ORC_STAT_LIMIT_SA);
                    if (kn == 0) stat_add(ORC_STAT_LIMIT_SA_KN0);
                    *e = kn == 0 ? LSAmax : s0 / kn;
                }
            }
}

/* source/LU_SGS_Marching.cpp:10-291 (with the nu~ equation of Spalart-Allmaras) */
int orc_lusgs(orc_ctx *c)
{
    const int sa = c->cfg.turbulence == 200;
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    b->cv_old[C3(b, i, j, k)] = b->cv[C3(b, i, j, k)];
                    if (sa) b->ev_old[C3(b, i, j, k)] = b->ev[C3(b, i, j, k)];
                }
    }
    if (!c->force) {
        if (orc_resflux(c)) return -1;
    } else {
        /* CMesh::ComputeForceFunction / AddForceFunction (source/ComputeAllBlockVars.cpp:
         * 269-309): the coarse residual becomes the restricted fine one, through
         * Res + (Res_2h - Res); lambda, dt and the Dummy layers stay as the last
         * coarse ResFlux left them */
        for (int n = 0; n < c->nb; ++n) {
            blk_t *b = &c->blk[n];
            for (int i = 1; i <= b->IM - 1; ++i)
                for (int j = 1; j <= b->JM - 1; ++j)
                    for (int k = 1; k <= b->KM - 1; ++k) {
                        long x = C3(b, i, j, k);
                        cv_t qf = csub(b->res2h[x], b->res[x]);
                        b->res[x] = cadd(b->res[x], qf);
                    }
        }
    }
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        lusgs_sweeps(c, b);
        const int par = PAR(b);
        #pragma omp parallel for schedule(static) if (par)
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    b->cv[x] = cadd(b->cv[x], b->dq[x]);
                    if (sa) b->ev[x] = b->ev[x] + b->dev[x];
                }
    }
    for (int n = 0; n < c->nb; ++n) primitive_vars(c, &c->blk[n]);
    if (c->cfg.limiter)
        for (int n = 0; n < c->nb; ++n) {
            limit_vars(c, &c->blk[n]);
            if (sa) limit_sa(&c->blk[n]);
        }
    return 0;
}

/* source/ComputeAllBlockVars.cpp:318-372 */
int orc_res_calculate(orc_ctx *c, double *res6)
{
    double r[5] = {0, 0, 0, 0, 0}, re = 0.0;
    for (int n = 0; n < c->nb; ++n) {
        blk_t *b = &c->blk[n];
        /* the OpenMP reduction sums each block from zero, then adds it to the total */
        double t[5] = {0, 0, 0, 0, 0}, te = 0.0;
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    if (c->cfg.turbulence == 200)
                        te = te + fabs((b->ev_old[x] - b->ev[x]) / b->ev_old[x]) * b->cell[x].vol;
                    const cv_t *o = &b->cv_old[x], *q = &b->cv[x];
                    cv_t m;
                    m.r = fabs((o->r - q->r) / o->r);
                    m.ru = fabs((o->ru - q->ru) / o->ru);
                    m.rv = fabs((o->rv - q->rv) / o->rv);
                    m.rw = fabs((o->rw - q->rw) / o->rw);
                    m.re = fabs((o->re - q->re) / o->re);
                    m = cscale(m, b->cell[x].vol);
                    t[0] = t[0] + m.r; t[1] = t[1] + m.ru; t[2] = t[2] + m.rv; t[3] = t[3] + m.rw; t[4] = t[4] + m.re;
                }
        for (int q = 0; q < 5; ++q) r[q] = r[q] + t[q];
        re = re + te;
    }
    for (int q = 0; q < 5; ++q) c->res[q] = r[q] / c->total_vol;
    c->res[5] = c->cfg.turbulence == 200 ? re / c->total_vol : 0.0;
    if (res6) for (int q = 0; q < 6; ++q) res6[q] = c->res[q];
    return 0;
}

/* ----------------------------------------------------- performance data */
/* source/Ouput.cpp:9-95 */
static void section(orc_ctx *c, blk_t *b, int K, double *mass, double *pt, double *tt)
{
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    double m = 0.0, mp = 0.0, ptm = 0.0, mt = 0.0, ttm = 0.0;
    for (int i = 1; i <= b->IM - 1; ++i)
        for (int j = 1; j <= b->JM - 1; ++j) {
            const pv_t *a = &b->pv[C3(b, i, j, K)], *z = &b->pv[C3(b, i, j, K - 1)];
            const face_t *f = &b->fk[C3(b, i, j, K)];
            double dens = F32(0.5) * (a->d + z->d), u = F32(0.5) * (a->u + z->u), v = F32(0.5) * (a->v + z->v),
                   w = F32(0.5) * (a->w + z->w);
            double mc = dens * (u * f->S.x + v * f->S.y + w * f->S.z);
            m = m + mc;
            double temp = F32(0.5) * (a->t + z->t);
            double Ma = sqrt(u * u + v * v + w * w) / sqrt(Gamma * Rgas * temp);
            double Pt_cell = a->p * pow(F32(1.0) + (Gamma - F32(1.0)) / F32(2.0) * Ma * Ma, Gamma / (Gamma - F32(1.0)));
            ptm = ptm + pow(Pt_cell, (Gamma - F32(1.0)) / Gamma) * mc;
            mp = mp + mc;
            double Tt_cell = a->t * pow(1 + (Gamma - F32(1.0)) / F32(2.0) * Ma * Ma, 1);
            ttm = ttm + Tt_cell * mc;
            mt = mt + mc;
        }
    *mass = m * c->nd.dens_ref * c->nd.vel_ref;
    *pt = pow(ptm / mp, Gamma / (Gamma - F32(1.0))) * c->nd.dens_ref * c->nd.vel_ref * c->nd.vel_ref;
    *tt = ttm / mt * c->nd.temp_ref;
}

/* source/Ouput.cpp:97-107,282-300 (first inlet and first outlet patch) */
int orc_perf(orc_ctx *c, double *out5)
{
    blk_t *bi = 0, *bo = 0;
    int ki = 0, ko = 0;
    for (int n = 0; n < c->nb; ++n)
        for (int q = 0; q < c->blk[n].npatch; ++q) {
            patch_t *P = &c->blk[n].patch[q];
            if (P->kind == 0 && !bi) { bi = &c->blk[n]; ki = P->ks; }
            if (P->kind == 1 && !bo) { bo = &c->blk[n]; ko = P->ks; }
        }
    if (!bi || !bo) return -1;
    double mi, pti, tti, mo, pto, tto;
    section(c, bi, ki, &mi, &pti, &tti);
    section(c, bo, ko, &mo, &pto, &tto);
    mi = mi * bi->nblade;
    mo = mo * bo->nblade;
    double Gamma = c->nd.gamma, Rgas = c->nd.rgas;
    double PR = pto / pti;
    double L_adk = Gamma / (Gamma - F32(1.0)) * Rgas * tti * (pow(PR, (Gamma - F32(1.0)) / Gamma) - F32(1.0));
    double L_k = Gamma / (Gamma - F32(1.0)) * Rgas * tti * (tto / tti - F32(1.0));
    out5[0] = mi;
    out5[1] = mo;
    out5[2] = pto / pti;
    out5[3] = L_adk / L_k;
    out5[4] = fabs((mi - mo) / mi);
    return 0;
}

double orc_cfl(orc_ctx *c) { return c->cfl; }

/* ------------------------------------------------------ two-grid cycle */
/* CSolver::Prolong, fine to coarse (source/Prolong_Interpolate.cpp:7-50): the
 * volume-weighted state and the summed residual of the eight fine cells. */
static void mg_restrict(orc_ctx *f, orc_ctx *m)
{
    for (int n = 0; n < m->nb; ++n) {
        blk_t *b = &f->blk[n], *B = &m->blk[n];
        for (int i = 1; i <= B->IM - 1; ++i)
            for (int j = 1; j <= B->JM - 1; ++j)
                for (int k = 1; k <= B->KM - 1; ++k) {
                    const int I[2] = {2 * i - 1, 2 * i}, J[2] = {2 * j - 1, 2 * j}, K[2] = {2 * k - 1, 2 * k};
                    cv_t sq = {0, 0, 0, 0, 0}, sr = {0, 0, 0, 0, 0};
                    int first = 1;
                    for (int a = 0; a < 2; ++a)
                        for (int bb = 0; bb < 2; ++bb)
                            for (int cc = 0; cc < 2; ++cc) {
                                long y = C3(b, I[a], J[bb], K[cc]);
                                cv_t q = cscale(b->cv[y], b->cell[y].vol);
                                sq = first ? q : cadd(sq, q);
                                sr = first ? b->res[y] : cadd(sr, b->res[y]);
                                first = 0;
                            }
                    long x = C3(B, i, j, k);
                    B->cv2h[x] = cdiv(sq, B->cell[x].vol);
                    B->res2h[x] = sr;
                    B->cv[x] = B->cv2h[x];
                }
    }
}

/* CSolver::ComputeDeltConsVars (source/Prolong_Interpolate.cpp:159-176) */
static void mg_delta(orc_ctx *m)
{
    for (int n = 0; n < m->nb; ++n) {
        blk_t *B = &m->blk[n];
        for (int i = 1; i <= B->IM - 1; ++i)
            for (int j = 1; j <= B->JM - 1; ++j)
                for (int k = 1; k <= B->KM - 1; ++k) {
                    long x = C3(B, i, j, k);
                    B->delt[x] = csub(B->cv[x], B->cv2h[x]);
                }
    }
}

/* CMesh::CornerTreatment (source/ComputeAllBlockVars.cpp:374-439) on one
 * array: the twelve edges of the ghost box from the two neighbouring face
 * ghosts, then the eight corners from three edges and three face ghosts. */
static void corner_one(const blk_t *b, cv_t *a)
{
    int IM = b->IM, JM = b->JM, KM = b->KM;
#define A(i, j, k) a[C3(b, i, j, k)]
#define HALF(p, q) cscale(cadd(p, q), F32(0.5))
    for (int i = 1; i <= IM - 1; ++i) {
        A(i, 0, 0) = HALF(A(i, 1, 0), A(i, 0, 1));
        A(i, JM, 0) = HALF(A(i, JM - 1, 0), A(i, JM, 1));
        A(i, 0, KM) = HALF(A(i, 0, KM - 1), A(i, 1, KM));
        A(i, JM, KM) = HALF(A(i, JM - 1, KM), A(i, JM, KM - 1));
    }
    for (int j = 1; j <= JM - 1; ++j) {
        A(0, j, 0) = HALF(A(1, j, 0), A(0, j, 1));
        A(0, j, KM) = HALF(A(1, j, KM), A(0, j, KM - 1));
        A(IM, j, KM) = HALF(A(IM - 1, j, KM), A(IM, j, KM - 1));
        A(IM, j, 0) = HALF(A(IM - 1, j, 0), A(IM, j, 1));
    }
    for (int k = 1; k <= KM - 1; ++k) {
        A(0, 0, k) = HALF(A(1, 0, k), A(0, 1, k));
        A(IM, 0, k) = HALF(A(IM, 1, k), A(IM - 1, 0, k));
        A(0, JM, k) = HALF(A(1, JM, k), A(0, JM - 1, k));
        A(IM, JM, k) = HALF(A(IM - 1, JM, k), A(IM, JM - 1, k));
    }
#define CORNER(i, j, k, i1, j1, k1) \
    A(i, j, k) = cdiv(csub(cscale(cadd(cadd(A(i1, j, k), A(i, j1, k)), A(i, j, k1)), F32(2.0)), \
                           cadd(cadd(A(i1, j1, k), A(i1, j, k1)), A(i, j1, k1))), F32(3.0))
    CORNER(0, 0, 0, 1, 1, 1);
    CORNER(IM, 0, 0, IM - 1, 1, 1);
    CORNER(0, JM, 0, 1, JM - 1, 1);
    CORNER(IM, JM, 0, IM - 1, JM - 1, 1);
    CORNER(IM, 0, KM, IM - 1, 1, KM - 1);
    CORNER(0, 0, KM, 1, 1, KM - 1);
    CORNER(0, JM, KM, 1, JM - 1, KM - 1);
    CORNER(IM, JM, KM, IM - 1, JM - 1, KM - 1);
#undef CORNER
#undef HALF
#undef A
}

static void mg_corner(orc_ctx *m)
{
    for (int n = 0; n < m->nb; ++n) {
        corner_one(&m->blk[n], m->blk[n].cv);
        corner_one(&m->blk[n], m->blk[n].delt);
    }
}

/* CSolver::Interpolate, coarse to fine (source/Prolong_Interpolate.cpp:52-156):
 * weights 27, 9, 3, 1 (/64) on the nearest coarse cell, the three that share a
 * face with it towards the fine cell, the three that share an edge and the one
 * that shares a corner.  state = 1 sets the fine state from the coarse one
 * (the end of the coarse phase), else the coarse correction is added. */
static void mg_interpolate(orc_ctx *f, orc_ctx *m, int state)
{
    const double a1 = (double)(27.0f / 64.0f), a2 = (double)(9.0f / 64.0f), a3 = (double)(3.0f / 64.0f),
                 a4 = (double)(1.0f / 64.0f);
    for (int n = 0; n < f->nb; ++n) {
        blk_t *b = &f->blk[n];
        const blk_t *B = &m->blk[n];
        const cv_t *src = state ? B->cv : B->delt;
        for (int i = 1; i <= b->IM - 1; ++i)
            for (int j = 1; j <= b->JM - 1; ++j)
                for (int k = 1; k <= b->KM - 1; ++k) {
                    int i1 = i % 2 == 0 ? i / 2 : i / 2 + 1, i2 = i % 2 == 0 ? i / 2 + 1 : i / 2;
                    int j1 = j % 2 == 0 ? j / 2 : j / 2 + 1, j2 = j % 2 == 0 ? j / 2 + 1 : j / 2;
                    int k1 = k % 2 == 0 ? k / 2 : k / 2 + 1, k2 = k % 2 == 0 ? k / 2 + 1 : k / 2;
#define S(p, q, r) src[C3(B, p, q, r)]
                    cv_t t1 = cscale(S(i1, j1, k1), a1);
                    cv_t t2 = cscale(cadd(cadd(S(i1, j1, k2), S(i1, j2, k1)), S(i2, j1, k1)), a2);
                    cv_t t3 = cscale(cadd(cadd(S(i2, j2, k1), S(i1, j2, k2)), S(i2, j1, k2)), a3);
                    cv_t t4 = cscale(S(i2, j2, k2), a4);
#undef S
                    cv_t v = cadd(cadd(cadd(t1, t2), t3), t4);
                    long x = C3(b, i, j, k);
                    b->cv[x] = state ? v : cadd(b->cv[x], v);
                }
    }
}

static void prim_and_limit(orc_ctx *c)
{
    for (int n = 0; n < c->nb; ++n) primitive_vars(c, &c->blk[n]);
    if (c->cfg.limiter)
        for (int n = 0; n < c->nb; ++n) limit_vars(c, &c->blk[n]);
}

/* One iteration of source/Solver.cpp:61-103 (reuse_solution = N): the coarse
 * phase up to Coarse_iter, then fine LU-SGS + coarse correction. */
static int two_grid(orc_ctx *c)
{
    orc_ctx *m = c->coarse;
    m->cfl = c->cfl;   /* CTimeDiscr::CFL is shared by the levels */
    if (c->iter <= c->cfg.coarse_iter) {
        if (orc_lusgs(m)) return -1;
        orc_res_calculate(m, c->res);
        if (c->iter == c->cfg.coarse_iter) {
            mg_corner(m);
            mg_interpolate(c, m, 1);
            prim_and_limit(c);
        }
        return 0;
    }
    m->force = 1;
    if (orc_lusgs(c)) return -1;
    orc_res_calculate(c, 0);
    if (orc_resflux(c)) return -1;
    mg_restrict(c, m);
    prim_and_limit(m);
    if (orc_lusgs(m)) return -1;
    mg_delta(m);
    mg_corner(m);
    mg_interpolate(c, m, 0);
    prim_and_limit(c);
    return 0;
}

/* source/Solver.cpp:47-66 and source/Structured_FV3D.cpp:142-158 */
int orc_iterate(orc_ctx *c, int n, double *hist)
{
    const double amp = F32(1.2), cmax = F32(10.0);
    for (int it = 0; it < n; ++it) {
        c->iter++;
        if (c->iter > 1) {
            if (c->res[0] <= c->cfl_res) {
                c->cfl = c->cfl * amp;
                c->cfl_res = c->cfl_res / F32(10.0);
            }
            if (c->cfl > cmax) c->cfl = cmax;
        }
        double p[5];
        if (c->ng == 2) {
            if (two_grid(c)) return -1;
            orc_perf(c->iter <= c->cfg.coarse_iter ? c->coarse : c, p);
        } else {
            if (orc_lusgs(c)) return -1;
            orc_res_calculate(c, 0);
            orc_perf(c, p);
        }
        if (hist) {
            double *h = hist + (long)it * 7;
            h[0] = c->iter; h[1] = c->res[0]; h[2] = p[0]; h[3] = p[1]; h[4] = p[2]; h[5] = p[3]; h[6] = c->cfl;
        }
    }
    return 0;
}

/* -------------------------------------------------------------- output */
long orc_get(orc_ctx *c, int block, const char *name, double *out)
{
    if (block < 1 || block > c->nb) return -1;
    blk_t *b = &c->blk[block - 1];
    int IM = b->IM, JM = b->JM, KM = b->KM;
    long n = 0;
    if (!strcmp(name, "cv") || !strcmp(name, "pv")) {
        int pvq = name[0] == 'p';
        for (int i = 0; i <= IM; ++i)
            for (int j = 0; j <= JM; ++j)
                for (int k = 0; k <= KM; ++k) {
                    long x = C3(b, i, j, k);
                    if (pvq) { const pv_t *w = &b->pv[x]; double t[6] = {w->d, w->u, w->v, w->w, w->p, w->t}; memcpy(out + n, t, sizeof t); n += 6; }
                    else { const cv_t *q = &b->cv[x]; double t[5] = {q->r, q->ru, q->rv, q->rw, q->re}; memcpy(out + n, t, sizeof t); n += 5; }
                }
        return n;
    }
    const cv_t *c5 = 0;
    const double *s1 = 0;
    if (!strcmp(name, "res")) c5 = b->res;
    else if (!strcmp(name, "dq")) c5 = b->dq;
    else if (!strcmp(name, "fc")) c5 = b->fc;
    else if (!strcmp(name, "dt")) s1 = b->dt;
    else if (!strcmp(name, "lamcI")) s1 = b->lamc[0];
    else if (!strcmp(name, "lamcJ")) s1 = b->lamc[1];
    else if (!strcmp(name, "lamcK")) s1 = b->lamc[2];
    else if (!strcmp(name, "lamvI")) s1 = b->lamv[0];
    else if (!strcmp(name, "lamvJ")) s1 = b->lamv[1];
    else if (!strcmp(name, "lamvK")) s1 = b->lamv[2];
    else if (!strcmp(name, "omega")) s1 = b->omega;
    else if (!strcmp(name, "fv")) c5 = b->fv;
    else if (!strcmp(name, "resev")) s1 = b->res_ev;
    else if (!strcmp(name, "cfctur")) s1 = b->cfc_tur;
    else if (!strcmp(name, "cfvtur")) s1 = b->cfv_tur;
    else if (!strcmp(name, "dev")) s1 = b->dev;
    else if (!strcmp(name, "dwall")) s1 = b->dwall;
    else if (!strcmp(name, "mut") || !strcmp(name, "ev")) {
        const double *src = name[0] == 'm' ? b->mut : b->ev;
        for (int i = 0; i <= IM; ++i)
            for (int j = 0; j <= JM; ++j)
                for (int k = 0; k <= KM; ++k) out[n++] = src[C3(b, i, j, k)];
        return n;
    }
    else if (!strcmp(name, "gevI") || !strcmp(name, "gevJ") || !strcmp(name, "gevK")) {
        int d = name[3] - 'I';
        for (int i = 1; i <= IM - (d != 0); ++i)
            for (int j = 1; j <= JM - (d != 1); ++j)
                for (int k = 1; k <= KM - (d != 2); ++k) {
                    memcpy(out + n, b->gev[d] + C3(b, i, j, k) * 3, 3 * sizeof(double));
                    n += 3;
                }
        return n;
    }
    else if (!strcmp(name, "fcI") || !strcmp(name, "fcJ") || !strcmp(name, "fcK") || !strcmp(name, "fvI") ||
             !strcmp(name, "fvJ") || !strcmp(name, "fvK")) {
        int d = name[2] - 'I';
        const cv_t *src = name[1] == 'c' ? b->face_fc[d] : b->face_fv[d];
        for (int i = 1; i <= IM - (d != 0); ++i)
            for (int j = 1; j <= JM - (d != 1); ++j)
                for (int k = 1; k <= KM - (d != 2); ++k) {
                    const cv_t *q = &src[C3(b, i, j, k)];
                    double t[5] = {q->r, q->ru, q->rv, q->rw, q->re};
                    memcpy(out + n, t, sizeof t);
                    n += 5;
                }
        return n;
    }
    else if (!strcmp(name, "gradI") || !strcmp(name, "gradJ") || !strcmp(name, "gradK")) {
        int d = name[4] - 'I';
        for (int i = 1; i <= IM - (d != 0); ++i)
            for (int j = 1; j <= JM - (d != 1); ++j)
                for (int k = 1; k <= KM - (d != 2); ++k) {
                    memcpy(out + n, b->grad[d] + C3(b, i, j, k) * 12, 12 * sizeof(double));
                    n += 12;
                }
        return n;
    }
    else if (!strcmp(name, "metI") || !strcmp(name, "metJ") || !strcmp(name, "metK") || !strcmp(name, "metC")) {
        int d = name[3] == 'C' ? 3 : name[3] - 'I';
        const face_t *F = d == 0 ? b->fi : d == 1 ? b->fj : b->fk;
        for (int i = 1; i <= IM - (d != 0); ++i)
            for (int j = 1; j <= JM - (d != 1); ++j)
                for (int k = 1; k <= KM - (d != 2); ++k) {
                    long x = C3(b, i, j, k);
                    memcpy(out + n, d == 3 ? b->cell[x].g : F[x].g, 9 * sizeof(double));
                    n += 9;
                }
        return n;
    }
    else if (!strcmp(name, "SI") || !strcmp(name, "SJ") || !strcmp(name, "SK")) {
        int d = name[1] - 'I';
        const face_t *F = d == 0 ? b->fi : d == 1 ? b->fj : b->fk;
        for (int i = 1; i <= IM - (d != 0); ++i)
            for (int j = 1; j <= JM - (d != 1); ++j)
                for (int k = 1; k <= KM - (d != 2); ++k) {
                    const face_t *f = &F[C3(b, i, j, k)];
                    out[n++] = f->S.x; out[n++] = f->S.y; out[n++] = f->S.z;
                }
        return n;
    }
    else if (!strcmp(name, "vol")) {
        for (int i = 1; i <= IM - 1; ++i)
            for (int j = 1; j <= JM - 1; ++j)
                for (int k = 1; k <= KM - 1; ++k) out[n++] = b->cell[C3(b, i, j, k)].vol;
        return n;
    }
    if (c5 || s1) {
        for (int i = 1; i <= IM - 1; ++i)
            for (int j = 1; j <= JM - 1; ++j)
                for (int k = 1; k <= KM - 1; ++k) {
                    long x = C3(b, i, j, k);
                    if (c5) { const cv_t *q = &c5[x]; double t[5] = {q->r, q->ru, q->rv, q->rw, q->re}; memcpy(out + n, t, sizeof t); n += 5; }
                    else out[n++] = s1[x];
                }
        return n;
    }
    static const char *dn[6] = {"dIS", "dIP", "dJS", "dJP", "dKS", "dKP"};
    for (int d = 0; d < 6; ++d) {
        if (strncmp(name, dn[d], 3) != 0) continue;
        int pvq = strstr(name, "_pv") != 0;
        int ra = d < 2 ? ND : IM - 1, rb = d < 2 ? JM - 1 : d < 4 ? ND : JM - 1, rc = d < 4 ? KM - 1 : ND;
        for (int a = 1; a <= ra; ++a)
            for (int bb = 1; bb <= rb; ++bb)
                for (int cc = 1; cc <= rc; ++cc) {
                    long x = DI(b, d, a, bb, cc);
                    if (pvq) { const pv_t *w = &b->dpv[d][x]; double t[6] = {w->d, w->u, w->v, w->w, w->p, w->t}; memcpy(out + n, t, sizeof t); n += 6; }
                    else { const cv_t *q = &b->dcv[d][x]; double t[5] = {q->r, q->ru, q->rv, q->rw, q->re}; memcpy(out + n, t, sizeof t); n += 5; }
                }
        return n;
    }
    return -2;
}

/* ------------------------------------------------------------ test hooks */
/* Point evaluations of the face flux and the reconstruction, so that tests can
 * check them against closed forms.  States are in the face frame for the flux
 * (q = rho, rho*u_n, rho*u_t1, rho*u_t2, rho*E); the reconstruction takes
 * primitive records (rho,u,v,w,p,T) for var 200 and conservative ones for var
 * 100, and returns conservative left/right states. */
int orc_test_flux(int scheme, double gamma, const double *ql, const double *qr, double que, double *out)
{
    cv_t L = {ql[0], ql[1], ql[2], ql[3], ql[4]}, R = {qr[0], qr[1], qr[2], qr[3], qr[4]}, F;
    if (scheme == 300) F = roe(gamma, L, R, que);
    else if (scheme == 200) F = steger_warming(gamma, L, R, que);
    else return -1;
    out[0] = F.r; out[1] = F.ru; out[2] = F.rv; out[3] = F.rw; out[4] = F.re;
    return 0;
}

/* Sets the whole primitive box (ghost cells included, (IM+1)(JM+1)(KM+1)x6,
 * i slowest) of a block, then evaluates the viscous stages on it without
 * applying boundary conditions: curl, face gradients and viscous fluxes. */
int orc_test_viscous(orc_ctx *c, int block, const double *pv_box)
{
    if (block < 1 || block > c->nb) return -1;
    blk_t *b = &c->blk[block - 1];
    long box = (long)(b->IM + 1) * (b->JM + 1) * (b->KM + 1);
    for (long x = 0; x < box; ++x) {
        const double *p = pv_box + x * 6;
        pv_t w = {p[0], p[1], p[2], p[3], p[4], p[5]};
        b->pv[x] = w;
    }
    curl(b);
    gradients(b);
    viscous_flux(c, b);
    return 0;
}

double orc_test_sutherland(orc_ctx *c, double temp) { return sutherland(&c->nd, temp); }

int orc_test_recon(int var, int scheme, double gamma, const double *ff, const double *f, const double *b,
                   const double *bb, double *l, double *r)
{
    cv_t L, R;
    if (var == 200) {
        pv_t a = {ff[0], ff[1], ff[2], ff[3], ff[4], ff[5]}, c = {f[0], f[1], f[2], f[3], f[4], f[5]};
        pv_t d = {b[0], b[1], b[2], b[3], b[4], b[5]}, e = {bb[0], bb[1], bb[2], bb[3], bb[4], bb[5]};
        recon_prim(scheme, gamma, a, c, d, e, &L, &R);
    } else if (var == 100) {
        cv_t a = {ff[0], ff[1], ff[2], ff[3], ff[4]}, c = {f[0], f[1], f[2], f[3], f[4]};
        cv_t d = {b[0], b[1], b[2], b[3], b[4]}, e = {bb[0], bb[1], bb[2], bb[3], bb[4]};
        recon_cons(scheme, gamma, a, c, d, e, &L, &R);
    } else if (var == 300) {
        /* primitive inputs (rho u v w p T); the conservative states are formed
         * from them, and Rgas is the non-dimensional 1/gamma */
        const double *in[4] = {ff, f, b, bb};
        pv_t w[4];
        cv_t q[4];
        for (int n = 0; n < 4; ++n) {
            w[n] = (pv_t){in[n][0], in[n][1], in[n][2], in[n][3], in[n][4], in[n][5]};
            q[n] = (cv_t){w[n].d, w[n].d * w[n].u, w[n].d * w[n].v, w[n].d * w[n].w,
                          w[n].p / (gamma - F32(1.0)) + w[n].d * (w[n].u * w[n].u + w[n].v * w[n].v + w[n].w * w[n].w) / F32(2.0)};
        }
        recon_char(scheme, gamma, 1.0 / gamma, w[1], w[2], q[0], q[1], q[2], q[3], &L, &R);
    } else {
        return -1;
    }
    l[0] = L.r; l[1] = L.ru; l[2] = L.rv; l[3] = L.rw; l[4] = L.re;
    r[0] = R.r; r[1] = R.ru; r[2] = R.rv; r[3] = R.rw; r[4] = R.re;
    return 0;
}
