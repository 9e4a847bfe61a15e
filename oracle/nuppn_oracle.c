/* CPU oracle for the network-integration path.  TEST INFRASTRUCTURE ONLY (see nuppn_oracle.h).
 *
 * Every routine follows the reference file:line given in its comment, in the reference's own
 * floating-point operation order (left-to-right Fortran expression evaluation, no contraction;
 * the Makefile builds with -ffp-contract=off), so that its results can be compared with the HIP
 * path bit for bit where the arithmetic allows it.
 */
#define _GNU_SOURCE
#include "nuppn_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* error codes, utils/source/errors.F90:5-16 */
enum { SLV_CNVRG = 1, MAX_SUBSTEPS, BAD_STEP, NEG_ABUNDS, HUGE_ABUNDS, SUM_X, GOT_NAN, BAD_YE,
       NSE_CNVRG, NSE_WEAK, SMALL_DT };

#define SUM_TOL 1.0e-6          /* solver_knobs.F90:26 */
#define SMALL_C 1.0e-30         /* constants.F90:81 */
#define NLEVELS 8               /* bader_deuflhard.F90:27 */
#define ORC_NT9 5001            /* num_t9_preproc, reaclib.F90:155 */

struct orc_ctx {
    int nsp, nre, igamma;
    int *k;                 /* [8*nre] */
    int *rtype;
    double *coef;           /* [7*nre] */
    double *aw, *zw;
    unsigned char *consider; /* considerisotope: all but gamma */
    orc_knobs kn;
    int lsolver;
    /* makemap */
    int nnz;
    int *rowptr, *colind, *map, *row_of;
    int *diag;              /* slot of (i,i) or -1 */
    /* work */
    double *v, *jac, *jac_lo, *jac_cpy, *xn, *dxdt, *dxdt_lo, *rhs, *red_val, *red_rhs, *red_xlo;
    int *jrow, *jcol, *red_rowptr, *red_colind, *iexpand;
    unsigned char *mask;
    int nvar1, red_nnz;
    /* SAVE'd state of the reference */
    int frst;               /* integrate_network local `frst` (solver.F90:75) */
    int reduce;             /* bader_deuflhard_integration local `reduce` (bader_deuflhard.F90:152) */
    int kopt, kmax;         /* module variables, bader_deuflhard.F90:29-32 */
    int nseq[NLEVELS + 1];
    double a[NLEVELS + 1], alf[NLEVELS][NLEVELS];
    double *qcol, *yerr;    /* [NLEVELS*nsp] column-major like the reference (nsp, nlevels) */
    double err[NLEVELS], xcol[NLEVELS];
    double errmax, abuerr, negerr, sumerr, xest;
    /* the REACLIB T9 cache (reaclib.F90:155-156, 609-646) */
    int use_cache;
    double *rcache;         /* [ORC_NT9][nre] */
    double rc_t9[ORC_NT9], rc_dt9;
    /* rate physics beyond the bare REACLIB sum (orc_set_rate_physics): partition functions on
     * REACLIB's T9 grid and the factors of its reverse rates, Chugunov screening, detailed
     * balance; all off by default */
    int pf_n;
    double *pf_t9, *pf_tab;  /* [pf_n], [nsp][pf_n] */
    int *rev_ch, *rev_part;  /* [nre], [6*nre] */
    int screen, detailed_balance;
    int *irev;               /* [nre] reverse_init's pairing, -1 */
    double *g_spin, *mass_g, *q_mev;   /* [nsp] */
};

/* ------------------------------------------------------------------------------------------ */
/* reference SuperLU, opened lazily (the library also holds BLAS-calling routines of the
 * transposed solve that the path never reaches; lazy binding leaves them unresolved)          */

typedef void (*slu_solve_t)(int, int, int, int, double *, int *, int *, double *, int *);
typedef void (*dgssv_t)(int *, int *, int *, int *, double *, int *, int *, double *, int *,
                        long long *, int *);
typedef int (*colperm_t)(int, int, const int *, const int *, int *, int *, int *);
typedef int (*fstats_t)(long long, int, int *, int *, int *, int *, int *);
typedef int (*fsup_t)(long long, int *);

static void *g_ref;
static slu_solve_t g_slu_solve;
static dgssv_t g_dgssv;
static colperm_t g_colperm;
static fstats_t g_fstats;
static fsup_t g_fsup;

int orc_ref_open(const char *path)
{
    if (g_ref) return 0;
    void *h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
    if (!h) return 1;
    g_slu_solve = (slu_solve_t) dlsym(h, "slu_solve");
    g_dgssv = (dgssv_t) dlsym(h, "c_fortran_dgssv_");
    g_colperm = (colperm_t) dlsym(h, "ref_colperm");
    g_fstats = (fstats_t) dlsym(h, "ref_factor_stats");
    g_fsup = (fsup_t) dlsym(h, "ref_factor_supernodes");
    if (!g_slu_solve || !g_dgssv || !g_colperm || !g_fstats || !g_fsup) { dlclose(h); return 2; }
    g_ref = h;
    return 0;
}

int orc_ref_available(void) { return g_ref != NULL; }

void orc_ref_slu_solve(int n, int nnz, int isrow, int shiftindex, double *val, int *colind,
                       int *rowptr, double *rhs, int *ierr)
{
    g_slu_solve(n, nnz, isrow, shiftindex, val, colind, rowptr, rhs, ierr);
}

void orc_ref_dgssv(int iopt, int n, int nnz, int nrhs, double *values, int *rowind, int *colptr,
                   double *b, int ldb, long long *factors, int *info)
{
    g_dgssv(&iopt, &n, &nnz, &nrhs, values, rowind, colptr, b, &ldb, factors, info);
}

int orc_ref_colperm(int n, int nnz, const int *colptr, const int *rowind, int *perm_colamd,
                    int *perm_final, int *etree)
{
    return g_colperm(n, nnz, colptr, rowind, perm_colamd, perm_final, etree);
}

int orc_ref_factor_stats(long long handle, int n, int *perm_c, int *perm_r, int *nnz_l,
                         int *nnz_u, int *nsuper)
{
    return g_fstats(handle, n, perm_c, perm_r, nnz_l, nnz_u, nsuper);
}

int orc_ref_factor_supernodes(long long handle, int *xsup) { return g_fsup(handle, xsup); }

/* ------------------------------------------------------------------------------------------ */

void orc_default_knobs(orc_knobs *k)
{
    /* readsolverinput defaults, solver_knobs.F90:43-57 */
    k->ittd = 6;
    k->dgrd = 1.0e-3;
    k->grdthreshold = 1.0e-15;
    k->irdn = 2;
    k->cyminformal = 1.0e-30;
    k->min_rate_ov_dt = 1.0e-99;
    k->max_sub_steps = 4000;
    k->min_t9_for_slv = 0.01;
    k->delta_state_max = 1.0e99;
    k->interpolate_temp = 1;
    k->interpolate_dens = 1;
    k->integration_method = 0;
}

static void apply_knobs(orc_ctx *c, const orc_knobs *kn)
{
    c->kn = *kn;
    /* static network switch, solver_knobs.F90:66-69 */
    if (c->kn.cyminformal < 1.0e-99) c->kn.min_rate_ov_dt = -1.0e99;
}

/* Fortran real ** integer as compiled by gfortran (__builtin_powi -> libgcc __powidf2) */
static double powi(double x, int m)
{
    unsigned int n = m < 0 ? -(unsigned int) m : (unsigned int) m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

static const double FACT[5] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0}; /* jac_rhs.F90:60-64 */

/* BD tables: bader_deuflhard_init, compute_work_estimates, compute_correction_factors,
 * determine_optimal_convergence (bader_deuflhard.F90:66-124), with use_weights = .true.,
 * cf=1, cj=0, cs=10, clr=cj, safe1=1/4 (lines 47-59). */
static void bd_init(int *nseq, double *a, double alf[NLEVELS][NLEVELS], int *kopt, int *kmax,
                    double dgrd)
{
    const double alpha = 5.0 / 7.0;
    const double cf = 1.0, cj = 0.0, cs = 10.0, clr = 0.0, safe1 = 0.25;
    nseq[0] = 2; nseq[1] = 6; nseq[2] = 10;
    for (int i = 3; i < NLEVELS + 1; ++i) nseq[i] = (int) lround((double) nseq[i - 1] / alpha);
    for (int i = 0; i < NLEVELS + 1; ++i) a[i] = 0.0;   /* a(9) is never assigned */
    a[0] = cj + clr + (nseq[0] + 1) * (cf + cs);
    for (int i = 1; i < NLEVELS; ++i) a[i] = a[i - 1] + nseq[i] * (cf + cs) + clr + cs;
    for (int i = 0; i < NLEVELS; ++i)
        for (int j = 0; j < NLEVELS; ++j) alf[i][j] = 0.0;
    for (int iq = 2; iq <= NLEVELS; ++iq)
        for (int k = 1; k <= iq - 1; ++k)
            alf[k - 1][iq - 1] = pow(dgrd * safe1, (a[k] - a[iq]) /
                                                       ((a[iq] - a[0] + 1.0) * (2.0 * k + 1)));
    int ko;
    *kmax = 0;
    for (ko = 2; ko <= NLEVELS - 1; ++ko) {
        if (a[ko] > a[ko - 1] * alf[ko - 2][ko - 1]) { *kmax = ko; break; }
    }
    *kopt = ko;
}

void orc_bd_tables(int *nseq9, double *a9, double *alf64, int *kopt, int *kmax)
{
    int nseq[NLEVELS + 1];
    double a[NLEVELS + 1], alf[NLEVELS][NLEVELS];
    bd_init(nseq, a, alf, kopt, kmax, 1.0e-3);
    for (int i = 0; i < NLEVELS + 1; ++i) { nseq9[i] = nseq[i]; a9[i] = a[i]; }
    for (int i = 0; i < NLEVELS; ++i)
        for (int j = 0; j < NLEVELS; ++j) alf64[i * NLEVELS + j] = alf[i][j];
}

/* ------------------------------------------------------------------------------------------ */
/* makemap, jac_rhs.F90:461-540: mark the 8 slots of each reaction and the diagonal of each
 * considered isotope, number them row-major, and record each reaction's slot per position.   */

static int cmp_ll(const void *a, const void *b)
{
    long long x = *(const long long *) a, y = *(const long long *) b;
    return (x > y) - (x < y);
}

static int find_slot(const orc_ctx *c, int r, int col)
{
    /* getelm binary search, sparsekit_module.F90:512-624 (sorted columns) */
    int lo = c->rowptr[r], hi = c->rowptr[r + 1] - 1;
    while (lo <= hi) {
        int m = (lo + hi) / 2;
        if (c->colind[m] == col) return m;
        if (c->colind[m] < col) lo = m + 1; else hi = m - 1;
    }
    return -1;
}

static void makemap(orc_ctx *c)
{
    int nsp = c->nsp, nre = c->nre;
    long long *keys = malloc(sizeof(long long) * (8 * (size_t) nre + nsp));
    size_t nk = 0;
    for (int l = 0; l < nre; ++l) {
        int in1 = c->k[0 * nre + l], in2 = c->k[2 * nre + l];
        int out1 = c->k[4 * nre + l], out2 = c->k[6 * nre + l];
        int rr[8] = {in1, in2, out1, out2, in1, in2, out1, out2};
        int cc[8] = {in1, in1, in1, in1, in2, in2, in2, in2};
        for (int m = 0; m < 8; ++m) keys[nk++] = (long long) rr[m] * nsp + cc[m];
    }
    for (int i = 0; i < nsp; ++i)
        if (c->consider[i]) keys[nk++] = (long long) i * nsp + i;
    qsort(keys, nk, sizeof(long long), cmp_ll);
    size_t nu = 0;
    for (size_t i = 0; i < nk; ++i)
        if (nu == 0 || keys[i] != keys[nu - 1]) keys[nu++] = keys[i];
    c->nnz = (int) nu;
    c->rowptr = calloc(nsp + 1, sizeof(int));
    c->colind = malloc(sizeof(int) * nu);
    c->row_of = malloc(sizeof(int) * nu);
    for (size_t i = 0; i < nu; ++i) {
        int r = (int) (keys[i] / nsp);
        c->colind[i] = (int) (keys[i] % nsp);
        c->row_of[i] = r;
        c->rowptr[r + 1]++;
    }
    for (int i = 0; i < nsp; ++i) c->rowptr[i + 1] += c->rowptr[i];
    free(keys);
    /* dns2csr_map(8,nre): in1in1, in2in1, out1in1, out2in1, in1in2, in2in2, out1in2, out2in2
     * (jac_rhs.F90:68-75, 531-538) */
    c->map = malloc(sizeof(int) * 8 * (size_t) nre);
    for (int l = 0; l < nre; ++l) {
        int in1 = c->k[0 * nre + l], in2 = c->k[2 * nre + l];
        int out1 = c->k[4 * nre + l], out2 = c->k[6 * nre + l];
        int rr[8] = {in1, in2, out1, out2, in1, in2, out1, out2};
        int cc[8] = {in1, in1, in1, in1, in2, in2, in2, in2};
        for (int m = 0; m < 8; ++m) c->map[m * nre + l] = find_slot(c, rr[m], cc[m]);
    }
    c->diag = malloc(sizeof(int) * nsp);
    for (int i = 0; i < nsp; ++i) c->diag[i] = find_slot(c, i, i);
}

int orc_nnz_full(const orc_ctx *c) { return c->nnz; }

void orc_full_pattern(const orc_ctx *c, int *rowptr, int *colind, int *map)
{
    memcpy(rowptr, c->rowptr, sizeof(int) * (c->nsp + 1));
    memcpy(colind, c->colind, sizeof(int) * c->nnz);
    memcpy(map, c->map, sizeof(int) * 8 * (size_t) c->nre);
}

orc_ctx *orc_create(int nsp, int nre, const int *k, const int *rtype, const double *coef,
                    const double *aw, const double *zw, int igamma, const orc_knobs *knobs)
{
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->nsp = nsp; c->nre = nre; c->igamma = igamma;
    c->k = malloc(sizeof(int) * 8 * (size_t) nre); memcpy(c->k, k, sizeof(int) * 8 * (size_t) nre);
    c->rtype = malloc(sizeof(int) * nre); memcpy(c->rtype, rtype, sizeof(int) * nre);
    c->coef = malloc(sizeof(double) * 7 * (size_t) nre);
    memcpy(c->coef, coef, sizeof(double) * 7 * (size_t) nre);
    c->aw = malloc(sizeof(double) * nsp); memcpy(c->aw, aw, sizeof(double) * nsp);
    c->zw = malloc(sizeof(double) * nsp); memcpy(c->zw, zw, sizeof(double) * nsp);
    c->consider = malloc(nsp);
    for (int i = 0; i < nsp; ++i) c->consider[i] = (i != igamma);
    orc_knobs d;
    if (!knobs) { orc_default_knobs(&d); knobs = &d; }
    apply_knobs(c, knobs);
    makemap(c);
    c->v = malloc(sizeof(double) * nre);
    c->jac = malloc(sizeof(double) * c->nnz);
    c->jac_lo = malloc(sizeof(double) * c->nnz);
    c->jac_cpy = malloc(sizeof(double) * c->nnz);
    c->jrow = malloc(sizeof(int) * c->nnz);
    c->jcol = malloc(sizeof(int) * c->nnz);
    c->red_val = malloc(sizeof(double) * c->nnz);
    c->red_colind = malloc(sizeof(int) * c->nnz);
    c->red_rowptr = malloc(sizeof(int) * (nsp + 1));
    c->iexpand = malloc(sizeof(int) * nsp);
    c->xn = malloc(sizeof(double) * nsp);
    c->dxdt = malloc(sizeof(double) * nsp);
    c->dxdt_lo = malloc(sizeof(double) * nsp);
    c->rhs = malloc(sizeof(double) * nsp);
    c->red_rhs = malloc(sizeof(double) * nsp);
    c->red_xlo = malloc(sizeof(double) * nsp);
    c->mask = malloc(nsp);
    c->qcol = calloc((size_t) nsp * NLEVELS, sizeof(double));
    c->yerr = calloc((size_t) nsp * NLEVELS, sizeof(double));
    c->lsolver = ORC_LS_GEPP;
    orc_reset_state(c);
    return c;
}

void orc_reset_state(orc_ctx *c)
{
    c->frst = 1;
    c->reduce = 0;
    bd_init(c->nseq, c->a, c->alf, &c->kopt, &c->kmax, c->kn.dgrd);
    c->errmax = c->abuerr = c->negerr = c->sumerr = 0.0;
}

void orc_set_knobs(orc_ctx *c, const orc_knobs *knobs)
{
    apply_knobs(c, knobs);
    bd_init(c->nseq, c->a, c->alf, &c->kopt, &c->kmax, c->kn.dgrd);
}

void orc_set_linear_solver(orc_ctx *c, int which) { c->lsolver = which; }

void orc_set_rates(orc_ctx *c, const double *v) { memcpy(c->v, v, sizeof(double) * c->nre); }

void orc_bd_state(const orc_ctx *c, int *iv, double *dv)
{
    iv[0] = c->frst; iv[1] = c->reduce; iv[2] = c->kopt; iv[3] = c->kmax;
    dv[0] = c->errmax; dv[1] = c->abuerr; dv[2] = c->negerr; dv[3] = c->sumerr;
    for (int k = 0; k < NLEVELS; ++k) dv[4 + k] = c->err[k];
}

void orc_bd_set_frst(orc_ctx *c, int frst) { c->frst = frst; }

void orc_destroy(orc_ctx *c)
{
    if (c) {
        free(c->pf_t9); free(c->pf_tab); free(c->rev_ch); free(c->rev_part); free(c->irev);
        free(c->g_spin); free(c->mass_g); free(c->q_mev);
    }
    if (!c) return;
    free(c->k); free(c->rtype); free(c->coef); free(c->aw); free(c->zw); free(c->consider);
    free(c->rowptr); free(c->colind); free(c->row_of); free(c->map); free(c->diag);
    free(c->v); free(c->jac); free(c->jac_lo); free(c->jac_cpy); free(c->jrow); free(c->jcol);
    free(c->red_val); free(c->red_colind); free(c->red_rowptr); free(c->iexpand);
    free(c->xn); free(c->dxdt); free(c->dxdt_lo); free(c->rhs); free(c->red_rhs); free(c->red_xlo); free(c->mask);
    free(c->qcol); free(c->yerr); free(c->rcache);
    free(c);
}

/* ------------------------------------------------------------------------------------------ */
/* rates: reaclib_calculate_rates (reaclib.F90:431-582), direct-evaluation branch, one REACLIB
 * set per reaction, Gfac = 1; then evaluate_all_rates' clips (evaluate_rates.F90:46,172).
 * With the T9 cache on (physics knob use_cache, reaclib.F90:446-470) the rates are interpolated
 * from the table of orc_set_rate_cache instead.                                              */
static void rates_direct(const orc_ctx *c, double t9in, double rho, double *v);

void orc_rates(const orc_ctx *c, double t9in, double rho, double ye, double *v)
{
    (void) ye; /* rho*ye scaling applies to '  ec' rates only; the network has none */
    if (c->use_cache && c->rcache) {
        /* reaclib.F90:443-470: i1 = floor((t9 - 0.01) / dt9) + 1 clamped to [1, n - 1]
         * (0-based here), linear interpolation, x rho for the two-body rates, the floor; then
         * evaluate_all_rates' clips */
        const int nre = c->nre;
        const double t9 = fmax(t9in, 0.01);
        int i1 = (int) floor((t9 - 0.01) / c->rc_dt9);
        i1 = i1 < 0 ? 0 : (i1 > ORC_NT9 - 2 ? ORC_NT9 - 2 : i1);
        const double dt9 = (t9 - c->rc_t9[i1]) / c->rc_dt9;
        const double *c1 = c->rcache + (size_t) i1 * nre, *c2 = c1 + nre;
        for (int l = 0; l < nre; ++l) {
            double rv = c1[l] + dt9 * (c2[l] - c1[l]);
            if (c->rtype[l] == 0) rv = rv * rho;
            rv = fmax(rv, 1.0e-99);
            rv = fmax(rv * 1.0, 1.0e-99);
            v[l] = fmax(rv, 1.0e-99);
        }
        return;
    }
    rates_direct(c, t9in, rho, v);
}

/* reaclib_preprocessor (reaclib.F90:609-646): the rates at rho = 1 on the 5,001-point grid
 * t9_i = 0.01 + i dt9, dt9 = (10 - 0.01) / 5000, each floored at 1e-99 */
void orc_set_rate_cache(orc_ctx *c, int use_cache)
{
    c->use_cache = use_cache;
    if (!use_cache || c->rcache) return;
    const int nre = c->nre;
    c->rcache = malloc(sizeof(double) * (size_t) nre * ORC_NT9);
    c->rc_dt9 = (10.0 - 0.01) / (ORC_NT9 - 1);
    for (int i = 0; i < ORC_NT9; ++i) {
        c->rc_t9[i] = 0.01 + i * c->rc_dt9;
        rates_direct(c, c->rc_t9[i], 1.0, c->rcache + (size_t) i * nre);
    }
}

const double *orc_rate_cache(const orc_ctx *c) { return c->rcache; }

/* reaclib_interpolate_partition_functions (reaclib.F90:586-605) with bsearch_r8
 * (utils.F90:687-722): the temperature clipped to the grid, the interval by bisection, linear
 * interpolation of every species' table */
static void partition_functions(const orc_ctx *c, double t9_in, double *pi)
{
    const int n = c->pf_n;
    const double *rt9 = c->pf_t9;
    const double t9 = fmax(rt9[0], fmin(t9_in, rt9[n - 1]));
    /* bsearch_r8, 1-based a, b, m */
    int a = 1, b = n, r = b - a, m = (a + b) / 2;
    while (t9 != rt9[m - 1] && r > 1) {
        if (rt9[m - 1] > t9) b = m;
        else a = m;
        r = b - a;
        m = (a + b) / 2;
    }
    if (rt9[m - 1] > t9) m = m - 1;
    const int i0 = m - 1, i1 = m;   /* 0-based i0 = m - 1, i1 = i0 + 1 */
    const double t90 = rt9[i0], t91 = rt9[i1];
    const double dt9 = (t9 - t90) / (t91 - t90);
    for (int s = 0; s < c->nsp; ++s) {
        const double *p = c->pf_tab + (size_t) s * n;
        pi[s] = p[i0] + (p[i1] - p[i0]) * dt9;
    }
}

/* Gfac of a REACLIB reverse ('v') rate by chapter, reaclib.F90:512-557 */
static double reverse_gfac(int chapter, const double *g)
{
    switch (chapter) {
    case 1: return g[1] / g[0];
    case 2: return g[1] * g[2] / g[0];
    case 3: return g[1] * g[2] * g[3] / g[0];
    case 4: return g[2] / (g[0] * g[1]);
    case 5: return g[2] * g[3] / (g[0] * g[1]);
    case 6: return g[2] * g[3] * g[4] / (g[0] * g[1]);
    case 7: return g[2] * g[3] * g[4] * g[5] / (g[0] * g[1]);
    case 8: return g[3] / (g[0] * g[1] * g[2]);
    case 9: return g[3] * g[4] / (g[0] * g[1] * g[2]);
    case 10: return g[4] * g[5] / (g[0] * g[1] * g[2] * g[3]);
    case 11: return g[1] * g[2] * g[3] * g[4] * g[5] / g[0];
    default: return 1.0;
    }
}

static void rates_direct(const orc_ctx *c, double t9in, double rho, double *v)
{
    double *pinter = NULL;
    if (c->pf_n > 0) {
        pinter = malloc(sizeof(double) * c->nsp);
        partition_functions(c, t9in, pinter);   /* reaclib_wrapper, reaclib.F90:177-190 */
    }
    const double t9 = fmax(t9in, 0.01);          /* t9_lowlimit_reaclib, reaclib.F90:443 */
    const double t9inv = 1.0 / t9;
    const double t9pow13 = pow(t9, 1.0 / 3.0);
    const double t9pow13inv = 1.0 / t9pow13;
    const double t9pow43 = t9 * t9pow13;
    const double t9pow53 = t9pow13 * t9pow43;
    const double lnt9 = log(t9);
    const int nre = c->nre;
    const double *a = c->coef;
    for (int l = 0; l < nre; ++l) {
        double r0 = a[0 * nre + l] + a[1 * nre + l] * t9inv;
        r0 = r0 + a[2 * nre + l] * t9pow13inv;
        r0 = r0 + a[3 * nre + l] * t9pow13;
        r0 = r0 + a[4 * nre + l] * t9;
        r0 = r0 + a[5 * nre + l] * t9pow53;
        r0 = r0 + a[6 * nre + l] * lnt9;
        double reacv = 0.0 + exp(r0);
        if (pinter && c->rev_ch[l] > 0) {
            double g[6];
            for (int e = 0; e < 6; ++e) {
                const int sp = c->rev_part[e * nre + l];
                g[e] = sp >= 0 ? pinter[sp] : 1.0;
            }
            reacv = reacv * reverse_gfac(c->rev_ch[l], g);
        }
        double rv = (c->rtype[l] == 0) ? reacv * rho : reacv;
        rv = fmax(rv, 1.0e-99);                     /* reaclib.F90:581 */
        rv = fmax(rv * 1.0, 1.0e-99);               /* rfac, evaluate_rates.F90:46 */
        v[l] = fmax(rv, 1.0e-99);                   /* evaluate_rates.F90:172 */
    }
    free(pinter);
}

/* physical constants, utils/source/constants.F90:11-45 (compile-time values) */
#define C_EV2ERG 1.602177e-12
#define C_MEV2ERG (C_EV2ERG * 1.0e6)
#define C_ERG2MEV (1.0 / C_MEV2ERG)
#define C_PI 3.141592653589793
#define C_AVOGADRO 6.022140857e23
#define C_BOLTZMANN 1.3806485279e-16
#define C_PLANCK_MEV 4.135667662e-21
#define C_PLANCK_ERG (C_PLANCK_MEV * C_MEV2ERG)
#define C_HBAR_ERG (C_PLANCK_ERG / 2.0 / C_PI)
#define C_AMU 1.6605402e-24
#define C_QE 4.80320425e-10
#define C_QE2 (C_QE * C_QE)

/* calculate_reverse_rates + reverse_rate (reverse.F90:107-337, use_partition_functions true,
 * no Coulomb corrections): every forward reaction i paired by reverse_init whose rate exceeds
 * 1e-97 sets its reverse partner's rate v(irev(i)) = max(v(i) fac, 1e-99), with G the
 * partition functions times the spin factors (nse_swj.F90:613-625) */
static void detailed_balance(const orc_ctx *c, double rho, double t9, double *v)
{
    const int nre = c->nre;
    double *G = malloc(sizeof(double) * c->nsp);
    partition_functions(c, t9, G);
    for (int s = 0; s < c->nsp; ++s) G[s] = G[s] * c->g_spin[s];
    const double kTerg = C_BOLTZMANN * t9 * 1.0e9;
    const double kTMeV = kTerg * C_ERG2MEV;
    for (int i = 0; i < nre; ++i) {
        if (c->irev[i] < 0 || !(v[i] > 1.0e-97)) continue;
        const int *k = c->k;
        const int a1 = k[0 * nre + i], na1 = k[1 * nre + i], a2 = k[2 * nre + i];
        const int b1 = k[4 * nre + i], nb1 = k[5 * nre + i], b2 = k[6 * nre + i];
        const int nb2 = k[7 * nre + i];
        double G_a1 = G[a1], m_a1 = c->mass_g[a1], G_a2, m_a2, G_b1, m_b1, G_b2, m_b2;
        if (na1 == 1) { G_a2 = G[a2]; m_a2 = c->mass_g[a2]; }
        else if (na1 == 2) { G_a2 = G_a1; m_a2 = m_a1; }
        else continue;
        G_b2 = G[b2];
        m_b2 = c->mass_g[b2];
        if (nb2 == 1) {
            if (nb1 == 1) { G_b1 = G[b1]; m_b1 = c->mass_g[b1]; }
            else { G_b1 = 0.0; m_b1 = 0.0; }
        } else if (nb2 == 2) {
            G_b1 = G_b2; m_b1 = m_b2;
        } else {
            continue;
        }
        double pfac, mfac, Q, rhofac;
        if (b1 == c->igamma) {   /* (n,g), (p,g), (a,g): one product */
            pfac = G_a1 * G_a2 / G_b2;
            mfac = pow(m_a1 * m_a2 / m_b2 * kTerg * 2.0 * C_PI / C_PLANCK_ERG / C_PLANCK_ERG, 1.5);
            Q = -c->q_mev[a1] - c->q_mev[a2] + c->q_mev[b2];
            rhofac = 1.0 / rho / C_AVOGADRO;
        } else {
            pfac = G_a1 * G_a2 / G_b1 / G_b2;
            mfac = pow(m_a1 * m_a2 / m_b1 / m_b2, 1.5);
            Q = -c->q_mev[a1] - c->q_mev[a2] + c->q_mev[b1] + c->q_mev[b2];
            rhofac = 1.0;
        }
        const double x = fmin(-Q / kTMeV, 200.0);
        const double efac = exp(x);
        double fac = mfac * efac * rhofac;
        fac = fac * pfac;
        fac = fmin(fmax(fac, 1.0e-99), 1.0e99);
        v[c->irev[i]] = fmax(v[i] * fac, 1.0e-99);
    }
    free(G);
}

/* Chugunov, DeWitt & Yakovlev screening (screening_chugunov.F90): the plasma quantities of the
 * zone (screen_precalc_chugunov, 86-152, with calculate_azbar utils.F90:504-525), then every
 * reaction's factor (calculate_screening_chugunov, 159-235) from the charges of its reactants,
 * as evaluate_all_rates loops (evaluate_rates.F90:65-121; the networks here have none of the
 * reference's fixed-index special cases 43 and 109) */
static void screening(const orc_ctx *c, double t9, double rho, const double *x, double *v)
{
    const double third = 1.0 / 3.0;
    const double fourov3pi = 4.0 / (3.0 * C_PI * C_PI);
    const double fourpiqe2 = 4.0 * C_PI * C_QE2;
    const double tkin = t9 * 1.0e9;
    double ytot = 0.0, sa = 0.0, sz = 0.0;
    for (int s = 0; s < c->nsp; ++s) {
        if (!c->consider[s]) continue;
        const double y = x[s] / c->aw[s];
        ytot = ytot + y;
    }
    for (int s = 0; s < c->nsp; ++s) {
        if (!c->consider[s]) continue;
        const double y = x[s] / c->aw[s];
        sa = sa + c->aw[s] * y;
    }
    for (int s = 0; s < c->nsp; ++s) {
        if (!c->consider[s]) continue;
        const double y = x[s] / c->aw[s];
        sz = sz + c->zw[s] * y;
    }
    const double abar = sa / ytot, zbar = sz / ytot;
    const double ntot = rho / (abar * C_AMU);
    const double denom = pow(3.0 / (4.0 * C_PI * zbar * ntot), third);
    const double a_av = denom * pow(zbar, third);
    const double wp = pow(fourpiqe2 * zbar * zbar * ntot / (abar * C_AMU), 0.5);
    const double tp = C_HBAR_ERG * wp / C_BOLTZMANN;
    double tn = tkin / tp, tk;
    const double t1s = 0.1, t2s = 0.2, t0s = (t1s + t2s) * 0.5;
    if (tn >= 2.0 * t2s) {
        tk = tkin;
    } else if (tn <= 0.5 * t1s) {
        tk = t1s * tp;
    } else {
        const double sg = 1.0 / (1.0 + exp(-(tn - t0s) / 0.01));
        tn = (1.0 - sg) * t1s + sg * tn;
        tk = tn * tp;
    }
    const double gamfac = C_QE2 / (a_av * tk) / C_BOLTZMANN;
    const double zeta = pow(fourov3pi * tp * tp / tk / tk, third);
    const double c_a1 = 2.7822, c_a2 = 98.34, c_a3 = sqrt(3.0) - c_a1 / sqrt(c_a2);
    const double c_b1 = -1.7476, c_b2 = 66.07, c_b3 = 1.12, c_b4 = 65.0, alfa = 0.022;
    const int nre = c->nre;
    for (int i = 0; i < nre; ++i) {
        if (v[i] <= 1.0e-99) continue;
        const int r1 = c->k[i], r2 = c->k[1 * nre + i] == 2 ? c->k[i] : c->k[2 * nre + i];
        const double z1 = c->zw[r1], z2 = c->zw[r2];
        double screen = 1.0;
        if (z1 > 0.0 && z2 > 0.0) {
            const double z1z2 = z1 * z2;
            const double gam = z1z2 * gamfac;
            const double beta = 0.41 - 0.6 / gam;
            const double gama = 0.06 + 2.2 / gam;
            double gamtild = gam / pow(1.0 + alfa * zeta + beta * (zeta * zeta) +
                                       gama * (zeta * zeta * zeta), third);
            if (gamtild >= 650.0) {
                gamtild = 600.0;
            } else if (gamtild <= 550.0) {
            } else {
                const double sg = 1.0 / (1.0 + exp(-(gamtild - 599.0) / 1.0));
                gamtild = sg * 600.0 + (1.0 - sg) * gamtild;
            }
            const double gamtild2 = gamtild * gamtild;
            const double t1 = pow(gamtild, 3.0 * 0.5) *
                              (c_a1 / sqrt(c_a2 + gamtild) + c_a3 / (1.0 + gamtild));
            const double t2 = c_b1 * gamtild2 / (c_b2 + gamtild);
            const double t3 = c_b3 * gamtild2 / (c_b4 + gamtild2);
            const double h0fit = t1 + t2 + t3;
            screen = fmin(exp(h0fit), 1.0e30);
        }
        screen = fmin(screen, 1.0e30);
        v[i] = v[i] * screen;
    }
}

/* evaluate_all_rates (evaluate_rates.F90:14-172) at the abundances x: the REACLIB rates (with
 * the partition-function factors of its reverse rates), detailed balance, screening, the clip.
 * Without rate physics this is orc_rates. */
void orc_rates_full(const orc_ctx *c, double t9, double rho, double ye, const double *x,
                    double *v)
{
    orc_rates(c, t9, rho, ye, v);
    if (c->pf_n > 0 && c->detailed_balance && c->irev) detailed_balance(c, rho, t9, v);
    if (c->screen && x) screening(c, t9, rho, x, v);
    for (int l = 0; l < c->nre; ++l) v[l] = fmax(v[l], 1.0e-99);
}

/* reverse_init (reverse.F90:25-105) on this network: forward reactions (not REACLIB reverse
 * rates, not weak, at most one heavy product of a kind) paired with the first reaction of the
 * mirrored topology, kept when that one is a REACLIB reverse rate (all species in their ground
 * state here) */
static void reverse_pairs(orc_ctx *c)
{
    const int nre = c->nre, *k = c->k;
    for (int i = 0; i < nre; ++i) c->irev[i] = -1;
    for (int i = 0; i < nre; ++i) {
        if (c->rtype[i] == 2) continue;              /* weak */
        if (k[7 * nre + i] > 1) continue;
        if (c->rev_ch[i] > 0) continue;
        for (int j = 0; j < nre; ++j) {
            const int ri[8] = {k[i], k[nre + i], k[2 * nre + i], k[3 * nre + i], k[4 * nre + i],
                               k[5 * nre + i], k[6 * nre + i], k[7 * nre + i]};
            const int rj[8] = {k[6 * nre + j], k[7 * nre + j], k[4 * nre + j], k[5 * nre + j],
                               k[2 * nre + j], k[3 * nre + j], k[j], k[nre + j]};
            int same = 1;
            for (int e = 0; e < 8; ++e) same = same && ri[e] == rj[e];
            if (same) {
                if (c->rev_ch[j] > 0) c->irev[i] = j;
                break;
            }
        }
    }
}

void orc_set_rate_physics(orc_ctx *c, int pf_n, const double *pf_t9, const double *pf_tab,
                          const int *rev_ch, const int *rev_part, int screen,
                          int detailed_balance, const double *g_spin, const double *mass_g,
                          const double *q_mev)
{
    const int nsp = c->nsp, nre = c->nre;
    free(c->pf_t9); free(c->pf_tab); free(c->rev_ch); free(c->rev_part); free(c->irev);
    free(c->g_spin); free(c->mass_g); free(c->q_mev);
    c->pf_t9 = c->pf_tab = c->g_spin = c->mass_g = c->q_mev = NULL;
    c->rev_ch = c->rev_part = c->irev = NULL;
    c->pf_n = pf_n;
    c->screen = screen;
    c->detailed_balance = detailed_balance;
    if (pf_n > 0) {
        c->pf_t9 = malloc(sizeof(double) * pf_n);
        memcpy(c->pf_t9, pf_t9, sizeof(double) * pf_n);
        c->pf_tab = malloc(sizeof(double) * (size_t) nsp * pf_n);
        memcpy(c->pf_tab, pf_tab, sizeof(double) * (size_t) nsp * pf_n);
        c->rev_ch = malloc(sizeof(int) * nre);
        memcpy(c->rev_ch, rev_ch, sizeof(int) * nre);
        c->rev_part = malloc(sizeof(int) * 6 * (size_t) nre);
        memcpy(c->rev_part, rev_part, sizeof(int) * 6 * (size_t) nre);
        c->g_spin = malloc(sizeof(double) * nsp);
        memcpy(c->g_spin, g_spin, sizeof(double) * nsp);
        c->mass_g = malloc(sizeof(double) * nsp);
        memcpy(c->mass_g, mass_g, sizeof(double) * nsp);
        c->q_mev = malloc(sizeof(double) * nsp);
        memcpy(c->q_mev, q_mev, sizeof(double) * nsp);
        c->irev = malloc(sizeof(int) * nre);
        reverse_pairs(c);
    }
    /* a cache built before holds the rates without the partition-function factors */
    if (c->rcache) {
        free(c->rcache);
        c->rcache = NULL;
        if (c->use_cache) orc_set_rate_cache(c, 1);
    }
}

void orc_reverse_pairs(const orc_ctx *c, int *irev)
{
    for (int i = 0; i < c->nre; ++i) irev[i] = c->irev ? c->irev[i] : -1;
}

/* s += t as an unevaluated double-double sum (hi, lo): TwoSum (Knuth), the error into lo */
static void dd_add(double *hi, double *lo, double t)
{
    const double s = *hi + t, bb = s - *hi;
    *lo += (*hi - (s - bb)) + (t - bb);
    *hi = s;
}

/* calculate_dxdt, jac_rhs.F90:217-278.  Under ORC_LS_EXACT each species' dX/dt is the value
 * of sum_l (count A) vx_l rounded once: the products split exactly (FMA) and summed in
 * double-double, dxdt = RN(S), dxdt_lo = S - RN(S) to double-double accuracy (a yardstick for
 * the rounding of the reference's sequential FP64 sum). */
void orc_dxdt(const orc_ctx *c, const double *x, const double *v, double *dxdt)
{
    const int nsp = c->nsp, nre = c->nre, g = c->igamma;
    const int exact = c->lsolver == ORC_LS_EXACT;
    double *xn = c->xn, *lo = c->dxdt_lo;
    const double *an = c->aw;
    for (int i = 0; i < nsp; ++i) {
        dxdt[i] = 0.0;
        lo[i] = 0.0;
        xn[i] = (c->consider[i] && c->aw[i] != 0.0) ? x[i] / c->aw[i] : 0.0;
    }
    xn[g] = 1.0;
    for (int l = 0; l < nre; ++l) {
        int k1 = c->k[l], k2 = c->k[nre + l], k3 = c->k[2 * nre + l];
        int k5 = c->k[4 * nre + l], k6 = c->k[5 * nre + l], k7 = c->k[6 * nre + l];
        int k8 = c->k[7 * nre + l];
        double vx = FACT[k2] * v[l] * powi(xn[k1], k2) * xn[k3];
        if (exact) {
            const int sp[4] = {k1, k3, k5, k7};
            const double cnt[4] = {-k2, -1.0, k6, k8};
            for (int r = 0; r < 4; ++r) {
                const double ca = cnt[r] * an[sp[r]], t = ca * vx;
                dd_add(&dxdt[sp[r]], &lo[sp[r]], t);
                lo[sp[r]] += fma(ca, vx, -t);
            }
            continue;
        }
        double dx1 = -k2 * vx, dx3 = -vx, dx5 = k6 * vx, dx7 = k8 * vx;
        dxdt[k1] = dxdt[k1] + dx1;
        dxdt[k3] = dxdt[k3] + dx3;
        dxdt[k5] = dxdt[k5] + dx5;
        dxdt[k7] = dxdt[k7] + dx7;
    }
    for (int i = 0; i < nsp; ++i) {
        if (!c->consider[i]) continue;
        if (exact) {
            const double hi = dxdt[i], r = hi + lo[i];
            lo[i] = lo[i] - (r - hi);
            dxdt[i] = r;
        } else {
            dxdt[i] = dxdt[i] * c->aw[i];
        }
    }
    dxdt[g] = 0.0;
    lo[g] = 0.0;
}

/* One Jacobian contribution cnt * stem * a_row / a_col as a real number, added to the
 * double-double (hi, lo): the ratio cnt a_row / a_col as a double-double (its remainder is
 * exact with an FMA), the product with stem split exactly.  ORC_LS_EXACT assembles J from
 * these and rounds each entry once. */
static void jterm_dd(double *hi, double *lo, double cnt, double stem, double ar, double ac)
{
    const double num = cnt * ar;   /* an exact integer */
    const double ch = num / ac, cl = fma(-ch, ac, num) / ac;
    const double t = stem * ch;
    dd_add(hi, lo, t);
    *lo += fma(stem, ch, -t) + stem * cl;
}

/* calculate_jacobian, jac_rhs.F90:282-455 (irdn = 2).  Under ORC_LS_EXACT every entry is the
 * once-rounded sum of its contributions as real values (jterm_dd). */
int orc_jacobian(const orc_ctx *c, double dzeit, const double *x, const double *v,
                 double *jac, int *row, int *col, unsigned char *mask)
{
    const int nsp = c->nsp, nre = c->nre, g = c->igamma, nnz = c->nnz;
    const double *an = c->aw;
    double *xn = c->xn;
    for (int i = 0; i < nsp; ++i) mask[i] = 0;
    const double dfdNmin = c->kn.cyminformal / dzeit;
    const int exact = c->lsolver == ORC_LS_EXACT;
    double *jlo = c->jac_lo;
    for (int i = 0; i < nnz; ++i) { jac[i] = 1.0e-99; jlo[i] = 0.0; row[i] = -1; col[i] = -1; }
    for (int i = 0; i < nsp; ++i)
        if (c->consider[i] && an[i] > 0.0) xn[i] = x[i] / an[i];
        else xn[i] = 0.0;   /* untouched in the reference; only gamma (reset below) hits this */
    xn[g] = 1.0;
    for (int l = 0; l < nre; ++l) {
        if (fabs(v[l]) / dzeit <= c->kn.min_rate_ov_dt) continue;
        const int in1 = c->k[l], Nin1 = c->k[nre + l], in2 = c->k[2 * nre + l];
        const int Nin2 = c->k[3 * nre + l], out1 = c->k[4 * nre + l];
        const int Nout1 = c->k[5 * nre + l], out2 = c->k[6 * nre + l], Nout2 = c->k[7 * nre + l];
        const double ain1 = an[in1], ain2 = an[in2], aout1 = an[out1], aout2 = an[out2];
        const double y1 = xn[in1], y2 = xn[in2];
        const int *m = c->map;
        const int ice1 = m[0 * nre + l], ice2 = m[1 * nre + l], ice3 = m[4 * nre + l];
        const int ice4 = m[2 * nre + l], ice5 = m[3 * nre + l], ice6 = m[5 * nre + l];
        const int ice7 = m[6 * nre + l], ice8 = m[7 * nre + l];

        double stem = FACT[Nin1] * FACT[Nin2] * v[l];
        double dfdN = stem * powi(y1, Nin1) * y2;
        if (fabs(dfdN) > dfdNmin) { mask[in1] = 1; mask[in2] = 1; mask[out1] = 1; mask[out2] = 1; }

        stem = (double) Nin1 * FACT[Nin1] * FACT[Nin2] * powi(y2, Nin2) * powi(y1, Nin1 - 1) * v[l];
        const double dfin1dy1 = -Nin1 * stem, dfin2dy1 = -Nin2 * stem;
        const double dfout1dy1 = Nout1 * stem, dfout2dy1 = Nout2 * stem;
        stem = (double) Nin2 * FACT[Nin1] * FACT[Nin2] * powi(y1, Nin1) * powi(y2, Nin2 - 1) * v[l];
        const double dfin1dy2 = -Nin1 * stem, dfin2dy2 = -Nin2 * stem;
        const double dfout1dy2 = Nout1 * stem, dfout2dy2 = Nout2 * stem;

        if (exact) {
            const double s1 = (double) Nin1 * FACT[Nin1] * FACT[Nin2] * powi(y2, Nin2) *
                              powi(y1, Nin1 - 1) * v[l];
            const double s2 = (double) Nin2 * FACT[Nin1] * FACT[Nin2] * powi(y1, Nin1) *
                              powi(y2, Nin2 - 1) * v[l];
            if (in1 != g) jterm_dd(&jac[ice1], &jlo[ice1], -Nin1, s1, ain1, ain1);
            if (in1 != g && in2 != g) {
                jterm_dd(&jac[ice2], &jlo[ice2], -Nin2, s1, ain2, ain1);
                jterm_dd(&jac[ice3], &jlo[ice3], -Nin1, s2, ain1, ain2);
            }
            if (in1 != g && out1 != g) jterm_dd(&jac[ice4], &jlo[ice4], Nout1, s1, aout1, ain1);
            if (in1 != g && out2 != g) jterm_dd(&jac[ice5], &jlo[ice5], Nout2, s1, aout2, ain1);
            if (in2 != g) jterm_dd(&jac[ice6], &jlo[ice6], -Nin2, s2, ain2, ain2);
            if (in2 != g && out1 != g) jterm_dd(&jac[ice7], &jlo[ice7], Nout1, s2, aout1, ain2);
            if (in2 != g && out2 != g) jterm_dd(&jac[ice8], &jlo[ice8], Nout2, s2, aout2, ain2);
        } else {
            if (in1 != g) jac[ice1] = jac[ice1] + dfin1dy1 * ain1 / ain1;
            if (in1 != g && in2 != g) {
                jac[ice2] = jac[ice2] + dfin2dy1 * ain2 / ain1;
                jac[ice3] = jac[ice3] + dfin1dy2 * ain1 / ain2;
            }
            if (in1 != g && out1 != g) jac[ice4] = jac[ice4] + dfout1dy1 * aout1 / ain1;
            if (in1 != g && out2 != g) jac[ice5] = jac[ice5] + dfout2dy1 * aout2 / ain1;
            if (in2 != g) jac[ice6] = jac[ice6] + dfin2dy2 * ain2 / ain2;
            if (in2 != g && out1 != g) jac[ice7] = jac[ice7] + dfout1dy2 * aout1 / ain2;
            if (in2 != g && out2 != g) jac[ice8] = jac[ice8] + dfout2dy2 * aout2 / ain2;
        }

        if (in2 != g && !mask[in2]) continue;
        if (out1 != g && !mask[out1]) continue;
        if (mask[in1] && mask[out2]) {
            if (in1 != g) { row[ice1] = in1; col[ice1] = in1; }
            if (in1 != g && in2 != g) {
                row[ice2] = in2; col[ice2] = in1;
                row[ice3] = in1; col[ice3] = in2;
            }
            if (in1 != g && out1 != g) { row[ice4] = out1; col[ice4] = in1; }
            if (in1 != g && out2 != g) { row[ice5] = out2; col[ice5] = in1; }
            if (in2 != g) { row[ice6] = in2; col[ice6] = in2; }
            if (in2 != g && out1 != g) { row[ice7] = out1; col[ice7] = in2; }
            if (in2 != g && out2 != g) { row[ice8] = out2; col[ice8] = in2; }
        }
    }
    mask[g] = 0;
    if (exact)
        for (int l = 0; l < nnz; ++l) jac[l] = jac[l] + jlo[l];
    for (int l = 0; l < nnz; ++l)
        if (jac[l] == 0.0 && row[l] != col[l]) return 1;
    return 0;
}

/* reduce_jacobian, jac_rhs.F90:134-197 */
int orc_reduce_jacobian(const orc_ctx *c, double *jac, int *row, const int *colv,
                        const unsigned char *mask, int *nvar1_out, int *rowptr, int *colind,
                        double *val)
{
    const int nsp = c->nsp, nnz = c->nnz;
    int *iexpand = c->iexpand;
    for (int i = 0; i < nnz; ++i)
        if ((fabs(jac[i]) == 0.0 || jac[i] == -1.0) && row[i] != -1) row[i] = -1;
    int nvar1 = 0;
    for (int i = 0; i < nsp; ++i) {
        iexpand[i] = -1;
        if (!mask[i]) continue;
        iexpand[i] = nvar1++;
        int ice = c->diag[i];
        row[ice] = i;
        jac[ice] = jac[ice] + 1.0;
    }
    /* pack in COO (= full row-major) order, renumber, CSR (coocsr keeps the order) */
    for (int i = 0; i <= nvar1; ++i) rowptr[i] = 0;
    int n_ice = 0;
    for (int i = 0; i < nnz; ++i) {
        if (row[i] == -1) continue;
        /* a set slot always carries its pattern position (makemap numbering) */
        int r = iexpand[c->row_of[i]], cl = iexpand[c->colind[i]];
        if (jac[i] == 0.0) return -1;
        rowptr[r + 1]++;
        colind[n_ice] = cl;
        val[n_ice] = jac[i];
        ++n_ice;
    }
    for (int i = 0; i < nvar1; ++i) rowptr[i + 1] += rowptr[i];
    *nvar1_out = nvar1;
    return n_ice;
}

/* ------------------------------------------------------------------------------------------ */
/* GEPP: left-looking column LU with SuperLU's pivot policy (dpivotL.c:114-184): the largest
 * |value| among the not-yet-pivoted rows of the current column (first occurrence wins), unless
 * the diagonal (original row == original column) is nonzero and >= threshold(1.0) * max.
 * L columns are scaled by the reciprocal of the pivot (dpivotL.c:178-180).                    */

struct orc_lu {
    int n;
    int *perm_c, *iperm_c, *perm_r, *pivrow;
    int *Lp, *Li, *Up, *Ui;
    double *Lx, *Ux;
    int nnz_l, nnz_u;
};

int orc_lu_factor(int n, const int *colptr, const int *rowind, const double *val,
                  const int *perm_c_in, orc_lu **out)
{
    orc_lu *lu = calloc(1, sizeof(orc_lu));
    lu->n = n;
    lu->perm_c = malloc(sizeof(int) * n); lu->iperm_c = malloc(sizeof(int) * n);
    lu->perm_r = malloc(sizeof(int) * n); lu->pivrow = malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) lu->perm_c[i] = perm_c_in ? perm_c_in[i] : i;
    for (int i = 0; i < n; ++i) lu->iperm_c[lu->perm_c[i]] = i;
    for (int i = 0; i < n; ++i) lu->perm_r[i] = -1;
    size_t capl = (size_t) colptr[n] * 4 + 16, capu = capl;
    lu->Lp = malloc(sizeof(int) * (n + 1)); lu->Up = malloc(sizeof(int) * (n + 1));
    lu->Li = malloc(sizeof(int) * capl); lu->Lx = malloc(sizeof(double) * capl);
    lu->Ui = malloc(sizeof(int) * capu); lu->Ux = malloc(sizeof(double) * capu);
    double *x = calloc(n, sizeof(double));
    unsigned char *flag = calloc(n, 1);
    int *touched = malloc(sizeof(int) * n);
    int info = 0;
    lu->Lp[0] = lu->Up[0] = 0;
    for (int j = 0; j < n; ++j) {
        int c = lu->iperm_c[j], nt = 0;
        for (int p = colptr[c]; p < colptr[c + 1]; ++p) {
            int r = rowind[p];
            x[r] = val[p];
            if (!flag[r]) { flag[r] = 1; touched[nt++] = r; }
        }
        size_t nu = lu->Up[j];
        for (int k = 0; k < j; ++k) {
            int pr = lu->pivrow[k];
            if (!flag[pr]) continue;
            double xk = x[pr];
            if (nu + 1 >= capu) {
                capu *= 2;
                lu->Ui = realloc(lu->Ui, sizeof(int) * capu);
                lu->Ux = realloc(lu->Ux, sizeof(double) * capu);
            }
            lu->Ui[nu] = k; lu->Ux[nu] = xk; ++nu;
            for (int p = lu->Lp[k]; p < lu->Lp[k + 1]; ++p) {
                int i = lu->Li[p];
                if (!flag[i]) { flag[i] = 1; x[i] = 0.0; touched[nt++] = i; }
                x[i] -= lu->Lx[p] * xk;
            }
        }
        /* pivot search over unpivoted touched rows, in touched order */
        double pivmax = 0.0;
        int pivptr = -1, diag = -1, first = -1;
        for (int t = 0; t < nt; ++t) {
            int r = touched[t];
            if (lu->perm_r[r] >= 0) continue;
            if (first < 0) first = t;
            double rt = fabs(x[r]);
            if (rt > pivmax) { pivmax = rt; pivptr = t; }
            if (r == c) diag = t;
        }
        int piv;
        if (first < 0) { info = j + 1; piv = -1; }
        else if (pivmax == 0.0) { info = j + 1; piv = touched[first]; }
        else {
            double thresh = 1.0 * pivmax;
            if (diag >= 0) {
                double rt = fabs(x[touched[diag]]);
                if (rt != 0.0 && rt >= thresh) pivptr = diag;
            }
            piv = touched[pivptr];
        }
        if (info) {  /* no nonzero pivot: the factorisation stops (the caller stops too) */
            for (int k2 = j; k2 < n; ++k2) { lu->Up[k2 + 1] = (int) nu; lu->Lp[k2 + 1] = (int) lu->Lp[j]; }
            break;
        }
        lu->perm_r[piv] = j;
        lu->pivrow[j] = piv;
        double pv = x[piv];
        lu->Ui[nu] = j; lu->Ux[nu] = pv; ++nu;
        lu->Up[j + 1] = (int) nu;
        double temp = 1.0 / pv;
        size_t nl = lu->Lp[j];
        for (int t = 0; t < nt; ++t) {
            int r = touched[t];
            if (lu->perm_r[r] >= 0) continue;
            if (nl + 1 >= capl) {
                capl *= 2;
                lu->Li = realloc(lu->Li, sizeof(int) * capl);
                lu->Lx = realloc(lu->Lx, sizeof(double) * capl);
            }
            lu->Li[nl] = r; lu->Lx[nl] = x[r] * temp; ++nl;
        }
        lu->Lp[j + 1] = (int) nl;
        for (int t = 0; t < nt; ++t) { flag[touched[t]] = 0; x[touched[t]] = 0.0; }
    }
    lu->nnz_l = lu->Lp[n];
    lu->nnz_u = lu->Up[n];
    free(x); free(flag); free(touched);
    *out = lu;
    return info;
}

/* dgstrs (NOTRANS), SuperLU_5.0/SRC/dgstrs.c:149-317: b <- Pr b, L solve, U solve, x <- Pc x */
void orc_lu_solve(const orc_lu *lu, double *b)
{
    int n = lu->n;
    double *y = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) y[lu->perm_r[i]] = b[i];
    for (int j = 0; j < n; ++j) {
        double yj = y[j];
        for (int p = lu->Lp[j]; p < lu->Lp[j + 1]; ++p) y[lu->perm_r[lu->Li[p]]] -= lu->Lx[p] * yj;
    }
    for (int j = n - 1; j >= 0; --j) {
        int pd = lu->Up[j + 1] - 1;
        y[j] = y[j] / lu->Ux[pd];
        double yj = y[j];
        for (int p = lu->Up[j]; p < pd; ++p) y[lu->Ui[p]] -= lu->Ux[p] * yj;
    }
    for (int i = 0; i < n; ++i) b[i] = y[lu->perm_c[i]];
    free(y);
}

void orc_lu_info(const orc_lu *lu, int *perm_r, int *nnz_l, int *nnz_u)
{
    if (perm_r) memcpy(perm_r, lu->perm_r, sizeof(int) * lu->n);
    if (nnz_l) *nnz_l = lu->nnz_l;
    if (nnz_u) *nnz_u = lu->nnz_u;
}

void orc_lu_free(orc_lu *lu)
{
    if (!lu) return;
    free(lu->perm_c); free(lu->iperm_c); free(lu->perm_r); free(lu->pivrow);
    free(lu->Lp); free(lu->Li); free(lu->Lx); free(lu->Up); free(lu->Ui); free(lu->Ux);
    free(lu);
}

/* CSR (0-based) -> CSC with sorted row indices (dCompRow_to_CompCol, dutil.c:154) */
static void csr_to_csc(int n, int nnz, const int *rowptr, const int *colind, const double *val,
                       int *colptr, int *rowind, double *cval)
{
    for (int i = 0; i <= n; ++i) colptr[i] = 0;
    for (int p = 0; p < nnz; ++p) colptr[colind[p] + 1]++;
    for (int i = 0; i < n; ++i) colptr[i + 1] += colptr[i];
    int *pos = malloc(sizeof(int) * (n + 1));
    memcpy(pos, colptr, sizeof(int) * (n + 1));
    for (int r = 0; r < n; ++r)
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            int q = pos[colind[p]]++;
            rowind[q] = r;
            cval[q] = val[p];
        }
    free(pos);
}

/* factor the reduced CSR system with the oracle GEPP, column order from the reference's
 * get_perm_c(COLAMD) + sp_preorder when the reference library is open, else natural */
static int gepp_factor_csr(int n, int nnz, const int *rowptr, const int *colind,
                           const double *val, orc_lu **lu)
{
    int *cp = malloc(sizeof(int) * (n + 1)), *ri = malloc(sizeof(int) * (nnz + 1));
    double *cv = malloc(sizeof(double) * (nnz + 1));
    csr_to_csc(n, nnz, rowptr, colind, val, cp, ri, cv);
    int *perm = NULL;
    if (g_ref) {
        int *pcol = malloc(sizeof(int) * n), *et = malloc(sizeof(int) * n);
        perm = malloc(sizeof(int) * n);
        g_colperm(n, nnz, cp, ri, pcol, perm, et);
        free(pcol); free(et);
    }
    int info = orc_lu_factor(n, cp, ri, cv, perm, lu);
    free(cp); free(ri); free(cv); free(perm);
    return info;
}

/* ORC_LS_EXACT: the solution of the FP64 system itself, carried as x + xlo (double-double).
 * x from the factors is refined, x += LU^-1 (b - A (x + xlo)), with the residual accumulated
 * in quad precision (__float128, 113-bit significand) and rounded once, and each correction
 * added to x + xlo by TwoSum, until the corrections are below 1e-30 |x_i| (or stop shrinking,
 * or 200 steps).  The refinement converges to the exact solution whenever cond(A) eps(FP64)
 * < 1 (Wilkinson; Demmel et al., "Error bounds from extra-precise iterative refinement", ACM
 * TOMS 32, 2006): the reference algorithm with exact linear algebra, the yardstick for the
 * rounding of any FP64 solve. */
static void exact_refine(int n, const int *rowptr, const int *colind, const double *val,
                         const double *b, double *x, double *xlo,
                         void (*solve)(const void *, double *), const void *fac)
{
    double *d = malloc(sizeof(double) * (n > 0 ? n : 1));
    double prev = HUGE_VAL;
    for (int i = 0; i < n; ++i) xlo[i] = 0.0;
    for (int it = 0; it < 200; ++it) {
        for (int i = 0; i < n; ++i) {
            __float128 r = (__float128) b[i];
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
                r -= (__float128) val[p] * ((__float128) x[colind[p]] + (__float128) xlo[colind[p]]);
            d[i] = (double) r;
        }
        solve(fac, d);
        double eta = 0.0;
        for (int i = 0; i < n; ++i) {
            const double sc = fabs(x[i]) > 0.0 ? fabs(x[i]) : 1e-300;
            const double e = fabs(d[i]) / sc;
            if (e > eta) eta = e;
            double hi = x[i], lo = xlo[i];
            dd_add(&hi, &lo, d[i]);
            x[i] = hi + lo;
            xlo[i] = lo - (x[i] - hi);
        }
        if (eta <= 1e-30 || !(eta < prev) || !(eta == eta)) break;
        prev = eta;
    }
    free(d);
}

static void solve_gepp_cb(const void *lu, double *b) { orc_lu_solve((const orc_lu *) lu, b); }

/* the reduced CSR system factored by the reference c_fortran_dgssv_ (1-based CSC copies) */
typedef struct {
    int n, nnz;
    int *cp, *ri;
    double *cv;
    long long handle;
} ref_handle;

static int ref_factor_csr(const orc_ctx *c, ref_handle *h)
{
    const int n = c->nvar1, nnz = c->red_nnz;
    h->n = n; h->nnz = nnz;
    h->cp = malloc(sizeof(int) * (n + 1)); h->ri = malloc(sizeof(int) * (nnz + 1));
    h->cv = malloc(sizeof(double) * (nnz + 1));
    csr_to_csc(n, nnz, c->red_rowptr, c->red_colind, c->red_val, h->cp, h->ri, h->cv);
    for (int i = 0; i <= n; ++i) h->cp[i] += 1;
    for (int i = 0; i < nnz; ++i) h->ri[i] += 1;
    int info = 0;
    double dummy = 0.0;
    orc_ref_dgssv(1, n, nnz, 1, h->cv, h->ri, h->cp, &dummy, n, &h->handle, &info);
    return info;
}

static void solve_ref_cb(const void *hv, double *b)
{
    ref_handle *h = (ref_handle *) hv;
    int info = 0;
    orc_ref_dgssv(2, h->n, h->nnz, 1, h->cv, h->ri, h->cp, b, h->n, &h->handle, &info);
}

static void ref_free(ref_handle *h)
{
    if (h->handle) {
        int info = 0;
        double dummy = 0.0;
        orc_ref_dgssv(3, h->n, h->nnz, 1, h->cv, h->ri, h->cp, &dummy, h->n, &h->handle, &info);
    }
    free(h->cp); free(h->ri); free(h->cv);
    h->cp = NULL; h->ri = NULL; h->cv = NULL; h->handle = 0;
}

/* solve_Ax_equals_b_csr option 5 (linalg.F90:110-124): rhs overwritten by the solution */
static int solve_reduced(orc_ctx *c, orc_stats *st)
{
    int n = c->nvar1, nnz = c->red_nnz, ierr = 0;
    if (st) { st->n_factor++; st->n_solve++; }
    if (c->lsolver == ORC_LS_EXACT) {
        double *b = malloc(sizeof(double) * (n > 0 ? n : 1));
        memcpy(b, c->red_rhs, sizeof(double) * n);
        if (g_ref) {   /* the reference SuperLU's factors (faster than the oracle GEPP) */
            ref_handle h = {0};
            ierr = ref_factor_csr(c, &h);
            if (ierr == 0) {
                solve_ref_cb(&h, c->red_rhs);
                exact_refine(n, c->red_rowptr, c->red_colind, c->red_val, b, c->red_rhs,
                             c->red_xlo, solve_ref_cb, &h);
            }
            ref_free(&h);
        } else {
            orc_lu *lu = NULL;
            ierr = gepp_factor_csr(n, nnz, c->red_rowptr, c->red_colind, c->red_val, &lu);
            if (ierr == 0) {
                orc_lu_solve(lu, c->red_rhs);
                exact_refine(n, c->red_rowptr, c->red_colind, c->red_val, b, c->red_rhs,
                             c->red_xlo, solve_gepp_cb, lu);
            }
            orc_lu_free(lu);
        }
        free(b);
        return ierr;
    }
    if (c->lsolver == ORC_LS_REF && g_ref) {
        /* slu_solve takes 1-based CSR and shifts it in place; the caller restores it */
        for (int i = 0; i <= n; ++i) c->red_rowptr[i] += 1;
        for (int i = 0; i < nnz; ++i) c->red_colind[i] += 1;
        g_slu_solve(n, nnz, 1, 1, c->red_val, c->red_colind, c->red_rowptr, c->red_rhs, &ierr);
        return ierr;
    }
    orc_lu *lu = NULL;
    ierr = gepp_factor_csr(n, nnz, c->red_rowptr, c->red_colind, c->red_val, &lu);
    if (ierr == 0) orc_lu_solve(lu, c->red_rhs);
    orc_lu_free(lu);
    return ierr;
}

/* ------------------------------------------------------------------------------------------ */
/* utilities                                                                                  */

/* indexx, sort.F90:73-125 (Numerical Recipes heapsort, 1-based internally) */
static void indexx(int n, const double *arrin, int *indx)
{
    for (int j = 1; j <= n; ++j) indx[j - 1] = j;
    if (n == 1) return;
    int l = n / 2 + 1, ir = n, indxt, i, j;
    double q;
    for (;;) {
        if (l > 1) {
            l = l - 1;
            indxt = indx[l - 1];
            q = arrin[indxt - 1];
        } else {
            indxt = indx[ir - 1];
            q = arrin[indxt - 1];
            indx[ir - 1] = indx[0];
            ir = ir - 1;
            if (ir == 1) { indx[0] = indxt; return; }
        }
        i = l;
        j = l + l;
        while (j <= ir) {
            if (j < ir)
                if (arrin[indx[j - 1] - 1] < arrin[indx[j] - 1]) j = j + 1;
            if (q < arrin[indx[j - 1] - 1]) {
                indx[i - 1] = indx[j - 1];
                i = j;
                j = j + j;
            } else {
                j = ir + 1;
            }
        }
        indx[i - 1] = indxt;
    }
}

/* sortedsum, utils.F90:481-501 */
double orc_sortedsum(int n, const double *a)
{
    int *ix = malloc(sizeof(int) * (n > 0 ? n : 1));
    indexx(n, a, ix);
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + a[ix[i] - 1];
    free(ix);
    return s;
}

/* calculate_ye, utils.F90:537-553 */
double orc_calculate_ye(const orc_ctx *c, const double *x)
{
    double *eta = malloc(sizeof(double) * c->nsp);
    for (int i = 0; i < c->nsp; ++i)
        eta[i] = c->consider[i] ? (c->aw[i] - 2.0 * c->zw[i]) * x[i] / c->aw[i] : 0.0;
    double ye = (1.0 - orc_sortedsum(c->nsp, eta)) * 0.5;
    free(eta);
    return ye;
}

/* ypsplus, utils.F90:726-740 */
void orc_ypsplus(int n, double *y)
{
    for (int i = 0; i < n; ++i)
        if (y[i] < -1.0e-27) y[i] = 1.0e-99;
    for (int i = 0; i < n; ++i)
        if (y[i] < 1.0e-27) y[i] = 1.0e-99;
}

static double plain_sum(int n, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + y[i];
    return s;
}

static double maxval_neg(int n, const double *y)
{
    double m = -DBL_MAX;
    for (int i = 0; i < n; ++i) if (-y[i] > m) m = -y[i];
    return m;
}

/* ------------------------------------------------------------------------------------------ */
/* newton_iteration, backward_euler.F90:120-176 */
static int newton_iteration(orc_ctx *c, const double *y_old, const double *y_cur, double dt,
                            double *grd, double *y_new, orc_stats *st)
{
    const int nsp = c->nsp;
    if (orc_jacobian(c, dt, y_cur, c->v, c->jac, c->jrow, c->jcol, c->mask)) return BAD_STEP;
    for (int i = 0; i < c->nnz; ++i) c->jac[i] = c->jac[i] * (-dt);
    int nn = orc_reduce_jacobian(c, c->jac, c->jrow, c->jcol, c->mask, &c->nvar1, c->red_rowptr,
                                 c->red_colind, c->red_val);
    if (nn < 0) return BAD_STEP;
    c->red_nnz = nn;
    orc_dxdt(c, y_cur, c->v, c->dxdt);
    for (int i = 0; i < nsp; ++i) {
        if (c->lsolver != ORC_LS_EXACT) {
            c->rhs[i] = y_old[i] - y_cur[i] + dt * c->dxdt[i];
            continue;
        }
        /* y_old - y_cur + dt S (S the double-double dX/dt), rounded once */
        double hi = y_old[i], lo = 0.0;
        dd_add(&hi, &lo, -y_cur[i]);
        const double p = dt * c->dxdt[i];
        lo += fma(dt, c->dxdt[i], -p) + dt * c->dxdt_lo[i];
        dd_add(&hi, &lo, p);
        c->rhs[i] = hi + lo;
    }
    int m = 0;
    for (int i = 0; i < nsp; ++i) if (c->mask[i]) c->red_rhs[m++] = c->rhs[i];
    int ierr = solve_reduced(c, st);
    if (ierr != 0) return BAD_STEP;
    /* reduced_y_new = reduced_y_cur + reduced_rhs; ddecompac back into y_new */
    m = 0;
    for (int i = 0; i < nsp; ++i)
        if (c->mask[i]) {
            if (c->lsolver == ORC_LS_EXACT) {   /* y_cur + (x + xlo), rounded once */
                double hi = y_cur[i], lo = c->red_xlo[m];
                dd_add(&hi, &lo, c->red_rhs[m]);
                y_new[i] = hi + lo;
            } else {
                y_new[i] = y_cur[i] + c->red_rhs[m];
            }
            ++m;
        }
    double g = 0.0;
    for (int i = 0; i < nsp; ++i) {
        double d = y_new[i] - y_cur[i];
        double ga = (y_cur[i] > c->kn.grdthreshold) ? fabs(d / y_cur[i]) : 0.0;
        if (ga > g) g = ga;   /* maxval */
    }
    *grd = g;
    if (st) st->nvar1 = c->nvar1;
    return 0;
}

/* backward_euler_newton, backward_euler.F90:29-109.  x ** HALF is evaluated as sqrt(x), the
 * correctly rounded value of the same quantity (libm pow differs from it by < 1 ulp). */
int orc_backward_euler(orc_ctx *c, double *yps, double dt, double *dtrcmd, orc_stats *st)
{
    const int nsp = c->nsp;
    const double rmin = 0.2, rmax = 2.0;
    const double dgrd = c->kn.dgrd;
    const int nvar = nsp - 1;   /* count(considerisotope) */
    const double sum_check = c->kn.grdthreshold * nvar * dgrd;
    double *y_old = malloc(sizeof(double) * nsp), *y_cur = malloc(sizeof(double) * nsp);
    double *y_new = malloc(sizeof(double) * nsp);
    memcpy(y_old, yps, sizeof(double) * nsp);
    memcpy(y_cur, yps, sizeof(double) * nsp);
    memcpy(y_new, yps, sizeof(double) * nsp);
    int iter = 0, ierr = 0;
    double grd = 0.0;
    if (st) { st->titers = 0; st->iter = 0; }
    for (;;) {
        if (iter > c->kn.ittd) {
            double err, tol;
            double mneg = maxval_neg(nsp, y_new);
            if (mneg >= sum_check) { err = mneg; tol = sum_check; ierr = NEG_ABUNDS; }
            else if (fabs(1.0 - plain_sum(nsp, y_new)) > SUM_TOL) {
                err = fabs(1.0 - plain_sum(nsp, y_new)); tol = SUM_TOL; ierr = SUM_X;
            } else { err = grd; tol = dgrd; ierr = SLV_CNVRG; }
            *dtrcmd = fmax(fmin(sqrt(tol / fmax(err, SMALL_C)), rmax), rmin) * dt;
            break;
        }
        int e = newton_iteration(c, y_old, y_cur, dt, &grd, y_new, st);
        if (e) { ierr = e; *dtrcmd = rmin * dt; break; }
        iter = iter + 1;
        if (st) { st->titers++; st->iter++; st->n_newton++; }
        if (grd <= dgrd && maxval_neg(nsp, y_new) < sum_check &&
            fabs(1.0 - plain_sum(nsp, y_new)) <= SUM_TOL) {
            memcpy(yps, y_new, sizeof(double) * nsp);
            *dtrcmd = 2.0 * fmax(fmin(sqrt(dgrd / fmax(grd, SMALL_C)), rmax), rmin) * dt;
            orc_ypsplus(nsp, yps);
            break;
        }
        memcpy(y_cur, y_new, sizeof(double) * nsp);
    }
    free(y_old); free(y_cur); free(y_new);
    return ierr;
}

/* ------------------------------------------------------------------------------------------ */
/* Bader-Deuflhard, bader_deuflhard.F90:139-488                                               */

typedef struct {
    orc_lu *lu;         /* oracle GEPP factors */
    long long handle;   /* reference c_fortran_dgssv_ handle */
    int *cp, *ri;       /* 1-based CSC for the reference */
    double *cv;
} bd_factors;

static int bd_step(orc_ctx *c, double *y_k, double *delta_k, double h, int stage,
                   bd_factors *f, orc_stats *st)
{
    const int nsp = c->nsp, n = c->nvar1, nnz = c->red_nnz;
    int ierr = 0;
    double *x = calloc(nsp, sizeof(double));
    orc_dxdt(c, y_k, c->v, c->dxdt);
    for (int i = 0; i < nsp; ++i)
        c->rhs[i] = (stage == 1) ? h * c->dxdt[i] : h * c->dxdt[i] - delta_k[i];
    int m = 0;
    for (int i = 0; i < nsp; ++i) if (c->mask[i]) c->red_rhs[m++] = c->rhs[i];
    const int use_ref = (c->lsolver == ORC_LS_REF && g_ref);
    if (stage == 1) {
        if (st) st->n_factor++;
        if (use_ref) {
            /* reduced CSR -> CSC (csrcsc, jac_rhs.F90:200-213), 1-based for c_fortran_dgssv */
            f->cp = malloc(sizeof(int) * (n + 1)); f->ri = malloc(sizeof(int) * (nnz + 1));
            f->cv = malloc(sizeof(double) * (nnz + 1));
            csr_to_csc(n, nnz, c->red_rowptr, c->red_colind, c->red_val, f->cp, f->ri, f->cv);
            for (int i = 0; i <= n; ++i) f->cp[i] += 1;
            for (int i = 0; i < nnz; ++i) f->ri[i] += 1;
            int info = 0;
            orc_ref_dgssv(1, n, nnz, 1, f->cv, f->ri, f->cp, c->red_rhs, n, &f->handle, &info);
        } else {
            gepp_factor_csr(n, nnz, c->red_rowptr, c->red_colind, c->red_val, &f->lu);
        }
    }
    if (st) st->n_solve++;
    if (use_ref) {
        int info = 0;
        orc_ref_dgssv(2, n, nnz, 1, f->cv, f->ri, f->cp, c->red_rhs, n, &f->handle, &info);
    } else if (c->lsolver == ORC_LS_EXACT) {
        double *b = malloc(sizeof(double) * (n > 0 ? n : 1));
        memcpy(b, c->red_rhs, sizeof(double) * n);
        orc_lu_solve(f->lu, c->red_rhs);
        exact_refine(n, c->red_rowptr, c->red_colind, c->red_val, b, c->red_rhs, c->red_xlo,
                     solve_gepp_cb, f->lu);
        for (int i = 0; i < n; ++i) c->red_rhs[i] = c->red_rhs[i] + c->red_xlo[i];
        free(b);
    } else {
        orc_lu_solve(f->lu, c->red_rhs);
    }
    m = 0;
    for (int i = 0; i < nsp; ++i) if (c->mask[i]) x[i] = c->red_rhs[m++];
    if (stage == 2) for (int i = 0; i < nsp; ++i) delta_k[i] = delta_k[i] + 2.0 * x[i];
    else for (int i = 0; i < nsp; ++i) delta_k[i] = x[i];
    for (int i = 0; i < nsp; ++i) y_k[i] = y_k[i] + delta_k[i];
    int huge = 0;
    for (int i = 0; i < nsp; ++i) if (y_k[i] > 2.0) huge = 1;
    if (stage == 3 || huge) {
        if (use_ref) {
            int info = 0;
            orc_ref_dgssv(3, n, nnz, 1, f->cv, f->ri, f->cp, c->red_rhs, n, &f->handle, &info);
            free(f->cp); free(f->ri); free(f->cv); f->cp = NULL; f->ri = NULL; f->cv = NULL;
        } else {
            orc_lu_free(f->lu); f->lu = NULL;
        }
    }
    if (huge) ierr = HUGE_ABUNDS;
    free(x);
    return ierr;
}

/* semi_implicit_extrapolation, bader_deuflhard.F90:276-316 */
static int semi_implicit_extrapolation(orc_ctx *c, double dt, int k, double *y_k, orc_stats *st)
{
    const int nsp = c->nsp;
    double h = dt / (double) c->nseq[k - 1];
    c->xest = (h / c->nseq[k - 1]) * (h / c->nseq[k - 1]);
    for (int i = 0; i < c->nnz; ++i) c->jac[i] = c->jac_cpy[i] * (-h);
    int nn = orc_reduce_jacobian(c, c->jac, c->jrow, c->jcol, c->mask, &c->nvar1, c->red_rowptr,
                                 c->red_colind, c->red_val);
    if (nn < 0) return BAD_STEP;
    c->red_nnz = nn;
    double *delta = calloc(nsp, sizeof(double));
    bd_factors f = {0};
    int ierr = bd_step(c, y_k, delta, h, 1, &f, st);
    for (int i = 0; !ierr && i < c->nseq[k - 1] - 1; ++i) ierr = bd_step(c, y_k, delta, h, 2, &f, st);
    if (!ierr) ierr = bd_step(c, y_k, delta, h, 3, &f, st);
    free(delta);
    return ierr;
}

/* extrapolate (Neville), bader_deuflhard.F90:371-407 */
static void extrapolate(orc_ctx *c, int k, const double *yest, double *y_extrap)
{
    const int nsp = c->nsp;
    double *qcol = c->qcol, *yerr = c->yerr;   /* (j, col) -> [col*nsp + j] */
    c->xcol[k - 1] = c->xest;
    double *yz = malloc(sizeof(double) * nsp), *d = malloc(sizeof(double) * nsp);
    for (int j = 0; j < nsp; ++j) { yz[j] = yest[j]; yerr[(k - 1) * nsp + j] = yest[j]; }
    if (k == 1) {
        for (int j = 0; j < nsp; ++j) { qcol[j] = yest[j]; y_extrap[j] = yest[j]; }
        free(yz); free(d);
        return;
    }
    for (int j = 0; j < nsp; ++j) d[j] = yest[j];
    for (int k1 = 1; k1 <= k - 1; ++k1) {
        double delta = 1.0 / (c->xcol[k - k1 - 1] - c->xest);
        double f1 = c->xest * delta;
        double f2 = c->xcol[k - k1 - 1] * delta;
        for (int j = 0; j < nsp; ++j) {
            double q = qcol[(k1 - 1) * nsp + j];
            qcol[(k1 - 1) * nsp + j] = yerr[(k - 1) * nsp + j];
            double dl = d[j] - q;
            yerr[(k - 1) * nsp + j] = f1 * dl;
            d[j] = f2 * dl;
            yz[j] = yz[j] + yerr[(k - 1) * nsp + j];
        }
    }
    for (int j = 0; j < nsp; ++j) { qcol[(k - 1) * nsp + j] = yerr[(k - 1) * nsp + j]; y_extrap[j] = yz[j]; }
    free(yz); free(d);
}

/* compute_error, bader_deuflhard.F90:325-366 */
static void compute_error(orc_ctx *c, int k, const double *y_k, const double *yscal)
{
    const int nsp = c->nsp;
    double ab = -DBL_MAX, ng = -DBL_MAX;
    for (int j = 0; j < nsp; ++j) {
        double e = fabs(c->yerr[(k - 1) * nsp + j] / yscal[j]);
        if (y_k[j] > 1.0e-27 && e > ab) ab = e;
        if (y_k[j] < c->kn.grdthreshold && -y_k[j] > ng) ng = -y_k[j];
    }
    c->abuerr = ab / c->kn.dgrd;
    c->negerr = ng / c->kn.grdthreshold;
    if (c->abuerr < 1.0e-3) {
        c->errmax = c->abuerr;
    } else {
        c->sumerr = fabs(1.0 - orc_sortedsum(nsp, y_k)) / SUM_TOL;
        c->errmax = fmax(c->abuerr, c->sumerr);
    }
    c->errmax = fmax(c->errmax, c->negerr);
}

/* bader_deuflhard_integration, bader_deuflhard.F90:139-272 */
int orc_bader_deuflhard(orc_ctx *c, double *yps, double dt, double *dtrcmd, orc_stats *st)
{
    const int nsp = c->nsp;
    const double safe1 = 0.25, safe2 = 0.7, redmin = 0.7, redmax = 1.0e-5, scalmx = 0.25;
    double scal = 1.0, red = -1.0, work, workmin, fact;
    int km = -1, ierr = 0, k;
    double *y_k = malloc(sizeof(double) * nsp), *y_extrap = malloc(sizeof(double) * nsp);
    double *yscal = malloc(sizeof(double) * nsp);
    c->nvar1 = 0;
    if (st) st->iter = 0;
    memcpy(y_k, yps, sizeof(double) * nsp);
    if (orc_jacobian(c, dt, yps, c->v, c->jac, c->jrow, c->jcol, c->mask)) {
        *dtrcmd = dt * 0.2; ierr = BAD_STEP; goto done;
    }
    memcpy(c->jac_cpy, c->jac, sizeof(double) * c->nnz);
    for (k = 1; k <= c->kmax; ++k) {
        memcpy(y_k, yps, sizeof(double) * nsp);
        ierr = semi_implicit_extrapolation(c, dt, k, y_k, st);
        if (ierr != 0) { *dtrcmd = dt * 0.2; goto done; }
        if (st) { st->iter += c->nseq[k - 1]; st->nvar1 = c->nvar1; }
        extrapolate(c, k, y_k, y_extrap);
        memcpy(y_k, y_extrap, sizeof(double) * nsp);
        for (int j = 0; j < nsp; ++j)
            if (y_k[j] != y_k[j]) { ierr = GOT_NAN; *dtrcmd = dt * 0.2; goto done; }
        for (int j = 0; j < nsp; ++j) yscal[j] = fmax(fabs(y_k[j]), c->kn.grdthreshold);
        if (k != 1) {
            compute_error(c, k, y_k, yscal);
            km = k - 1;
            c->err[km - 1] = pow(c->errmax / safe1, 1.0 / (2.0 * km + 1.0));
        }
        if (k != 1 && (k >= c->kopt - 1 || c->frst)) {
            if (c->errmax <= 1.0) {
                c->reduce = 0;
                break;
            } else {
                if (k == c->kmax || k == c->kopt + 1) {
                    c->reduce = 1; red = safe2 / c->err[km - 1];
                } else if (k == c->kopt) {
                    if (c->alf[c->kopt - 2][c->kopt - 1] < c->err[km - 1]) {
                        c->reduce = 1; red = safe2 / c->err[km - 1];
                    }
                } else if (c->kopt == c->kmax) {
                    if (c->alf[km - 1][c->kmax - 2] < c->err[km - 1]) {
                        c->reduce = 1; red = c->alf[km - 1][c->kmax - 2] * safe2 / c->err[km - 1];
                    }
                } else if (c->alf[km - 1][c->kopt - 1] < c->err[km - 1]) {
                    c->reduce = 1; red = c->alf[km - 1][c->kopt - 2] / c->err[km - 1];
                }
            }
        }
    }
    if (c->reduce) {
        red = fmax(redmax, fmin(red, redmin));
        *dtrcmd = dt * red;
        if (c->negerr > c->sumerr && c->negerr > c->abuerr) ierr = NEG_ABUNDS;
        else if (c->sumerr > c->abuerr) ierr = SUM_X;
        else ierr = SLV_CNVRG;
    } else {
        workmin = 1.0e35;
        for (int kk = 1; kk <= km; ++kk) {
            fact = fmax(c->err[kk - 1], scalmx);
            work = fact * c->a[kk];
            if (work < workmin) { scal = fact; workmin = work; c->kopt = kk + 1; }
        }
        *dtrcmd = 2.0 * dt / scal;
        if (c->kopt >= k && c->kopt != c->kmax) {
            fact = fmax(scal / c->alf[c->kopt - 2][c->kopt - 1], scalmx);
            if (c->a[c->kopt] * fact <= workmin) {
                *dtrcmd = *dtrcmd * scal / fact;
                c->kopt = c->kopt + 1;
            }
        }
    }
    memcpy(yps, y_k, sizeof(double) * nsp);
done:
    free(y_k); free(y_extrap); free(yscal);
    return ierr;
}

/* ------------------------------------------------------------------------------------------ */
/* linear interpolant over [0, dtwant] (interp.F90:17-50 with the 2-point table) */
static double interp2(double x0, double x1, double y0, double y1, double xw)
{
    double x = fmax(xw, x0);
    x = fmin(x, x1);
    /* bsearch_r8 on a 2-point table always returns idx 1 (or 0 clipped) */
    return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
}

int orc_integrate_zone(orc_ctx *c, double *yps, double t9_0, double t9_1, double rho_0,
                       double rho_1, double *ye, double dtwant, orc_stats *st)
{
    return orc_integrate_zone_v(c, NULL, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, st);
}

int orc_integrate_zone_v(orc_ctx *c, const double *v_mid, double *yps, double t9_0, double t9_1,
                         double rho_0, double rho_1, double *ye, double dtwant, orc_stats *st)
{
    const int nsp = c->nsp;
    orc_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    /* caller side: calculate_ye + evaluate_all_rates at the cycle mid-point (ppn.F90:289-298) */
    double t9m = 0.5 * (t9_0 + t9_1), rhom = 0.5 * (rho_0 + rho_1);
    *ye = orc_calculate_ye(c, yps);
    if (v_mid) memcpy(c->v, v_mid, sizeof(double) * c->nre);
    else orc_rates_full(c, t9m, rhom, *ye, yps, c->v);

    /* integrate_network, solver.F90:53-238 */
    double dtdone = 0.0, dtrcmd = 0.0, ye_old = *ye;
    double *yps_old = malloc(sizeof(double) * nsp);
    memcpy(yps_old, yps, sizeof(double) * nsp);
    double my_t9_last = 0.5 * (t9_0 + t9_1), my_rho_last = 0.5 * (rho_0 + rho_1);
    double dttry = dtwant, my_t9_0 = 0, my_t9_1 = 0, my_rho_0 = 0, my_rho_1 = 0, my_t9, my_rho;
    int i, ierr = 0;
    for (i = 1; i <= c->kn.max_sub_steps; ++i) {
        if (i > 1) c->frst = 0;
        if (c->kn.interpolate_temp) {
            double t_0 = dtdone, t_1 = fmin(dtdone + dttry, dtwant);
            my_t9_0 = interp2(0.0, dtwant, t9_0, t9_1, t_0);
            my_t9_1 = interp2(0.0, dtwant, t9_0, t9_1, t_1);
            my_t9 = 0.5 * (my_t9_0 + my_t9_1);
        } else {
            my_t9 = 0.5 * (t9_0 + t9_1);
        }
        if (c->kn.interpolate_dens) {
            double t_0 = dtdone, t_1 = fmin(dtdone + dttry, dtwant);
            my_rho_0 = interp2(0.0, dtwant, rho_0, rho_1, t_0);
            my_rho_1 = interp2(0.0, dtwant, rho_0, rho_1, t_1);
            my_rho = 0.5 * (my_rho_0 + my_rho_1);
        } else {
            my_rho = 0.5 * (rho_0 + rho_1);
        }
        double diff = fabs(my_t9_0 - my_t9_1) / my_t9_0;
        diff = fmax(diff, fabs(my_rho_0 - my_rho_1) / my_rho_0);
        if (diff > c->kn.delta_state_max && t9_0 > c->kn.min_t9_for_slv &&
            t9_1 > c->kn.min_t9_for_slv) {
            dttry = 0.5 * dttry;
            continue;
        }
        if (my_t9 != my_t9_last || my_rho != my_rho_last) {
            orc_rates_full(c, my_t9, my_rho, *ye, yps, c->v);
            my_rho_last = my_rho; my_t9_last = my_t9;
        }
        ierr = 0;
        if (my_t9 >= c->kn.min_t9_for_slv) {
            if (c->kn.integration_method == 0) ierr = orc_backward_euler(c, yps, dttry, &dtrcmd, st);
            else ierr = orc_bader_deuflhard(c, yps, dttry, &dtrcmd, st);
            *ye = orc_calculate_ye(c, yps);
            /* where the reference stops (zero in the matrix, singular factor) the oracle
             * returns BAD_STEP to its caller */
            if (ierr == BAD_STEP) goto out;
            if (dttry < SMALL_C) { ierr = SMALL_DT; goto out; }
        } else {
            ierr = 0;
        }
        if (ierr == 0) {
            dtdone = dtdone + dttry;
            for (int j = 0; j < nsp; ++j) yps[j] = fmax(yps[j], 1.0e-99);
            memcpy(yps_old, yps, sizeof(double) * nsp);
            ye_old = *ye;
            st->nsucc++;
            if (dtdone >= dtwant) break;
        } else {
            memcpy(yps, yps_old, sizeof(double) * nsp);
            *ye = ye_old;
            st->nfail++;
        }
        dttry = fmin(dtrcmd, dtwant - dtdone);
        if (i == c->kn.max_sub_steps) { ierr = MAX_SUBSTEPS; goto out; }
    }
    st->nsubt = i;
    if (c->kn.integration_method == 1) {
        double s = plain_sum(nsp, yps);
        for (int j = 0; j < nsp; ++j) yps[j] = yps[j] / s;
    }
    for (int j = 0; j < nsp; ++j) yps[j] = fmax(yps[j], 1.0e-99);
    ierr = 0;
out:
    free(yps_old);
    return ierr;
}

/* ------------------------------------------------------------------------------------------ */
/* The mppnp cycle around the integration: zone selection, per-zone renormalisation, diffusive
 * mixing over the zones (frames/mppnp/source/mppnp.F90:637-781, 1080-1097;
 * frames/mppnp/source/mixing.F90; utils/source/utils.F90:387-469,670-684) */

/* zoneselect, mppnp.F90:1080-1097: zones hotter than the threshold; returns mmax (1-based) */
int orc_zoneselect(double t9threshold, const double *t9, int m, unsigned char *sel)
{
    int mmax = 0;
    for (int j = 1; j <= m; ++j) {
        sel[j - 1] = 0;
        if (t9[j - 1] > t9threshold) {
            sel[j - 1] = 1;
            mmax = j;
        }
    }
    return mmax;
}

/* ksum (Kahan summation), utils.F90:670-684 */
double orc_ksum(int n, const double *a)
{
    double c = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double y = a[i] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    return s;
}

/* check_1d_mass_cons without sorting, utils.F90:387-397: |1 - sum(yps)| */
double orc_mass_residual(int n, const double *y)
{
    return fabs(1.0 - plain_sum(n, y));
}

/* renorm_1d, utils.F90:433-444 (the message aside): yps = yps / ksum(yps) */
void orc_renorm_1d(int n, double *y)
{
    const double s = orc_ksum(n, y);
    for (int i = 0; i < n; ++i) y[i] = y[i] / s;
}

/* mix_calculate_D_Lagrangian, mixing.F90:68-110: D_lagr = D_eul (4 pi Rsun^2 / Msun)^2
 * exp(2 (2 ln r + ln rho)), r and rho averaged in logarithm over the zone and its inner
 * neighbour; 0 where D_eul <= 0.  (dconst ** 2 is the compile-time constant dconst * dconst.) */
void orc_mix_d_lagrangian(int m, const double *rho, const double *rad, const double *d_eul,
                          double *d_lagr)
{
    const double dconst = 3.0589703419947736e-11;
    const double dc2 = dconst * dconst;
    for (int j = 0; j < m; ++j) {
        d_lagr[j] = 0.0;
        if (d_eul[j] <= 0.0) continue;
        double lnrho, lnrad;
        if (j == 0) {
            lnrho = log(rho[j]);
            lnrad = log(rad[j]);
        } else {
            lnrho = (log(rho[j]) + log(rho[j - 1])) / 2.0;
            lnrad = (log(rad[j]) + log(rad[j - 1])) / 2.0;
        }
        d_lagr[j] = d_eul[j] * dc2 * exp(2.0 * (2.0 * lnrad + lnrho));
    }
}

/* matrixdiffus, mixing.F90:180-250 (code_source GNV / EVL): the implicit diffusion matrix on
 * mass coordinates xm; D is limited in place where D(j+1) h / (xmm xr) exceeds sig_term_limit.
 * Returns 0, or 1 where the reference stops (a zero zone width). */
int orc_matrixdiffus(int m, double h, const double *xm, double *D, double sig, double *a,
                     double *b, double *c)
{
    /* Fortran indices 1..m as [j - 1] */
    double xr = xm[1] - xm[0];
    double alpha = D[1] * h / (xr * xr);
    if (sig < alpha) alpha = sig;
    a[0] = 0.0;
    b[0] = 1.0 + alpha;
    c[0] = -1.0 * alpha;
    double xl = 0.0;
    for (int j = 2; j <= m - 1; ++j) {
        xl = xm[j - 1] - xm[j - 2];
        xr = xm[j] - xm[j - 1];
        const double xmm = 0.5 * (xm[j] - xm[j - 2]);
        alpha = h / xmm;
        if (sig < D[j] * alpha / xr) D[j] = sig / alpha * xr;
        a[j - 1] = -1.0 * D[j - 1] * alpha / xl;
        b[j - 1] = 1.0 + (D[j] / xr + D[j - 1] / xl) * alpha;
        c[j - 1] = -1.0 * D[j] * alpha / xr;
        if (xl == 0.0 || xr == 0.0) return 1;
    }
    xl = xm[m - 1] - xm[m - 2];
    alpha = D[m - 1] * h / (xl * xl);
    a[m - 1] = -1.0 * alpha;
    b[m - 1] = 1.0 + alpha;
    c[m - 1] = 0.0;
    if (xl == 0.0 || xr == 0.0) return 1;
    return 0;
}

/* reduce_term, mixing.F90:325-340 */
static void reduce_term(double alpha, double onedxr, double *d_coeff, double sig)
{
    if (alpha * onedxr * *d_coeff > sig) *d_coeff = sig / (alpha * onedxr);
}

/* matrixdiffus_mesa, mixing.F90:254-322 (code_source MES): the same on shell masses dq, with
 * the limited coefficients kept apart (d_reduce) */
void orc_matrixdiffus_mesa(int m, double h, const double *dq, const double *d, double sig,
                           double *a, double *b, double *c)
{
    double *dr = calloc((size_t) m + 1, sizeof(double));
    dr[0] = d[0];
    double xr = dq[0];
    double alpha = h / xr;
    double term_c = 1.0 / xr;
    double d_coeff = d[1];
    reduce_term(alpha, term_c, &d_coeff, sig);
    dr[1] = d_coeff;
    a[0] = 0.0;
    b[0] = 1.0 + term_c * alpha * dr[1];
    c[0] = -1.0 * term_c * alpha * dr[1];
    for (int j = 2; j <= m - 1; ++j) {
        const double xl = dq[j - 2];
        xr = dq[j - 1];
        const double xmm = 0.5 * (dq[j - 1] + dq[j - 2]);
        alpha = h / xmm;
        const double term_a = 1.0 / xl;
        term_c = 1.0 / xr;
        d_coeff = d[j];
        reduce_term(alpha, term_c, &d_coeff, sig);
        dr[j] = d_coeff;
        const double term_b = term_c * dr[j] + term_a * dr[j - 1];
        a[j - 1] = -1.0 * term_a * dr[j - 1] * alpha;
        b[j - 1] = 1.0 + term_b * alpha;
        c[j - 1] = -1.0 * term_c * dr[j] * alpha;
    }
    const double xl = dq[m - 1];
    alpha = h / (xl * xl);
    a[m - 1] = -1.0 * alpha * d[m - 1];
    b[m - 1] = 1.0 + alpha * d[m - 1];
    c[m - 1] = 0.0;
    free(dr);
}

/* tridag, mixing.F90:140-176 (Numerical Recipes' Thomas algorithm).  Returns 0, or 1 where the
 * reference stops (a zero pivot). */
int orc_tridag(int n, const double *a, const double *b, const double *c, const double *r,
               double *u)
{
    double *gam = malloc(sizeof(double) * (size_t) (n + 1));
    if (b[0] == 0.0) { free(gam); return 1; }
    double bet = b[0];
    u[0] = r[0] / bet;
    for (int j = 1; j < n; ++j) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        if (bet == 0.0) { free(gam); return 1; }
        u[j] = (r[j] - a[j] * u[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; --j) u[j] = u[j] - gam[j + 1] * u[j + 1];
    free(gam);
    return 0;
}

/* do_mixing with idifmix = 1 (mixing.F90:30-64): D_lagr from D_eul, the matrix
 * (matrixdiffus, or matrixdiffus_mesa for code_source MES), and one tridag per considered
 * species over the zones, species 1..NSP-1 (the loop of mixdiff08, mixing.F90:118-130, leaves
 * the last species out).  yps[m][nsp] zone-major; consider[nsp].  Returns 0 or 1 (stop). */
int orc_do_mixing(int m, int nsp, double *yps, const unsigned char *consider, const double *xm,
                  const double *dq, const double *rho, const double *rad, const double *d_eul,
                  double dzeit, int mesa, double sig)
{
    double *dl = malloc(sizeof(double) * (size_t) m), *a = malloc(sizeof(double) * (size_t) m);
    double *b = malloc(sizeof(double) * (size_t) m), *c = malloc(sizeof(double) * (size_t) m);
    double *yn = malloc(sizeof(double) * (size_t) m), *u = malloc(sizeof(double) * (size_t) m);
    int rc = 0;
    orc_mix_d_lagrangian(m, rho, rad, d_eul, dl);
    if (mesa) {
        if (!(dq[0] > 0.0)) rc = 1;
        else orc_matrixdiffus_mesa(m, dzeit, dq, dl, sig, a, b, c);
    } else {
        rc = orc_matrixdiffus(m, dzeit, xm, dl, sig, a, b, c);
    }
    for (int s = 0; s < nsp - 1 && rc == 0; ++s) {
        if (!consider[s]) continue;
        for (int j = 0; j < m; ++j) yn[j] = yps[(size_t) j * nsp + s];
        rc = orc_tridag(m, a, b, c, yn, u);
        if (rc == 0)
            for (int j = 0; j < m; ++j) yps[(size_t) j * nsp + s] = u[j];
    }
    free(dl); free(a); free(b); free(c); free(yn); free(u);
    return rc;
}

/* One mppnp cycle (mppnp.F90:637-781) on zones that share this network: zoneselect; for every
 * selected zone calculate_ye + evaluate_all_rates at (T9, rho) + integrate_network over dzeit
 * at constant T9 and rho (719-724), then the mass check and renorm_1d above 5e-7 (737-738);
 * a zone's integration error ends the cycle there (the reference stops, 726-735);
 * do_mixing over all m zones (765), the 2-d mass check and renorm_2d when the largest residual
 * reaches SUM_TOL (768-773), and the clip max(yps, 1e-99) (781).  yps[m][nsp]; ierr[m] (0 for
 * unselected zones); stats[m] may be NULL.  Returns 0, or 1 where mixing stops. */
int orc_mppnp_cycle(orc_ctx *c, int m, double *yps, const double *t9, const double *rho,
                    const double *xm, const double *dq, const double *rad, const double *d_eul,
                    double dzeit, double t9threshold, int mesa, double sig, int *ierr,
                    orc_stats *stats)
{
    const int nsp = c->nsp;
    unsigned char *sel = malloc((size_t) m);
    orc_zoneselect(t9threshold, t9, m, sel);
    for (int j = 0; j < m; ++j) {
        ierr[j] = 0;
        if (stats) memset(&stats[j], 0, sizeof(orc_stats));
        if (!sel[j]) continue;
        double *y0 = yps + (size_t) j * nsp;
        double ye = 0.0;
        ierr[j] = orc_integrate_zone(c, y0, t9[j], t9[j], rho[j], rho[j], &ye, dzeit,
                                     stats ? &stats[j] : NULL);
        if (orc_mass_residual(nsp, y0) > 5.0e-7) orc_renorm_1d(nsp, y0);
    }
    free(sel);
    /* any integration error stops mppnp before mixing (mppnp.F90:726-735) */
    for (int j = 0; j < m; ++j)
        if (ierr[j] != 0) return 0;
    int rc = orc_do_mixing(m, nsp, yps, c->consider, xm, dq, rho, rad, d_eul, dzeit, mesa, sig);
    if (rc) return rc;
    double res = 0.0;
    for (int j = 0; j < m; ++j) res = fmax(res, orc_mass_residual(nsp, yps + (size_t) j * nsp));
    if (res >= SUM_TOL)
        for (int j = 0; j < m; ++j) orc_renorm_1d(nsp, yps + (size_t) j * nsp);
    for (size_t i = 0; i < (size_t) m * nsp; ++i) yps[i] = fmax(yps[i], 1.0e-99);
    return 0;
}
