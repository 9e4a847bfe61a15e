/* CPU oracle for the network-integration path.  TEST INFRASTRUCTURE ONLY: it is the checker that
 * the HIP path is compared against (tests/, the smoke check and the CPU-baseline leg of the
 * benchmark); no product code links or calls it.
 *
 * It restates, in plain C and in the reference's own operation order, the reference NuPPN
 * routines listed beside each declaration.  Species and reaction indices are 0-based here; the
 * counts k2, k4, k6, k8 are as in the reference.  All arrays are caller-owned.
 *
 * Linear solves use either the oracle's own left-looking GEPP (Gilbert-Peierls column solve with
 * SuperLU's threshold-1.0 / diagonal-preference pivot rule, SuperLU_5.0/SRC/dpivotL.c:114-156)
 * or, when the reference SuperLU compiled from source (_ref/libslu_ref.so) has been opened with
 * orc_ref_open, the reference entry points slu_solve / c_fortran_dgssv_ themselves.
 */
#ifndef NUPPN_ORACLE_H
#define NUPPN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* solver knobs, solver/source/solver_knobs.F90:36-81 (defaults in orc_default_knobs) */
typedef struct {
    int ittd;                 /* max Newton iterations (6) */
    int irdn;                 /* 2 */
    int max_sub_steps;        /* 4000 */
    int integration_method;   /* 0 backward Euler, 1 Bader-Deuflhard */
    int interpolate_temp;     /* 1 */
    int interpolate_dens;     /* 1 */
    double dgrd;              /* 1e-3 */
    double grdthreshold;      /* 1e-15 */
    double cyminformal;       /* 1e-30 (reference default; <1e-99 means static network) */
    double min_rate_ov_dt;    /* 1e-99, set to -1e99 when cyminformal < 1e-99 */
    double min_t9_for_slv;    /* 0.01 */
    double delta_state_max;   /* 1e99 */
} orc_knobs;

/* per-call diagnostics, solver/source/solver_diagnostics.F90:7-17 plus counters */
typedef struct {
    int nsubt;       /* sub-steps taken (index of the last sub-step) */
    int titers;      /* Newton iterations in the last backward-Euler step */
    int iter;        /* Newton iterations (BE) or sum of nseq (BD) of the last step */
    int nvar1;       /* size of the reduced system */
    int nsucc, nfail;
    long long n_newton;   /* Newton iterations over the whole call */
    long long n_factor;   /* LU factorisations over the whole call */
    long long n_solve;    /* triangular solves over the whole call */
} orc_stats;

typedef struct orc_ctx orc_ctx;
typedef struct orc_lu orc_lu;

/* linear solvers: the oracle GEPP; the reference SuperLU (when opened); ORC_LS_EXACT, the
 * reference algorithm with its rounding taken out where it is not the algorithm's: each Newton
 * system solved exactly (quad-precision residual refinement) and dX/dt and the Newton rhs
 * rounded once from double-double sums - a yardstick for the rounding error of FP64 paths */
enum { ORC_LS_GEPP = 0, ORC_LS_REF = 1, ORC_LS_EXACT = 2 };

void orc_default_knobs(orc_knobs *k);

/* k: [8*nre] laid out k[m*nre + l] for m = 0..7 (k1..k8); rtype: 0 two-body (x rho), 1
 * photodisintegration, 2 decay; coef: [7*nre] a0..a6; aw/zw: mass numbers and charges;
 * igamma: index of the gamma pseudo-species (not a considered isotope). */
orc_ctx *orc_create(int nsp, int nre, const int *k, const int *rtype, const double *coef,
                    const double *aw, const double *zw, int igamma, const orc_knobs *knobs);
void orc_destroy(orc_ctx *c);
void orc_set_knobs(orc_ctx *c, const orc_knobs *knobs);
void orc_set_linear_solver(orc_ctx *c, int which);
void orc_reset_state(orc_ctx *c);  /* SAVE'd state: frst, BD kopt / reduce */

/* makemap, solver/source/jac_rhs.F90:461-540: full pattern (row-major CSR, 0-based) and the
 * per-reaction slot map dns2csr_map(8, nre) -> map[m*nre + l]. */
int  orc_nnz_full(const orc_ctx *c);
void orc_full_pattern(const orc_ctx *c, int *rowptr, int *colind, int *map);

/* reaclib_calculate_rates, physics/source/reaclib.F90:431-582, with the clips of
 * evaluate_rates.F90:46,172 (rfac = 1) */
void orc_rates(const orc_ctx *c, double t9, double rho, double ye, double *v);

/* The REACLIB T9 cache (physics knob use_cache): reaclib_preprocessor (reaclib.F90:609-646)
 * tabulates the rates at rho = 1 on 5,001 points T9 = 0.01 .. 10, and orc_rates then interpolates
 * them linearly (reaclib.F90:446-470).  orc_rate_cache: the table [5001][nre] or NULL. */
void orc_set_rate_cache(orc_ctx *c, int use_cache);
const double *orc_rate_cache(const orc_ctx *c);

/* calculate_dxdt, solver/source/jac_rhs.F90:217-278 */
void orc_dxdt(const orc_ctx *c, const double *x, const double *v, double *dxdt);

/* calculate_jacobian, solver/source/jac_rhs.F90:282-455.  jac/row/col have orc_nnz_full
 * entries (row/col 0-based, -1 = unset).  Returns 0, or 1 on an off-diagonal exact zero (the
 * reference stops there). */
int orc_jacobian(const orc_ctx *c, double dt, const double *x, const double *v,
                 double *jac, int *row, int *col, unsigned char *mask);

/* reduce_jacobian, solver/source/jac_rhs.F90:134-197, applied to a scaled Jacobian (row is
 * modified as in the reference).  Output reduced CSR, 0-based, sorted columns.  Returns the
 * reduced nnz, or -1 on an exact zero in the reduced matrix (the reference stops). */
int orc_reduce_jacobian(const orc_ctx *c, double *jac, int *row, const int *col,
                        const unsigned char *mask, int *nvar1, int *rowptr, int *colind,
                        double *val);

/* rate vector used by orc_backward_euler / orc_bader_deuflhard (module array rates::v) */
void orc_set_rates(orc_ctx *c, const double *v);

/* Bader-Deuflhard state carried between calls: iv = frst, reduce, kopt, kmax; dv = errmax,
 * abuerr, negerr, sumerr, err(1:8) (bader_deuflhard.F90:29-38,152; solver.F90:75) */
void orc_bd_state(const orc_ctx *c, int *iv, double *dv);
void orc_bd_set_frst(orc_ctx *c, int frst);

/* One Newton iteration / the backward-Euler step / the BD step / the sub-step controller.
 * ierr codes as utils/source/errors.F90:5-16 (0 = success). */
int orc_backward_euler(orc_ctx *c, double *yps, double dt, double *dtrcmd, orc_stats *st);
int orc_bader_deuflhard(orc_ctx *c, double *yps, double dt, double *dtrcmd, orc_stats *st);
/* evaluate_all_rates at the cycle mid-point (ppn.F90:289-301, mppnp.F90:719-724) followed by
 * integrate_network (solver/source/solver.F90:53-238) */
int orc_integrate_zone(orc_ctx *c, double *yps, double t9_0, double t9_1, double rho_0,
                       double rho_1, double *ye, double dtwant, orc_stats *st);
/* the same with the mid-point rate vector supplied by the caller (v_mid [nre]) instead of
 * evaluated here: isolates the integrator from libm differences in the rate formula */
int orc_integrate_zone_v(orc_ctx *c, const double *v_mid, double *yps, double t9_0, double t9_1,
                         double rho_0, double rho_1, double *ye, double dtwant, orc_stats *st);

/* sortedsum (utils.F90:481-501, heapsort indexx of sort.F90:73-125) and calculate_ye
 * (utils.F90:537-553), ypsplus (utils.F90:726-740) */
double orc_sortedsum(int n, const double *a);
double orc_calculate_ye(const orc_ctx *c, const double *x);
void   orc_ypsplus(int n, double *y);

/* The mppnp cycle around the integration (frames/mppnp/source/mppnp.F90:637-781,
 * mixing.F90, utils.F90:387-469,670-684).  yps[m][nsp] zone-major. */
int    orc_zoneselect(double t9threshold, const double *t9, int m, unsigned char *sel);
double orc_ksum(int n, const double *a);
double orc_mass_residual(int n, const double *y);
void   orc_renorm_1d(int n, double *y);
void   orc_mix_d_lagrangian(int m, const double *rho, const double *rad, const double *d_eul,
                            double *d_lagr);
int    orc_matrixdiffus(int m, double h, const double *xm, double *D, double sig, double *a,
                        double *b, double *c);
void   orc_matrixdiffus_mesa(int m, double h, const double *dq, const double *d, double sig,
                             double *a, double *b, double *c);
int    orc_tridag(int n, const double *a, const double *b, const double *c, const double *r,
                  double *u);
int    orc_do_mixing(int m, int nsp, double *yps, const unsigned char *consider, const double *xm,
                     const double *dq, const double *rho, const double *rad, const double *d_eul,
                     double dzeit, int mesa, double sig);
int    orc_mppnp_cycle(orc_ctx *c, int m, double *yps, const double *t9, const double *rho,
                       const double *xm, const double *dq, const double *rad, const double *d_eul,
                       double dzeit, double t9threshold, int mesa, double sig, int *ierr,
                       orc_stats *stats);

/* Rate physics beyond the bare REACLIB sum (evaluate_rates.F90:14-172): partition functions on
 * REACLIB's T9 grid (pf_t9 [pf_n], pf_tab [nsp][pf_n]) and the factors of its reverse ('v') rates
 * (rev_ch [nre] chapter 1..11 or 0, rev_part [6][nre] participants e1..e6 or -1; reaclib.F90:
 * 512-557, 586-605), Chugunov screening (screen 1; screening_chugunov.F90), detailed balance
 * (reverse.F90:25-337; spin factors g_spin, nuclear masses mass_g in g, q_mev as ppn_q).
 * pf_n = 0 switches the partition functions and detailed balance off. */
void orc_set_rate_physics(orc_ctx *c, int pf_n, const double *pf_t9, const double *pf_tab,
                          const int *rev_ch, const int *rev_part, int screen,
                          int detailed_balance, const double *g_spin, const double *mass_g,
                          const double *q_mev);
void orc_reverse_pairs(const orc_ctx *c, int *irev);
/* evaluate_all_rates at abundances x (orc_rates without rate physics) */
void orc_rates_full(const orc_ctx *c, double t9, double rho, double ye, const double *x,
                    double *v);

/* Bader-Deuflhard tables, bader_deuflhard.F90:66-124 */
void orc_bd_tables(int *nseq9, double *a9, double *alf64, int *kopt, int *kmax);

/* GEPP on a CSC matrix (0-based).  perm_c in SuperLU's convention (perm_c[i] = position of
 * original column i); NULL = natural order.  Returns 0 or SuperLU's info (j+1 for an exactly
 * zero pivot in column j). */
int  orc_lu_factor(int n, const int *colptr, const int *rowind, const double *val,
                   const int *perm_c, orc_lu **out);
void orc_lu_solve(const orc_lu *lu, double *b);
void orc_lu_info(const orc_lu *lu, int *perm_r, int *nnz_l, int *nnz_u);
void orc_lu_free(orc_lu *lu);

/* Reference SuperLU (compiled from the reference sources) opened lazily from `path`. */
int  orc_ref_open(const char *path);
int  orc_ref_available(void);
void orc_ref_slu_solve(int n, int nnz, int isrow, int shiftindex, double *val, int *colind,
                       int *rowptr, double *rhs, int *ierr);
void orc_ref_dgssv(int iopt, int n, int nnz, int nrhs, double *values, int *rowind, int *colptr,
                   double *b, int ldb, long long *factors, int *info);
int  orc_ref_colperm(int n, int nnz, const int *colptr, const int *rowind, int *perm_colamd,
                     int *perm_final, int *etree);
int  orc_ref_factor_stats(long long handle, int n, int *perm_c, int *perm_r, int *nnz_l,
                          int *nnz_u, int *nsuper);
int  orc_ref_factor_supernodes(long long handle, int *xsup);

#ifdef __cplusplus
}
#endif
#endif
