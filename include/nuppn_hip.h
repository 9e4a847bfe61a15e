/* nuppn_hip.h - C ABI of the MI355X (gfx950) network-integration path.
 *
 * Plain pointers and sizes only.  Unless a name ends in _dev, array arguments are host memory
 * owned by the caller and are only read or written during the call.  All indices are 0-based
 * except in the two SuperLU entry points, which keep the reference's Fortran conventions.
 *
 * Every entry point replaces a reference interface, cited beside it (paths relative to the
 * NuPPN checkout).  Errors come back as codes, never as a process exit: the reference's
 * integration codes (utils/source/errors.F90:5-16) are positive, library errors negative.
 */
#ifndef NUPPN_HIP_H
#define NUPPN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* integration error codes, utils/source/errors.F90:5-16 */
#define NUPPN_OK            0
#define NUPPN_SLV_CNVRG     1
#define NUPPN_MAX_SUBSTEPS  2
#define NUPPN_BAD_STEP      3
#define NUPPN_NEG_ABUNDS    4
#define NUPPN_HUGE_ABUNDS   5
#define NUPPN_SUM_X         6
#define NUPPN_GOT_NAN       7
#define NUPPN_SMALL_DT      11
/* library errors */
#define NUPPN_ERR_ARG          (-1)   /* bad argument */
#define NUPPN_ERR_HIP          (-2)   /* HIP runtime error / no device / library not loadable */
#define NUPPN_ERR_PATTERN      (-3)   /* reserved (not returned): entries that reduce_jacobian
                                         drops (exactly 0 or -1) are dropped, and a zero where
                                         the reference stops ends the zone with NUPPN_BAD_STEP */
#define NUPPN_ERR_SINGULAR     (-4)   /* exactly zero pivot (slu_solve returns info > 0 there) */
#define NUPPN_ERR_UNSUPPORTED  (-5)   /* configuration the device path does not serve (irdn != 2,
                                         an integration_method other than 0 / 1, a path or
                                         network shape a call names but the context lacks) */
#define NUPPN_ERR_MIXING       (-6)   /* the mixing step stops where mixing.F90 does: a zero zone
                                         width or shell mass, or a zero pivot of tridag */

/* Reaction network (physics/source/rates.F90:13-16, ppn_physics.F90:746-815 conventions).
 * k[m*nre + l], m = 0..7 holds k1..k8 of reaction l: k1,k3,k5,k7 are 0-based species indices,
 * k2,k4,k6,k8 counts.  rtype: 0 two-body (rate x rho), 1 photodisintegration, 2 decay.
 * coef[m*nre + l] = REACLIB a0..a6.  aw/zw: mass number and charge of each species.  igamma:
 * the gamma pseudo-species (molar abundance 1, not a considered isotope). */
typedef struct {
    int nsp;
    int nre;
    const int *k;
    const int *rtype;
    const double *coef;
    const double *aw;
    const double *zw;
    int igamma;
} nuppn_network;

/* solver knobs, solver/source/solver_knobs.F90:36-81 (same names and meaning) */
typedef struct {
    int ittd;
    int irdn;
    int max_sub_steps;
    int integration_method;   /* 0 backward Euler, 1 Bader-Deuflhard */
    int interpolate_temp;
    int interpolate_dens;
    double dgrd;
    double grdthreshold;
    double cyminformal;       /* >= 0: the dynamic network (jac_rhs.F90:308-436), the mask and
                                 the reduced entries recomputed at every Jacobian evaluation;
                                 < 0: the static network (every species of every reaction) */
    double min_rate_ov_dt;    /* reactions with |v| / dt at or below it are left out of the
                                 Jacobian; forced to -1e99 for cyminformal < 1e-99
                                 (solver_knobs.F90:66-69) */
    double min_t9_for_slv;
    double delta_state_max;
} nuppn_knobs;

/* per-zone diagnostics (solver/source/solver_diagnostics.F90:7-17 plus work counters) */
typedef struct {
    int nsubt;
    int titers;
    int iter;
    int nvar1;
    int nsucc;
    int nfail;
    long long n_newton;
    long long n_factor;
    long long n_solve;
    long long lu_nnz;     /* sum of nnz(L) + nnz(U) over the zone's factorisations (plain path) */
    long long n_refine;   /* FP64 refinement residuals (batched path) */
    long long n_factor_fp64;  /* factorisations done with FP64 factors (batched path) */
} nuppn_zone_stats;

typedef struct nuppn_ctx nuppn_ctx;

/* readsolverinput defaults (solver_knobs.F90:43-57), except the static network (cyminformal = -1;
 * the namelist's default 1e-30 selects the dynamic one) */
void nuppn_default_knobs(nuppn_knobs *k);

/* Network set-up: makemap (solver/source/jac_rhs.F90:461-540), the static reduced pattern of
 * I - dt J, the inverse (pull) maps for dX/dt and J, the COLAMD column order with etree
 * post-order (SuperLU get_perm_c.c:405 + sp_preorder.c:140-150) and device upload.
 * Replaces solver_init + makemap + the network broadcast (jac_rhs.F90:565-580). */
nuppn_ctx *nuppn_ctx_create(const nuppn_network *net, const nuppn_knobs *knobs, int *status);
void nuppn_ctx_destroy(nuppn_ctx *ctx);

/* Rate physics beyond the bare REACLIB sum, as evaluate_all_rates applies it
 * (physics/source/evaluate_rates.F90:14-172; the reference's defaults detailed_balance = .true.,
 * screen_option = CHUGUNOV, physics_knobs.F90:110-111):
 *   - partition functions on REACLIB's T9 grid (pf_t9 [pf_n] ascending, pf_tab [nsp][pf_n]),
 *     interpolated linearly at the clipped temperature (reaclib.F90:586-605), multiply the REACLIB
 *     reverse ('v') rates by their chapter's Gfac (reaclib.F90:512-557): rev_ch [nre] is the
 *     chapter 1..11 of such a rate (0 otherwise), rev_part [6][nre] its participants e1..e6
 *     (species, -1 none);
 *   - detailed_balance: every forward reaction paired as reverse_init pairs it (reverse.F90:25-105)
 *     sets its partner's rate from its own (reverse.F90:107-337): g_spin [nsp] spin factors,
 *     mass_g [nsp] nuclear masses (g), q_mev [nsp] binding energies (MeV);
 *   - screen 1: Chugunov, DeWitt & Yakovlev (2007) screening of every reaction between charged
 *     reactants at the zone's abundances (screening_chugunov.F90:86-235).
 * The batch and zone entry points then evaluate the rates at each zone's abundances.  pf_n = 0
 * and screen = 0 (or a null pointer) switch it all off: the bare REACLIB sum. */
typedef struct {
    int pf_n;
    const double *pf_t9;
    const double *pf_tab;
    const int *rev_ch;
    const int *rev_part;
    int detailed_balance;
    const double *g_spin;
    const double *mass_g;
    const double *q_mev;
    int screen;
} nuppn_rate_physics;

int nuppn_set_rate_physics(nuppn_ctx *ctx, const nuppn_rate_physics *rp);
/* reverse_init's pairing in use: irev[nre] = the reverse partner of a forward reaction, -1 */
int nuppn_reverse_pairs(nuppn_ctx *ctx, int *irev);
/* evaluate_all_rates (evaluate_rates.F90:14-172) for nz states at abundances x [nz][nsp]:
 * v [nz][nre]; equals nuppn_rates_eval without rate physics */
int nuppn_evaluate_all_rates(nuppn_ctx *ctx, int nz, const double *t9, const double *rho,
                             const double *x, double *v);

/* Column order used by the factorisation, SuperLU convention (perm_c[i] = position of reduced
 * column i).  set: caller-supplied order, like SuperLU's MY_PERMC; get: the order in use. */
int nuppn_set_column_order(nuppn_ctx *ctx, const int *perm_c);
int nuppn_get_column_order(nuppn_ctx *ctx, int *perm_c);

/* Host-only (no GPU needed): the column order SuperLU's dgssv would use for an n x n CSC
 * pattern - get_perm_c(COLAMD) then sp_preorder's etree post-order (SuperLU_5.0/SRC/
 * get_perm_c.c:405, sp_preorder.c:140-150).  perm_c [n] (SuperLU convention), etree [n] or
 * NULL (post-ordered parents, root parent = n). */
int nuppn_column_order(int n, const int *colptr, const int *rowind, int *perm_c, int *etree);

/* Reduced system pattern (CSC of I - dt J over the masked species, reduce_jacobian +
 * csrcsc, jac_rhs.F90:134-213): sizes, then arrays of n+1 / nnz / n / nsp entries. */
int nuppn_reduced_size(nuppn_ctx *ctx, int *n, int *nnz);
int nuppn_reduced_pattern(nuppn_ctx *ctx, int *colptr, int *rowind, int *red2sp);

/* REACLIB rates for nz thermodynamic states (reaclib_calculate_rates, physics/source/
 * reaclib.F90:431-582, with evaluate_rates.F90:46,172 clips).  v: [nz*nre], zone-major. */
int nuppn_rates_eval(nuppn_ctx *ctx, int nz, const double *t9, const double *rho,
                     const double *ye, double *v);

/* The REACLIB T9 cache (physics knob use_cache, physics/source/physics_knobs.F90:34): on the
 * first switch-on, reaclib_preprocessor (reaclib.F90:609-646) tabulates every rate at rho = 1 on
 * 5,001 points T9 = 0.01 .. 10 on the device; while it is on, every rate evaluation of the
 * context (nuppn_rates_eval and the integration entry points) interpolates the table linearly
 * (reaclib.F90:446-470) instead of evaluating the formula.  use_cache 0 switches back. */
int nuppn_set_rate_cache(nuppn_ctx *ctx, int use_cache);
/* the table, [5001 * nre] (row i: T9 = 0.01 + i (10 - 0.01) / 5000); NUPPN_ERR_UNSUPPORTED if
 * it was never built */
int nuppn_rate_cache_table(nuppn_ctx *ctx, double *table);

/* dX/dt (calculate_dxdt, jac_rhs.F90:217-278) for one state; x, dxdt: [nsp], v: [nre]. */
int nuppn_dxdt(nuppn_ctx *ctx, const double *x, const double *v, double *dxdt);

/* Values of the reduced matrix I - dt J in the nuppn_reduced_pattern order
 * (calculate_jacobian + val*(-dt) + reduce_jacobian, backward_euler.F90:133-137). */
int nuppn_jacobian_reduced(nuppn_ctx *ctx, double dt, const double *x, const double *v,
                           double *csc_val);

/* GEPP of a reduced-pattern matrix on the device (SuperLU dgstrf pivot policy, dpivotL.c) and
 * solve of nrhs right-hand sides b [nrhs*n]; perm_r [n] receives the row pivots (SuperLU
 * convention), nnz_lu [2] the sizes of L (without unit diagonal) and U (with diagonal). */
int nuppn_lu_solve_reduced(nuppn_ctx *ctx, const double *csc_val, int nrhs, double *b,
                           int *perm_r, int *nnz_lu);

/* One zone: evaluate_all_rates at the cycle mid-point + integrate_network
 * (frames/ppn/source/ppn.F90:289-301, solver/source/solver.F90:53-238).  yps: [nsp] in/out,
 * ye in/out, *ierr: integration code.  Returns NUPPN_OK or a library error. */
int nuppn_integrate_zone(nuppn_ctx *ctx, double *yps, double t9_0, double t9_1, double rho_0,
                         double rho_1, double *ye, double dtwant, int *ierr,
                         nuppn_zone_stats *stats);

/* All zones of an mppnp cycle in one launch (replaces the worker loop of
 * frames/mppnp/source/mppnp.F90:713-753).  yps: [nz*nsp], zone-major (Fortran yps(nsp,nz));
 * t9_0/t9_1/rho_0/rho_1/dtwant/ye: [nz]; ierr: [nz]; stats: [nz] or NULL. */
int nuppn_integrate_batch(nuppn_ctx *ctx, int nz, double *yps, const double *t9_0,
                          const double *t9_1, const double *rho_0, const double *rho_1,
                          double *ye, const double *dtwant, int *ierr, nuppn_zone_stats *stats);

/* Same with device-resident arrays, enqueued on `stream` (a hipStream_t, NULL = default). */
int nuppn_integrate_batch_dev(nuppn_ctx *ctx, int nz, double *yps, const double *t9_0,
                              const double *t9_1, const double *rho_0, const double *rho_1,
                              double *ye, const double *dtwant, int *ierr,
                              nuppn_zone_stats *stats, void *stream);

/* Options of the batched path (nuppn_integrate_batch_mixed): the reduced system is factored
 * without pivoting in a block-tridiagonal + border form (molar units, columns scaled by the
 * diagonal), in FP32 (factor_precision 32) or FP64 (64).  Every solve is refined against the
 * FP64 matrix, x_{k+1} = x_k + (LU)^-1 (rhs - A x_k), the residual rhs - A x carried in twice the
 * FP64 precision and rounded once, so the refinement converges to the solution of the FP64
 * system itself whatever the factor precision.  It stops when
 *   - each correction satisfies |d_i| <= refine_tol * (|x_i| + |y_i| + 1e-30), y the current
 *     abundances, or
 *   - (accept_tol > 0 only) x is backward stable: max_i |rhs - A x|_i / (|A| |x| + |rhs| +
 *     1e-30)_i <= accept_tol.  Off by default: a backward-stable x can still be far from the
 *     solution in ill-conditioned components, differently for each factor precision.
 * When the corrections stop contracting, or max_refine steps pass, with FP32 factors, the zone
 * switches to FP64 factors for the rest of the call and the iteration is redone; with FP64
 * factors the last iterate is kept.
 * team_zones: a pass whose work list holds at most this many zones runs every zone on a
 * workgroup of several wavefronts (shared assembly and residuals, a twisted two-ended
 * factorisation: a shorter iteration), a longer list on one wavefront per zone.  -1 (default):
 * every pass; 0: never.
 * team_waves: the wavefronts of such a workgroup, 2 or 4; 0 (default): 4.
 * queue_passes: controller/worker passes the host queues before it reads the work-list counts
 * back (1..16, default 16; the counts stay on the device, so queued passes need no host round
 * trip).  While the list holds more zones than team_zones the host still reads back every pass.
 * use_graph: 1 (default): a chunk of 16 passes with a short list is captured once as a hipGraph
 * and replayed while the list stays within the grids it was captured with; 0: plain launches.
 * Every choice gives the same results bit for bit. */
typedef struct {
    int factor_precision;   /* 32 (default) or 64 */
    int max_refine;         /* refinement steps per solve (12) */
    double refine_tol;      /* 1e-17: below half an ulp of every new abundance */
    double accept_tol;      /* 0 (off) */
    int team_zones;         /* -1 (default), 0 or a zone count */
    int team_waves;         /* 0 (default), 2 or 4 */
    int queue_passes;       /* 16 (default): 1..16 */
    int use_graph;          /* 1 (default) or 0 */
} nuppn_batch_opts;

void nuppn_default_batch_opts(nuppn_batch_opts *o);

/* Autotuner of the batched path (no counterpart in the reference): times one cycle of the
 * sample batch (host arrays, as nuppn_integrate_batch_mixed) under each candidate shape - block
 * size 32 and 64 where the network's band allows it, and the multi-wavefront kernels for short
 * work lists (team_zones default / 0 / 2 x CUs with two wavefronts) - and keeps the fastest: the
 * context is rebuilt with that block size and its options become the defaults of later calls
 * with opts NULL.  cache_path (may be NULL): a text file of earlier choices keyed by network
 * (nsp, nre, nnz) and CU count; a hit applies it without timing, a new choice is appended.
 * Every candidate gives the same results bit for bit (res->consistent reports it). */
typedef struct {
    int block_size, team_zones, team_waves;   /* the choice */
    int from_cache;                           /* 1: read from cache_path, nothing timed */
    int ncand;                                /* candidates timed */
    int consistent;                           /* 1: all gave the same abundances bit for bit */
    int cand_nb[8], cand_tz[8], cand_tw[8];
    double ms[8];                             /* their cycle times */
} nuppn_tune_result;

int nuppn_autotune(nuppn_ctx *ctx, int nz, const double *yps, const double *t9_0,
                   const double *t9_1, const double *rho_0, const double *rho_1,
                   const double *dtwant, const char *cache_path, nuppn_tune_result *res);

/* Block structure of the batched path: info[0..6] = block size NB (0: unavailable; -1: the
 * band form, for interiors too wide for blocks, whose window is half-bandwidth + 1), blocks,
 * border species, interior species, interior half-bandwidth, resident slots, LDS bytes per
 * slot.  Returns NUPPN_OK or NUPPN_ERR_UNSUPPORTED. */
int nuppn_batch_block_info(nuppn_ctx *ctx, int *info, int ninfo);

/* Rebuilds the batched path in the given form: nb = 32 or 64 (block tridiagonal with that block
 * size) or -1 (the band form); the context otherwise takes the first of 32, 64, band that holds.
 * No counterpart in the reference (a shape choice, like nuppn_autotune's).  Returns NUPPN_OK,
 * NUPPN_ERR_UNSUPPORTED (the form does not hold for the network) or NUPPN_ERR_ARG. */
int nuppn_set_block_form(nuppn_ctx *ctx, int nb);

/* Batched mixed-precision path: all zones of an mppnp cycle (replaces the worker loop of
 * frames/mppnp/source/mppnp.F90:713-753).  Same control flow, decisions code and outputs as
 * nuppn_integrate_batch; the linear solves are the refined block factorisation above instead of
 * the reference's GEPP, so results agree with the plain path to within the FP64 rounding spread
 * of the solve, not bit for bit. */
int nuppn_integrate_batch_mixed(nuppn_ctx *ctx, int nz, double *yps, const double *t9_0,
                                const double *t9_1, const double *rho_0, const double *rho_1,
                                double *ye, const double *dtwant, int *ierr,
                                nuppn_zone_stats *stats, const nuppn_batch_opts *opts);
int nuppn_integrate_batch_mixed_dev(nuppn_ctx *ctx, int nz, double *yps, const double *t9_0,
                                    const double *t9_1, const double *rho_0, const double *rho_1,
                                    double *ye, const double *dtwant, int *ierr,
                                    nuppn_zone_stats *stats, const nuppn_batch_opts *opts,
                                    void *stream);

/* One mppnp cycle (frames/mppnp/source/mppnp.F90:637-781) on m >= 2 zones of a stellar model
 * that share this network, zone-major yps[m][nsp] (Fortran yps(msl, nsp) holds it the other way:
 * pass the transpose):
 *   - zoneselect (637, 1080-1097): the zones with t9 > t9threshold are integrated, all in one batch
 *     (path 1: the batched path with bopts, as nuppn_integrate_batch_mixed; path 0: the plain FP64
 *     path), at constant T9 and rho over dzeit, Ye from the zone's abundances (719-724);
 *   - per integrated zone, renorm_1d when |1 - sum X| > 5e-7 (737-738); ierr[z] and stats[z]
 *     (zeros for the zones not selected).  Any non-zero ierr ends the call here (the reference
 *     stops, 726-735): the zones keep their integrated abundances and mixing is not applied;
 *   - do_mixing (765; mixing.F90:30-176 with idifmix 1): D_lagr from the Eulerian coefficient
 *     d_eul, radius rad (solar units) and density rho, the implicit diffusion matrix over the mass
 *     coordinates xm (code_source 0: GNV / EVL, matrixdiffus) or the shell masses dq (1: MES,
 *     matrixdiffus_mesa) with coefficients limited at sig_term_limit, one tridiagonal solve per
 *     considered species 1..nsp-1 (the last species is not mixed, as in mixdiff08);
 *   - renorm_2d of every zone when the largest |1 - sum X| reaches 1e-6, then max(yps, 1e-99)
 *     (768-781).
 * Returns NUPPN_OK, NUPPN_ERR_MIXING, NUPPN_ERR_UNSUPPORTED (path 1 without a batched path) or a
 * library error.  Replaces the master/worker cycle of mppnp.F90:653-781. */
typedef struct {
    double t9threshold;      /* zoneselect's threshold (ppn_frame.input t9threshold) */
    double sig_term_limit;   /* 1e10 (ppn_frame.input) */
    int code_source;         /* 0: GNV / EVL (mass coordinates), 1: MES (shell masses) */
    int path;                /* 1 (default): batched path; 0: plain FP64 path */
} nuppn_mppnp_opts;

void nuppn_default_mppnp_opts(nuppn_mppnp_opts *o);
int nuppn_mppnp_cycle(nuppn_ctx *ctx, int m, double *yps, const double *t9, const double *rho,
                      const double *xm, const double *dq, const double *rad, const double *d_eul,
                      double dzeit, const nuppn_mppnp_opts *opts, const nuppn_batch_opts *bopts,
                      int *ierr, nuppn_zone_stats *stats);
/* Same with device-resident arrays, enqueued on `stream` (a hipStream_t, NULL = default). */
int nuppn_mppnp_cycle_dev(nuppn_ctx *ctx, int m, double *yps, const double *t9, const double *rho,
                          const double *xm, const double *dq, const double *rad,
                          const double *d_eul, double dzeit, const nuppn_mppnp_opts *opts,
                          const nuppn_batch_opts *bopts, int *ierr, nuppn_zone_stats *stats,
                          void *stream);

/* Timing of the last batch call from HIP events on its stream: out[0] rate kernel ms,
 * out[1] whole integration ms, out[2] summed Newton-kernel ms, out[3] summed control-kernel
 * ms, out[4] Newton-kernel launches, out[5] controller passes, out[6] launches of captured
 * pass graphs.  nout: entries wanted (<= 7). */
int nuppn_last_timing(nuppn_ctx *ctx, double *out, int nout);

/* Link-level drop-ins with the reference signatures and semantics.
 * slu_solve: solver/source/slu_f90/slu_f90.c:3-54 (Fortran interface slu_f90/slu.f90:6-11);
 *   colind/rowptr are shifted to 0-based in place when shiftindex != 0, as the reference does.
 *   Its device buffers (pattern, values, dense-capacity factors) are kept per GPU from call to
 *   call and reallocated only for a larger system (held until the process ends); calls are
 *   serialised by a process-wide lock.
 * c_fortran_dgssv_: solver/source/SuperLU_5.0/FORTRAN/c_fortran_dgssv.c:35-190
 *   (iopt 1 factor -> handle, 2 solve, 3 free; 1-based CSC). */
void slu_solve(int n, int nnz, int isrow, int shiftindex, double val[], int colind[],
               int rowptr[], double rhs[], int *ierr);
void c_fortran_dgssv_(int *iopt, int *n, int *nnz, int *nrhs, double *values, int *rowind,
                      int *colptr, double *b, int *ldb, long long *f_factors, int *info);

/* slu_solve without device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, index shift, results and error codes as slu_solve;
 * every device buffer of the call lives in one fixed 48 MiB array per GPU (held until the process
 * ends), which bounds the system to n up to about 2,000 (dense-capacity factors); larger systems
 * return *ierr = NUPPN_ERR_UNSUPPORTED without touching the caller's arrays.  Calls are serialised
 * by a process-wide lock (they share the array); host-side ordering work uses host memory. */
void slu_solve_noalloc(int n, int nnz, int isrow, int shiftindex, double val[], int colind[],
                       int rowptr[], double rhs[], int *ierr);

#ifdef __cplusplus
}
#endif
#endif
