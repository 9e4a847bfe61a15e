// Batched path: the Newton iteration of the backward-Euler step (newton_iteration,
// solver/source/backward_euler.F90:120-176) with the pivot-free block factorisation of
// block_lu.hpp and FP64 residual refinement, for every zone of a cycle that needs one.
//
// It plugs into the same per-zone state machine as the plain path (zone.hpp: zone_init /
// zone_control run integrate_network's sub-step and Newton decisions in FP64), so both paths
// take their decisions with one implementation; only the linear solve differs:
//     x_0 = 0, r_0 = rhs;  d_k = (LU)^-1 r_k (factors in precision T),  x_{k+1} = x_k + d_k,
//     r_{k+1} = rhs - A x_{k+1} (the FP64 assembled matrix; products and sums in twice the FP64
//     precision, rounded once: block_residual)
// stopping when every correction is small against its own species, |d_i| <= refine_tol *
// (|x_i| + |y_i| + 1e-30) with y the current abundances (a normwise test would leave the
// abundances of rare species, 20 orders of magnitude below the dominant ones, unconverged; the
// 1e-30 floor ignores species far below any reported abundance).  With the extra-precise
// residual the iteration converges to the solution of the FP64 system itself, whatever the
// precision of the factors, as long as they contract it: FP32-factor and FP64-factor zones then
// take the same Newton iterates to within a few units of the last place.  (An FP64 residual
// would stop it at cond(A) eps(FP64) of the solution, different for every factorisation.)
// When the corrections stop shrinking (a contraction worse than 1/2 per step) or the step limit
// passes with FP32 factors, cond(A) * eps(FP32) is not small: the zone is marked for FP64
// factors (for the rest of the call) and the same wavefront redoes the iteration with them.
// With FP64 factors the last iterate is kept.
#pragma once
#include "bd.hpp"
#include "block_lu.hpp"
#include "band_lu.hpp"
#include "zone.hpp"

namespace nuppn {

struct BatchOpts {
    int fp64_factors;   // 1: FP64 factors from the start (no FP32 stage)
    int max_refine;     // refinement steps per solve
    double refine_tol;  // stop when every |d_i| <= refine_tol * (|x_i| + |y_i| + 1e-30)
    double accept_tol;  // > 0: stop also when max_i |r_i| / (|A| |x| + |rhs|)_i <= accept_tol
};

// per-slot scratch of the block Newton kernels (one slot per resident wavefront)
// (B_JAC .. B_SORT: the Bader-Deuflhard step's J, tableau and work vectors)
// (B_LO: the low parts of the extra-precise residual; B_XLO: those of the solution; B_APHYS:
// the assembled matrix in the residual stream's physical order)
// (B_FIRST, B_RFLAG, B_EFLAG: the dynamic network's marks, dynamic.hpp)
enum BSlotArr { B_XN, B_VX, B_ST, B_AVAL, B_RHS, B_X, B_RES, B_D, B_JAC, B_YK, B_YEXT, B_YSCAL,
                B_DELTA, B_QCOL, B_YERR, B_SORT, B_LO, B_XLO, B_APHYS, B_FIRST, B_RFLAG, B_EFLAG,
                B_COUNT };

// the factors, in the same slot after the arrays above
// (F_DIN .. F_BIN: the dense blocks D_i, C_i, B_i of the scaled matrix, the factorisation's
// inputs, one set per factor precision (FP64, then FP32); only their pattern positions are ever
// written, so the zeros elsewhere persist from the slots' allocation)
enum BFacArr { F_CS, F_COLSC, F_ROWINV, F_SINV, F_W, F_BM, F_E, F_XF, F_G, F_GPIV, F_DIN, F_CIN,
               F_BIN, F_DIN32, F_CIN32, F_BIN32, F_COUNT };

struct BLayout {
    size_t off[B_COUNT];
    size_t foff[F_COUNT];
    size_t bytes;   // one slot: arrays, then factors
};

// band_l / band_u > 0: the band form (nblk = 0), its L and U bands of that many entries in the
// areas of Sinv and W
inline BLayout block_layout(int nsp, int nre, int n, int nnzA, int NB, int nblk, int nip,
                            int nbd, int nres_phys, size_t band_l = 0, size_t band_u = 0) {
    const size_t bs = (size_t) nblk * NB * NB * 8;   // FP64-sized: FP32 factors use half
    size_t pow2 = 1;
    while (pow2 < (size_t) nsp) pow2 <<= 1;
    const size_t len[B_COUNT] = {8ull * nsp, 8ull * nre, 16ull * nre, 8ull * nnzA, 8ull * n,
                                 8ull * n, 8ull * n, 8ull * n, 8ull * nnzA,
                                 8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * n, 64ull * nsp,
                                 64ull * nsp, 8ull * pow2, 8ull * n, 8ull * n,
                                 8ull * nres_phys, 4ull * nsp, (size_t) nre, (size_t) nnzA};
    const size_t flen[F_COUNT] = {8ull * n, 8ull * (nip + nbd), 8ull * (nip + nbd),
                                  band_l ? 8 * band_l : bs, band_u ? 8 * band_u : bs, bs,
                                  8ull * nbd * nip + 8, 8ull * nbd * nip + 8, 8ull * nbd * nbd + 8,
                                  4ull * nbd + 4, bs, bs, bs, bs / 2, bs / 2, bs / 2};
    BLayout L;
    size_t o = 0;
    for (int i = 0; i < B_COUNT; ++i) {
        L.off[i] = o;
        o += (len[i] + 255) & ~(size_t) 255;
    }
    for (int i = 0; i < F_COUNT; ++i) {
        L.foff[i] = o;
        o += (flen[i] + 255) & ~(size_t) 255;
    }
    L.bytes = o;
    return L;
}

struct BKArgs {
    DevBlock B;
    DevBand D;   // the band form's structure (ni = 0 for the block form)
    BatchOpts O;
    char *slot_base;
    size_t slot_bytes;
    size_t off[B_COUNT];
    size_t foff[F_COUNT];
};

template <class T>
__device__ inline BlockFactors<T> zone_factors(char *base, const size_t *foff) {
    BlockFactors<T> F;
    F.Sinv = (T *) (base + foff[F_SINV]);
    F.W = (T *) (base + foff[F_W]);
    F.Bm = (T *) (base + foff[F_BM]);
    F.E = (T *) (base + foff[F_E]);
    F.XF = (T *) (base + foff[F_XF]);
    F.G = (double *) (base + foff[F_G]);
    F.gpiv = (int *) (base + foff[F_GPIV]);
    F.cs = (double *) (base + foff[F_CS]);
    F.colsc = (double *) (base + foff[F_COLSC]);
    F.rowinv = (double *) (base + foff[F_ROWINV]);
    const bool f64 = sizeof(T) == 8;
    F.Din = (T *) (base + foff[f64 ? F_DIN : F_DIN32]);
    F.Cin = (T *) (base + foff[f64 ? F_CIN : F_CIN32]);
    F.Bin = (T *) (base + foff[f64 ? F_BIN : F_BIN32]);
    return F;
}

// the band form's factors in the same factor area (the block form's areas reused: Sinv -> L's
// band, W -> U's band, E -> L's border rows, XF -> U's border columns)
template <class T>
__device__ inline BandFactors<T> band_factors(char *base, const size_t *foff) {
    BandFactors<T> F;
    F.Lb = (T *) (base + foff[F_SINV]);
    F.Uc = (T *) (base + foff[F_W]);
    F.LE = (T *) (base + foff[F_E]);
    F.UF = (T *) (base + foff[F_XF]);
    F.G = (double *) (base + foff[F_G]);
    F.gpiv = (int *) (base + foff[F_GPIV]);
    F.cs = (double *) (base + foff[F_CS]);
    F.colsc = (double *) (base + foff[F_COLSC]);
    F.rowinv = (double *) (base + foff[F_ROWINV]);
    return F;
}

// Factors of the assembled matrix (A in CSR order) in precision T into the factor area fac.
// Returns 0, or 1 when a pivot or the border Schur complement is not finite or zero.
template <class T, int NB, class TM = Team<1>>
__device__ __forceinline__ int block_factor_zone(const BKArgs &b, const double *Aval, char *fac,
                                 NUPPN_LDS char *lds, ZoneStats &st) {
    const BlockFactors<T> F = zone_factors<T>(fac, b.foff);
    const BlockLds<T> L = make_block_lds<T, NB>(lds, b.B.nip);
    const int fail = block_factor<T, NB, TM>(b.B, Aval, F, L);
    st.n_factor++;
    if (sizeof(T) == 8) st.n_factor_fp64++;
    return fail;
}

// Refined solve of A x = rhs with the precision-T factors (fac: the zone's slot; A is read there
// in the residual stream's order, B_APHYS).  y: current abundances (species order).  The solution is kept as x + xlo, an unevaluated sum carried in
// twice the FP64 precision, so that y + x is accurate relative to itself even for a species
// that the step burns down by many orders of magnitude (y + x << |x|).  The corrections stop
// when every |d_i| <= refine_tol (|y_i + x_i| + kRefineFloor): each new abundance has converged
// relative to its own size.  Returns 0 (x, xlo hold the solution), 1 (FP32 factors cannot
// deliver it: the caller switches to FP64 factors) or -4 (non-finite with FP64 factors: the
// reference's singular-matrix outcome).
template <bool FP64, class TM, class S>
__device__ __forceinline__ int refine_loop(const DevNet &N, const BKArgs &b, const double *rhs,
                                           const double *y, double *x, double *xlo, double *res,
                                           double *d, char *fac, NUPPN_LDS double *red,
                                           ZoneStats &st, S solve) {
    const DevBlock &B = b.B;
    const int n = B.n;
    NUPPN_GLB double *xg = gptr(x), *xl = gptr(xlo), *resg = gptr(res);
    const NUPPN_GLB double *dg = gptr((const double *) d), *yg = gptr(y), *rhsg = gptr(rhs);
    const NUPPN_GLB int *red2sp = gptr(N.red2sp);
    const int t0 = TM::tid();
    batched_for<4>(
        t0, n, TM::NT, [&](int i) { return rhsg[i]; },
        [&](int i, double r) {
            xg[i] = 0.0;
            xl[i] = 0.0;
            resg[i] = r;
        });
    TM::sync();
    double eta_prev = 0.0;
    for (int k = 0; k < b.O.max_refine; ++k) {
        solve(res, d);
        st.n_solve++;
        double eta = 0.0;
        bool fin = true;
        struct Uq {
            double di, xh, xlo, y;
        };
        batched_for<4>(
            t0, n, TM::NT, [&](int i) { return Uq{dg[i], xg[i], xl[i], yg[red2sp[i]]}; },
            [&](int i, const Uq &q) {
                const double di = q.di, xh = q.xh;
                // (xh + xlo) + d: TwoSum, the rounding error into the low part
                const double sh = xh + di;
                const double bb = sh - xh;
                const double lo = q.xlo + ((xh - (sh - bb)) + (di - bb));
                const double hi = sh + lo;
                xg[i] = hi;
                xl[i] = lo - (hi - sh);
                const double scale = fabs(q.y + hi) + kRefineFloor;
                eta = fmax(eta, fabs(di) / scale);
                fin = fin && isfinite(hi);
            });
        eta = team_max<TM>(eta, red);
        if (team_any<TM>(!fin, red)) return FP64 ? -4 : 1;
        if (eta <= b.O.refine_tol) return 0;   // converged corrections
        // the next residual, extra-precise (d, the correction just applied, serves as its
        // scratch); with accept_tol > 0, x is also accepted once it is backward stable: it
        // solves a system within accept_tol, entry by entry, of A and rhs
        const double omega = block_residual<TM>(B, (const double *) (fac + b.off[B_APHYS]),
                                                rhs, x, xlo, res, d,
                                                (double *) (fac + b.off[B_LO]), red);
        st.n_refine++;
        if (b.O.accept_tol > 0.0 && omega <= b.O.accept_tol) return 0;
        // otherwise the corrections must keep contracting (the first correction's eta is
        // about 1 by construction, so contraction is judged from the second one on).  A
        // stall or the step limit with FP32 factors means cond(A) * eps(FP32) is not small
        // and the FP64 factors take over; with FP64 factors the last iterate is the result.
        if ((k > 1 && eta > 0.5 * eta_prev) || k == b.O.max_refine - 1) return FP64 ? 0 : 1;
        eta_prev = eta;
    }
    return 0;
}

template <class T, int NB, class TM = Team<1>>
__device__ __forceinline__ int block_refine(const DevNet &N, const BKArgs &b,
                            const double *rhs, const double *y, double *x, double *xlo,
                            double *res, double *d, char *fac, NUPPN_LDS char *lds,
                            ZoneStats &st) {
    const BlockFactors<T> F = zone_factors<T>(fac, b.foff);
    const BlockLds<T> L = make_block_lds<T, NB>(lds, b.B.nip);
    return refine_loop<sizeof(T) == 8, TM>(
        N, b, rhs, y, x, xlo, res, d, fac, L.red, st,
        [&](const double *r, double *dd) __attribute__((always_inline)) {
            block_solve<T, NB, TM>(b.B, F, L, r, dd);
        });
}


// Refined solve with fresh factors (the backward-Euler Newton iteration: one solve per
// factorisation).  Same return codes as block_refine.
template <class T, int NB, class TM>
__device__ __forceinline__ int block_refined_solve(const DevNet &N, const BKArgs &b, const double *Aval,
                                   const double *rhs, const double *y, double *x, double *xlo,
                                   double *res, double *d, char *fac, NUPPN_LDS char *lds,
                                   ZoneStats &st) {
    if (block_factor_zone<T, NB, TM>(b, Aval, fac, lds, st)) return sizeof(T) == 8 ? -4 : 1;
    return block_refine<T, NB, TM>(N, b, rhs, y, x, xlo, res, d, fac, lds, st);
}

// dX/dt of the reduced species from the per-reaction fluxes vx, each species' sum carried in
// twice the FP64 precision (products split exactly, Dot2 sums) and left as hi (out) + lo (lo).
// Every reaction moves integer multiples of its flux between species whose mass numbers weight
// them to zero, so these sums conserve mass to the rounding of each species' net rate, where
// an FP64 sum loses the rounding of its gross fluxes (forward and reverse rates of nuclei near
// equilibrium exceed the net by many orders): the mass drift of a hot zone over a cycle falls
// from ~1e-11 to the last bits.
template <class TM = Team<1>>
__device__ __forceinline__ void block_dxdt_dd(const DevBlock &B, const double *vxp, double *outp,
                                              double *lop) {
    const DevSeg &D = B.dxdt;
    const NUPPN_GLB double *coef = gptr(D.e_coef), *vxg = gptr(vxp);
    NUPPN_GLB double *o = gptr(outp), *lo = gptr(lop);
    seg_stream<SumDD2, TM>(
        D,
        [&](int e, int idx) {
            const double c = coef[e], f = vxg[idx];
            const double t = c * f;
            return SumDD2{t, fma(c, f, -t)};
        },
        [&](int r, const SumDD2 &t) {
            o[r] = t.hi;
            lo[r] = t.lo;
        });
}

// The Jacobian entries (CSR order) from the per-reaction stems: every entry is the sum of its
// contributions cnt stem a_row / a_col (jac_rhs.F90:365-392) plus the 1e-99 seed of
// calculate_jacobian (jac_rhs.F90:309).  All contributions of an entry share the ratio
// a_row / a_col, so the stream carries the counts (packed in the operand words: bits 0..19 the
// stem, bits 20..23 the count + 8), the sum of cnt stem is carried in double-double (each product
// split exactly by an FMA, Dot2 sums), and the entry is that sum times the ratio (a double-double
// from the host) plus the seed, rounded once.  jlo: [nnzA] scratch.
template <class TM = Team<1>>
__device__ __forceinline__ void block_jacobian_dd(const DevBlock &B, const double *stemp,
                                                  double *jacp, double *jlop) {
    const DevSeg &J = B.jac;
    const NUPPN_GLB double *st = gptr(stemp);
    const NUPPN_GLB double *rhi = gptr(B.jac_rhi), *rlo = gptr(B.jac_rlo);
    NUPPN_GLB double *jac = gptr(jacp), *jlo = gptr(jlop);
    seg_stream<SumDD2, TM>(
        J,
        [&](int e, int idx) {
            (void) e;
            const double s = st[idx & 0xFFFFF];
            const double c = (double) (((idx >> 20) & 0xF) - 8);
            const double t = s * c;
            return SumDD2{t, fma(s, c, -t)};
        },
        [&](int k, const SumDD2 &t) {
            jac[k] = t.hi;
            jlo[k] = t.lo;
        });
    struct Jq {
        double sh, sl, rh, rl;
    };
    batched_for<8>(
        TM::tid(), B.nnzA, TM::NT, [&](int k) { return Jq{jac[k], jlo[k], rhi[k], rlo[k]}; },
        [&](int k, const Jq &q) {
            // (rh + rl) (sh + sl) + seed: the leading product split exactly, the cross terms
            // and the seed's TwoSum error added to it, rounded once
            const double p = q.rh * q.sh;
            const double pe = fma(q.rh, q.sh, -p) + (q.rh * q.sl + q.rl * q.sh);
            const double seed = 1.0e-99;
            const double s = p + seed, bb = s - p;
            jac[k] = s + (((p - (s - bb)) + (seed - bb)) + pe);
        });
    TM::sync();
}

// the batched path's assembly: fluxes and stems per reaction, then I - dt J (CSR order, with
// reduce_jacobian's drops) and the Newton rhs y_old - y_cur + dt dX/dt as streamed sums over
// precomputed coefficient lists (J, dX/dt and the rhs each rounded once from double-double
// sums).  lo: [n], jlo: [nnzA] scratch.  Returns 1 where the reference stops (reduced_value).
template <class TM>
__device__ __forceinline__ int block_assemble(const DevNet &N, const DevKnobs &K,
                                      const DevBlock &B, double dt,
                                      const double *y_old, const double *y_cur, const double *v,
                                      double *xn, double *vx, double *stem, double *Aval,
                                      double *rhs, double *lop, double *jlo, double *aphys,
                                      const DynWork &DW, int &nvar1, NUPPN_LDS double *red) {
    const int t0 = TM::tid();
    molar(N, y_cur, xn, t0, TM::NT);
    TM::sync();
    reaction_terms(N, xn, v, vx, stem, stem + N.nre, t0, TM::NT);
    TM::sync();
    nvar1 = N.n;
    if (K.dynamic)
        nvar1 = dyn_network<TM>(N, K, dt, xn, v, stem, stem + N.nre, B.dyn_k, B.nnzA, DW, red);
    // the streamed sums store raw totals; the scalings follow in coalesced passes
    const NUPPN_GLB unsigned char *isdiag = gptr(B.csr_isdiag);
    NUPPN_GLB double *Ag = gptr(Aval);
    block_jacobian_dd<TM>(B, stem, Aval, jlo);
    int stop = 0;
    const NUPPN_GLB int *rpos = gptr(B.resid_pos);
    NUPPN_GLB double *Ap = gptr(aphys);
    if (K.dynamic) {
        // the reduced matrix of the dynamic network embedded in the static one (dynamic.hpp)
        const NUPPN_GLB int *crow = gptr(B.csr_row), *r2s = gptr(N.red2sp);
        const NUPPN_GLB unsigned char *eflag = gptr(DW.eflag);
        for (int k = t0; k < B.nnzA; k += TM::NT) {
            const bool diag = isdiag[k];
            double val;
            if (!dyn_marked(DW, r2s[crow[k]]))
                val = diag ? 1.0 : 0.0;
            else if (!diag && !eflag[k])
                val = 0.0;
            else
                val = reduced_value(Ag[k], dt, diag, stop);
            Ag[k] = val;
            Ap[rpos[k]] = val;
        }
    } else {
        struct Rq {
            double a;
            int pos;
            bool diag;
        };
        batched_for<8>(
            t0, B.nnzA, TM::NT, [&](int k) { return Rq{Ag[k], rpos[k], isdiag[k] != 0}; },
            [&](int k, const Rq &q) {
                const double val = reduced_value(q.a, dt, q.diag, stop);
                Ag[k] = val;
                Ap[q.pos] = val;
            });
    }
    stop = team_any<TM>(stop != 0, red) ? 1 : 0;
    const NUPPN_GLB int *red2sp = gptr(N.red2sp);
    const NUPPN_GLB double *yo = gptr(y_old), *yc = gptr(y_cur), *lo = gptr((const double *) lop);
    NUPPN_GLB double *rg = gptr(rhs);
    block_dxdt_dd<TM>(B, vx, rhs, lop);
    struct Hq {
        double a, b, fh, fl;
        int sp;
    };
    batched_for<4>(
        t0, N.n, TM::NT,
        [&](int r) {
            const int sp = red2sp[r];
            return Hq{yo[sp], -yc[sp], rg[r], lo[r], sp};
        },
        [&](int r, const Hq &q) {
            // y_old - y_cur + dt (hi + lo): each step split exactly (TwoSum, FMA), rounded once
            const double a = q.a, b = q.b, fh = q.fh;
            const double s1 = a + b, b1 = s1 - a;
            const double e1 = (a - (s1 - b1)) + (b - b1);
            const double p = dt * fh;
            const double e2 = fma(dt, fh, -p) + dt * q.fl;
            const double s2 = s1 + p, b2 = s2 - s1;
            const double e3 = (s1 - (s2 - b2)) + (p - b2);
            // reduce_rhs packs the masked species only: the others solve to an exact zero
            rg[r] = (!K.dynamic || dyn_marked(DW, q.sp)) ? s2 + (e1 + (e2 + e3)) : 0.0;
        });
    TM::sync();
    return stop;
}

// One Newton iteration (newton_iteration, backward_euler.F90:120-176) for zone z in PH_NEWTON:
// the solve with FP32 factors while the zone's factor precision (state pad0) is FP32, and with
// FP64 factors when that fails (the zone then stays on FP64 for the rest of the call) or when
// the zone is on FP64 already.  Writes y_new and grd (or lib_err) and sets AFTER_NEWTON.
// The factorisation forms of the batched path: the block-tridiagonal one (block size NB) and
// the band one (band_lu.hpp), each a refined solve with fresh factors and its LDS reduction slots.
template <int NB>
struct BlockForm {
    template <class T, class TM>
    __device__ __forceinline__ static int refined_solve(const DevNet &N, const BKArgs &b, const double *Aval,
                                        const double *rhs, const double *y, double *x,
                                        double *xlo, double *res, double *d, char *fac,
                                        NUPPN_LDS char *lds, ZoneStats &st) {
        return block_refined_solve<T, NB, TM>(N, b, Aval, rhs, y, x, xlo, res, d, fac, lds, st);
    }
    __device__ __forceinline__ static NUPPN_LDS double *red(const BKArgs &b, NUPPN_LDS char *lds) {
        return make_block_lds<double, NB>(lds, b.B.nip).red;
    }
};

struct BandForm {
    template <class T, class TM>
    __device__ __forceinline__ static int refined_solve(const DevNet &N, const BKArgs &b, const double *Aval,
                                        const double *rhs, const double *y, double *x,
                                        double *xlo, double *res, double *d, char *fac,
                                        NUPPN_LDS char *lds, ZoneStats &st) {
        const BandFactors<T> F = band_factors<T>(fac, b.foff);
        const BandLds<T> L = make_band_lds<T>(lds, b.B.n, b.D.W);
        st.n_factor++;
        if (sizeof(T) == 8) st.n_factor_fp64++;
        if (band_factor<T>(b.B, b.D, Aval, F, L)) return sizeof(T) == 8 ? -4 : 1;
        return refine_loop<sizeof(T) == 8, TM>(
            N, b, rhs, y, x, xlo, res, d, fac, L.red, st,
            [&](const double *r, double *dd) __attribute__((always_inline)) {
                band_solve<T>(b.B, b.D, F, L, r, dd);
            });
    }
    __device__ __forceinline__ static NUPPN_LDS double *red(const BKArgs &b, NUPPN_LDS char *lds) {
        return make_band_lds<double>(lds, b.B.n, b.D.W).red;
    }
};

template <class FORM, class TM>
__device__ __forceinline__ void zone_newton_form(const KArgs &a, const BKArgs &b, char *slot,
                                                 NUPPN_LDS char *lds, int z) {
    const DevNet &N = a.N;
    const DevKnobs &K = a.K;
    const ZoneData &Z = a.Z;
    ZoneState s = Z.state[z];
    if (s.phase != PH_NEWTON) return;
    const double *y_old = zrow(Z.y_old, Z.nsp_stride, z);
    const double *y_cur = zrow(Z.y_cur, Z.nsp_stride, z);
    double *y_new = zrow(Z.y_new, Z.nsp_stride, z);
    const double *v = zrow(Z.v, Z.nre_stride, z);
    double *xn = (double *) (slot + b.off[B_XN]);
    double *vx = (double *) (slot + b.off[B_VX]);
    double *stem = (double *) (slot + b.off[B_ST]);
    double *Aval = (double *) (slot + b.off[B_AVAL]);
    double *rhs = (double *) (slot + b.off[B_RHS]);
    double *x = (double *) (slot + b.off[B_X]);
    double *xlo = (double *) (slot + b.off[B_XLO]);
    double *res = (double *) (slot + b.off[B_RES]);
    double *d = (double *) (slot + b.off[B_D]);
    const double dt = s.dttry;
    NUPPN_LDS double *red = FORM::red(b, lds);
    const int t0 = TM::tid();
    const DynWork DW{(int *) (slot + b.off[B_FIRST]), (unsigned char *) (slot + b.off[B_RFLAG]),
                     (unsigned char *) (slot + b.off[B_EFLAG])};
    const int stop = block_assemble<TM>(N, K, b.B, dt, y_old, y_cur, v, xn, vx, stem, Aval, rhs,
                                        (double *) (slot + b.off[B_LO]),
                                        (double *) (slot + b.off[B_JAC]),
                                        (double *) (slot + b.off[B_APHYS]), DW, s.st.nvar1, red);
    int rc = stop ? 2 : 1;
    if (!stop && s.pad0 == 0) {
        rc = FORM::template refined_solve<float, TM>(N, b, Aval, rhs, y_cur, x, xlo, res, d,
                                                     slot, lds, s.st);
        if (rc == 1) s.pad0 = 1;
    }
    if (rc == 1)
        rc = FORM::template refined_solve<double, TM>(N, b, Aval, rhs, y_cur, x, xlo, res, d,
                                                      slot, lds, s.st);
    if (rc == 0) {
        // y_new = y_cur + (x + xlo), rounded once
        struct Yq {
            double yc, xh, xl;
            int sp;
        };
        batched_for<4>(
            t0, N.n, TM::NT,
            [&](int r) {
                const int sp = N.red2sp[r];
                return Yq{y_cur[sp], x[r], xlo[r], sp};
            },
            [&](int r, const Yq &q) {
                (void) r;
                const double sh = q.yc + q.xh;
                const double bb = sh - q.yc;
                y_new[q.sp] = sh + (((q.yc - (sh - bb)) + (q.xh - bb)) + q.xl);
            });
        TM::sync();
        double g = 0.0;
        for (int i = t0; i < N.nsp; i += TM::NT) {
            const double yc = y_cur[i];
            const double dd = y_new[i] - yc;
            const double ga = (yc > K.grdthreshold) ? fabs(dd / yc) : 0.0;
            g = fmax(g, ga);
        }
        s.grd = team_max<TM>(g, red);
    }
    s.lib_err = stop ? 0 : rc;
    s.step_ierr = stop ? E_BAD_STEP : 0;
    s.phase = PH_AFTER_NEWTON;
    if (t0 == 0) Z.state[z] = s;
    TM::sync();
}

template <int NB, class TM>
__device__ __forceinline__ void zone_newton_block(const KArgs &a, const BKArgs &b, char *slot,
                                                  NUPPN_LDS char *lds, int z) {
    zone_newton_form<BlockForm<NB>, TM>(a, b, slot, lds, z);
}

// The batched path's solver for the Bader-Deuflhard step: J once per step by the chunked sums
// (CSR order), I - h J per level factored pivot-free in the zone's precision, every stage solve
// refined in FP64 (FP32 factors that cannot deliver a solve switch the zone to FP64 factors,
// which redo that solve).
template <int NB>
struct BlockBDSolver {
    const KArgs &a;
    const BKArgs &b;
    char *slot;
    NUPPN_LDS char *lds;
    ZoneState &s;
    const double *v;
    const double *ycur;   // abundances of the current stage (the correction test's scale)
    int nvar1;            // the static system's size (the dynamic network is not served here)

    __device__ double *arr(int i) const { return (double *) (slot + b.off[i]); }

    __device__ void fluxes(const double *y) {
        const DevNet &N = a.N;
        molar(N, y, arr(B_XN));
        wave_sync();
        reaction_terms(N, arr(B_XN), v, arr(B_VX), arr(B_ST), arr(B_ST) + N.nre);
        wave_sync();
    }

    __device__ int jacobian(const double *y, double dt) {
        (void) dt;   // min_rate_ov_dt is -1e99 on the static network
        const DevBlock &B = b.B;
        fluxes(y);
        block_jacobian_dd(B, arr(B_ST), arr(B_JAC), arr(B_AVAL));   // A is set per level
        const NUPPN_GLB unsigned char *isdiag = gptr(B.csr_isdiag);
        const NUPPN_GLB double *jac = gptr((const double *) arr(B_JAC));
        int bad = 0;
        for (int k = lane_id(); k < B.nnzA; k += WAVE)
            if (!isdiag[k] && jac[k] == 0.0) bad = 1;   // calculate_jacobian stops
        wave_sync();
        return ballot(bad) ? E_BAD_STEP : 0;
    }

    __device__ int level(double h) {
        const DevBlock &B = b.B;
        const double *jac = arr(B_JAC);
        double *Aval = arr(B_AVAL);
        double *Ap = arr(B_APHYS);
        int stop = 0;
        for (int k = lane_id(); k < B.nnzA; k += WAVE) {
            const double val = reduced_value(jac[k], h, B.csr_isdiag[k], stop);
            Aval[k] = val;
            Ap[B.resid_pos[k]] = val;
        }
        wave_sync();
        if (ballot(stop)) return E_BAD_STEP;
        if (s.pad0 == 0 && block_factor_zone<float, NB>(b, Aval, slot, lds, s.st)) s.pad0 = 1;
        if (s.pad0 == 1 && block_factor_zone<double, NB>(b, Aval, slot, lds, s.st)) return -4;
        return 0;
    }

    // dX/dt summed in twice the FP64 precision (block_dxdt_dd), rounded once
    __device__ void dxdt(const double *y, double *out) {
        fluxes(y);
        block_dxdt_dd(b.B, arr(B_VX), out, arr(B_LO));
        const double *lo = arr(B_LO);
        for (int r = lane_id(); r < b.B.n; r += WAVE) out[r] = out[r] + lo[r];
        wave_sync();
    }

    // x = the refined solution, rounded to FP64 (the stage updates are FP64, as the reference's)
    __device__ int solve(const double *rhs, double *x) {
        const DevNet &N = a.N;
        const double *Aval = arr(B_AVAL);
        int rc = 1;
        if (s.pad0 == 0) {
            rc = block_refine<float, NB>(N, b, rhs, ycur, x, arr(B_XLO), arr(B_RES),
                                         arr(B_D), slot, lds, s.st);
            if (rc == 1) {
                s.pad0 = 1;
                if (block_factor_zone<double, NB>(b, Aval, slot, lds, s.st)) return -4;
            }
        }
        // a non-finite FP64 solution is returned as it is: the step then ends in GOT_NAN, as
        // the reference's does when its solve overflows
        if (rc == 1)
            (void) block_refine<double, NB>(N, b, rhs, ycur, x, arr(B_XLO), arr(B_RES),
                                            arr(B_D), slot, lds, s.st);
        else if (rc != 0)
            return rc;
        const double *xl = arr(B_XLO);
        for (int r = lane_id(); r < N.n; r += WAVE) x[r] = x[r] + xl[r];
        wave_sync();
        return 0;
    }
};

template <int NB>
__device__ void zone_bd_block(const KArgs &a, const BKArgs &b, char *slot, NUPPN_LDS char *lds,
                              int z) {
    const ZoneData &Z = a.Z;
    ZoneState s = Z.state[z];
    if (s.phase != PH_NEWTON) return;
    BDState bd = Z.bd[z];
    auto arr = [&](int i) { return (double *) (slot + b.off[i]); };
    BlockBDSolver<NB> sv{a, b, slot, lds, s, zrow(Z.v, Z.nre_stride, z), arr(B_YK), a.N.n};
    BDWork W{arr(B_YK), arr(B_YEXT), arr(B_YSCAL), arr(B_DELTA), arr(B_RHS), arr(B_X),
             arr(B_QCOL), arr(B_YERR), arr(B_SORT)};
    bd_step(a, z, s, bd, W, sv);
    s.phase = PH_AFTER_NEWTON;
    if (lane_id() == 0) {
        Z.state[z] = s;
        Z.bd[z] = bd;
    }
    wave_sync();
}

// every zone in PH_NEWTON (the control pass's work list); each wavefront owns one scratch slot
// (arrays and factors) and strides over the list.  W = resident wavefronts per SIMD the registers are held to: W = 2
// doubles the zones in flight when many need an iteration; W = 1 leaves the compiler the full
// register file and the shortest iteration, for passes with few zones.
template <int NB, int W>
__global__ void __launch_bounds__(64, W) zone_newton_block_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x)
        zone_newton_block<NB, Team<1>>(a, b, slot, lds, a.Z.active[i]);
}

// The same iteration with NW = 2 or 4 wavefronts per zone (one workgroup per zone; with the
// full register file one wavefront per SIMD, so 4 / NW zones per CU): the assembly streams, the
// residuals, the dense-block set-up, the border columns and every element-wise pass are shared,
// and the block factorisation and solves run twisted, top-down on wavefront 0 and bottom-up on
// wavefront 1.  For passes whose list is short, where one zone's iteration latency sets the pass
// time.  MINB = 2 holds the registers to two such workgroups per CU (long lists: more zones in
// flight, each iteration a little slower).
template <int NB, int NW, int MINB = 1>
__global__ void __launch_bounds__(64 * NW, MINB) zone_newton_team_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x)
        zone_newton_block<NB, Team<NW>>(a, b, slot, lds, a.Z.active[i]);
}

// The band form's iteration: one workgroup of four wavefronts per zone (the factorisation's
// 16 x 16 thread grid), the zones of the list strided over the grid.
__global__ void __launch_bounds__(256, 1) zone_newton_band_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x)
        zone_newton_form<BandForm, BandTeam>(a, b, slot, lds, a.Z.active[i]);
}

// every zone in PH_NEWTON under integration_method 1: one Bader-Deuflhard step each (a kernel
// of its own: the step's code would otherwise raise the Newton kernel's register use)
template <int NB>
__global__ void __launch_bounds__(64) zone_bd_block_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x)
        zone_bd_block<NB>(a, b, slot, lds, a.Z.active[i]);
}

}  // namespace nuppn
