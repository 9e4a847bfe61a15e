// Batched path: the Newton iteration of the backward-Euler step (newton_iteration,
// solver/source/backward_euler.F90:120-176) with the pivot-free block factorisation of
// block_lu.hpp and FP64 residual refinement, for every zone of a cycle that needs one.
//
// It plugs into the same per-zone state machine as the plain path (zone.hpp: zone_init /
// zone_control run integrate_network's sub-step and Newton decisions in FP64), so both paths
// take their decisions with one implementation; only the linear solve differs:
//     x_0 = 0, r_0 = rhs;  d_k = (LU)^-1 r_k (factors in precision T),  x_{k+1} = x_k + d_k,
//     r_{k+1} = rhs - A x_{k+1} (FP64, the assembled matrix)
// stopping when every correction is small against its own species, |d_i| <= refine_tol *
// (|x_i| + |y_i| + 1e-30) with y the current abundances (a normwise test would leave the
// abundances of rare species, 20 orders of magnitude below the dominant ones, unconverged; the
// 1e-30 floor ignores species far below any reported abundance), or when x is backward stable
// entry by entry, max_i |r_i| / (|A| |x| + |rhs|)_i <= accept_tol (Oettli-Prager; the
// correction test alone cannot stop on components whose condition number is near 1/eps(FP64),
// where the FP64 residual's own rounding keeps moving them).  The same stopping rule holds for
// both factor precisions, so FP32-factor and FP64-factor zones converge to the same solution.
// When the corrections stop shrinking (a contraction worse than 1/2 per step) or the step limit
// passes with FP32 factors, cond(A) * eps(FP32) is not small: the zone is marked for FP64
// factors (for the rest of the call) and left in the Newton phase, and the FP64 kernel,
// launched after the FP32 one in the same pass, redoes the iteration.  With FP64 factors the
// last iterate is kept.
#pragma once
#include "bd.hpp"
#include "block_lu.hpp"
#include "zone.hpp"

namespace nuppn {

struct BatchOpts {
    int fp64_factors;   // 1: FP64 factors from the start (no FP32 stage)
    int max_refine;     // refinement steps per solve
    double refine_tol;  // stop when every |d_i| <= refine_tol * (|x_i| + |y_i| + 1e-30)
    double accept_tol;  // stop when max_i |r_i| / (|A| |x| + |rhs|)_i <= accept_tol
};

// per-slot scratch of the block Newton kernels (one slot per resident wavefront)
// (B_JAC .. B_SORT: the Bader-Deuflhard step's J, tableau and work vectors)
enum BSlotArr { B_XN, B_VX, B_ST, B_AVAL, B_RHS, B_X, B_RES, B_D, B_JAC, B_YK, B_YEXT, B_YSCAL,
                B_DELTA, B_QCOL, B_YERR, B_SORT, B_COUNT };

// the factors, in the same slot after the arrays above
enum BFacArr { F_CS, F_COLSC, F_ROWINV, F_SINV, F_W, F_BM, F_E, F_XF, F_G, F_GPIV, F_COUNT };

struct BLayout {
    size_t off[B_COUNT];
    size_t foff[F_COUNT];
    size_t bytes;   // one slot: arrays, then factors
};

inline BLayout block_layout(int nsp, int nre, int n, int nnzA, int NB, int nblk, int nip,
                            int nbd) {
    const size_t bs = (size_t) nblk * NB * NB * 8;   // FP64-sized: FP32 factors use half
    size_t pow2 = 1;
    while (pow2 < (size_t) nsp) pow2 <<= 1;
    const size_t len[B_COUNT] = {8ull * nsp, 8ull * nre, 16ull * nre, 8ull * nnzA, 8ull * n,
                                 8ull * n, 8ull * n, 8ull * n, 8ull * nnzA,
                                 8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * n, 64ull * nsp,
                                 64ull * nsp, 8ull * pow2};
    const size_t flen[F_COUNT] = {8ull * n, 8ull * (nip + nbd), 8ull * (nip + nbd), bs, bs, bs,
                                  8ull * nbd * nip + 8, 8ull * nbd * nip + 8, 8ull * nbd * nbd + 8,
                                  4ull * nbd + 4};
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
    return F;
}

// Factors of the assembled matrix (A in CSR order) in precision T into the factor area fac.
// Returns 0, or 1 when a pivot or the border Schur complement is not finite or zero.
template <class T, int NB>
__device__ __forceinline__ int block_factor_zone(const BKArgs &b, const double *Aval, char *fac,
                                 NUPPN_LDS char *lds, ZoneStats &st) {
    const BlockFactors<T> F = zone_factors<T>(fac, b.foff);
    const BlockLds<T> L = make_block_lds<T, NB>(lds, b.B.nip);
    const int fail = block_factor<T, NB>(b.B, Aval, F, L);
    st.n_factor++;
    if (sizeof(T) == 8) st.n_factor_fp64++;
    return fail;
}

// Refined solve of A x = rhs with the precision-T factors in fac.  y: current abundances
// (species order) for the correction test.  Returns 0 (x holds the solution), 1 (FP32 factors
// cannot deliver it: the caller switches to FP64 factors) or -4 (non-finite with FP64
// factors: the reference's singular-matrix outcome).
template <class T, int NB>
__device__ __forceinline__ int block_refine(const DevNet &N, const BKArgs &b, const double *Aval,
                            const double *rhs, const double *y, double *x, double *res,
                            double *d, char *fac, NUPPN_LDS char *lds,
                            ZoneStats &st) {
    constexpr bool fp64 = sizeof(T) == 8;
    const DevBlock &B = b.B;
    const int n = B.n;
    const BlockFactors<T> F = zone_factors<T>(fac, b.foff);
    const BlockLds<T> L = make_block_lds<T, NB>(lds, B.nip);
    NUPPN_GLB double *xg = gptr(x), *resg = gptr(res);
    const NUPPN_GLB double *dg = gptr((const double *) d), *yg = gptr(y), *rhsg = gptr(rhs);
    const NUPPN_GLB int *red2sp = gptr(N.red2sp);
    for (int i = lane_id(); i < n; i += WAVE) {
        xg[i] = 0.0;
        resg[i] = rhsg[i];
    }
    wave_sync();
    double eta_prev = 0.0;
    for (int k = 0; k < b.O.max_refine; ++k) {
        block_solve<T, NB>(B, F, L, res, d);
        st.n_solve++;
        double eta = 0.0;
        bool fin = true;
        for (int i = lane_id(); i < n; i += WAVE) {
            const double di = dg[i];
            const double xi = xg[i] + di;
            xg[i] = xi;
            const double scale = fabs(xi) + fabs(yg[red2sp[i]]) + kAbundanceFloor;
            eta = fmax(eta, fabs(di) / scale);
            fin = fin && isfinite(xi);
        }
        eta = wave_max(eta);
        wave_sync();
        if (ballot(!fin)) return fp64 ? -4 : 1;
        if (eta <= b.O.refine_tol) return 0;   // converged corrections
        // x is accepted once it is backward stable in FP64: it solves a system within
        // accept_tol, entry by entry, of A and rhs
        // (d, the correction just applied, serves as the residual's scratch)
        const double omega = block_residual(B, Aval, rhs, x, res, d);
        st.n_refine++;
        if (omega <= b.O.accept_tol) return 0;
        // otherwise the corrections must keep contracting (the first correction's eta is
        // about 1 by construction, so contraction is judged from the second one on).  A
        // stall or the step limit with FP32 factors means cond(A) * eps(FP32) is not small
        // and the FP64 factors take over; with FP64 factors the last iterate is the result.
        if ((k > 1 && eta > 0.5 * eta_prev) || k == b.O.max_refine - 1) return fp64 ? 0 : 1;
        eta_prev = eta;
    }
    return 0;
}


// Refined solve with fresh factors (the backward-Euler Newton iteration: one solve per
// factorisation).  Same return codes as block_refine.
template <class T, int NB>
__device__ __forceinline__ int block_refined_solve(const DevNet &N, const BKArgs &b, const double *Aval,
                                   const double *rhs, const double *y, double *x, double *res,
                                   double *d, char *fac, NUPPN_LDS char *lds,
                                   ZoneStats &st) {
    if (block_factor_zone<T, NB>(b, Aval, fac, lds, st)) return sizeof(T) == 8 ? -4 : 1;
    return block_refine<T, NB>(N, b, Aval, rhs, y, x, res, d, fac, lds, st);
}

// the batched path's assembly: fluxes and stems per reaction, then I - dt J (CSR order) and the
// Newton rhs y_old - y_cur + dt dX/dt as chunked sums over precomputed coefficient lists
__device__ __forceinline__ void block_assemble(const DevNet &N, const DevBlock &B, double dt,
                                      const double *y_old, const double *y_cur, const double *v,
                                      double *xn, double *vx, double *stem, double *Aval,
                                      double *rhs) {
    molar(N, y_cur, xn);
    wave_sync();
    reaction_terms(N, xn, v, vx, stem, stem + N.nre);
    wave_sync();
    // the streamed sums store raw totals; the scalings follow in coalesced passes
    const DevSeg &J = B.jac;
    const NUPPN_GLB unsigned char *isdiag = gptr(B.csr_isdiag);
    NUPPN_GLB double *Ag = gptr(Aval);
    {
        const NUPPN_GLB double *coef = gptr(J.e_coef), *st = gptr(stem);
        seg_stream<Sum1>(
            J, [&](int e, int idx) { return Sum1{coef[e] * st[idx]}; },
            [&](int k, const Sum1 &t) { Ag[k] = t.s; });
    }
    for (int k = lane_id(); k < B.nnzA; k += WAVE) {
        const double jv = 1.0e-99 + Ag[k];
        double val = jv * (-dt);
        if (isdiag[k]) val = val + 1.0;
        Ag[k] = val;
    }
    const DevSeg &D = B.dxdt;
    const NUPPN_GLB int *red2sp = gptr(N.red2sp);
    const NUPPN_GLB double *yo = gptr(y_old), *yc = gptr(y_cur);
    NUPPN_GLB double *rg = gptr(rhs);
    {
        const NUPPN_GLB double *coef = gptr(D.e_coef), *vxg = gptr(vx);
        seg_stream<Sum1>(
            D, [&](int e, int idx) { return Sum1{coef[e] * vxg[idx]}; },
            [&](int r, const Sum1 &t) { rg[r] = t.s; });
    }
    for (int r = lane_id(); r < N.n; r += WAVE) {
        const int sp = red2sp[r];
        rg[r] = yo[sp] - yc[sp] + dt * rg[r];
    }
    wave_sync();
}

// One Newton iteration (newton_iteration, backward_euler.F90:120-176) for zone z in PH_NEWTON:
// the solve with FP32 factors while the zone's factor precision (state pad0) is FP32, and with
// FP64 factors when that fails (the zone then stays on FP64 for the rest of the call) or when
// the zone is on FP64 already.  Writes y_new and grd (or lib_err) and sets AFTER_NEWTON.
template <int NB>
__device__ __forceinline__ void zone_newton_block(const KArgs &a, const BKArgs &b, char *slot,
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
    double *res = (double *) (slot + b.off[B_RES]);
    double *d = (double *) (slot + b.off[B_D]);
    const double dt = s.dttry;
    block_assemble(N, b.B, dt, y_old, y_cur, v, xn, vx, stem, Aval, rhs);
    int rc = 1;
    if (s.pad0 == 0) {
        rc = block_refined_solve<float, NB>(N, b, Aval, rhs, y_cur, x, res, d, slot, lds,
                                            s.st);
        if (rc == 1) s.pad0 = 1;
    }
    if (rc == 1)
        rc = block_refined_solve<double, NB>(N, b, Aval, rhs, y_cur, x, res, d, slot, lds,
                                             s.st);
    if (rc == 0) {
        for (int r = lane_id(); r < N.n; r += WAVE) {
            const int sp = N.red2sp[r];
            y_new[sp] = y_cur[sp] + x[r];
        }
        wave_sync();
        double g = 0.0;
        for (int i = lane_id(); i < N.nsp; i += WAVE) {
            const double yc = y_cur[i];
            const double dd = y_new[i] - yc;
            const double ga = (yc > K.grdthreshold) ? fabs(dd / yc) : 0.0;
            g = fmax(g, ga);
        }
        s.grd = wave_max(g);
    }
    s.lib_err = rc;
    s.phase = PH_AFTER_NEWTON;
    if (lane_id() == 0) Z.state[z] = s;
    wave_sync();
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
        const DevSeg &J = B.jac;
        const NUPPN_GLB double *coef = gptr(J.e_coef), *stem = gptr((const double *) arr(B_ST));
        const NUPPN_GLB unsigned char *isdiag = gptr(B.csr_isdiag);
        NUPPN_GLB double *jac = gptr(arr(B_JAC));
        seg_stream<Sum1>(
            J, [&](int e, int idx) { return Sum1{coef[e] * stem[idx]}; },
            [&](int k, const Sum1 &t) { jac[k] = t.s; });
        int bad = 0;
        for (int k = lane_id(); k < B.nnzA; k += WAVE) {
            const double jv = 1.0e-99 + jac[k];
            jac[k] = jv;
            if (!isdiag[k] && jv == 0.0) bad = 1;   // calculate_jacobian stops
        }
        wave_sync();
        return ballot(bad) ? E_BAD_STEP : 0;
    }

    __device__ int level(double h) {
        const DevBlock &B = b.B;
        const double *jac = arr(B_JAC);
        double *Aval = arr(B_AVAL);
        for (int k = lane_id(); k < B.nnzA; k += WAVE) {
            double val = jac[k] * (-h);
            if (B.csr_isdiag[k]) val = val + 1.0;
            Aval[k] = val;
        }
        wave_sync();
        if (s.pad0 == 0 && block_factor_zone<float, NB>(b, Aval, slot, lds, s.st)) s.pad0 = 1;
        if (s.pad0 == 1 && block_factor_zone<double, NB>(b, Aval, slot, lds, s.st)) return -4;
        return 0;
    }

    __device__ void dxdt(const double *y, double *out) {
        const DevNet &N = a.N;
        fluxes(y);
        const DevSeg &D = b.B.dxdt;
        const NUPPN_GLB double *coef = gptr(D.e_coef), *vx = gptr((const double *) arr(B_VX));
        NUPPN_GLB double *o = gptr(out);
        (void) N;
        seg_stream<Sum1>(
            D, [&](int e, int idx) { return Sum1{coef[e] * vx[idx]}; },
            [&](int r, const Sum1 &t) { o[r] = t.s; });
        wave_sync();
    }

    __device__ int solve(const double *rhs, double *x) {
        const DevNet &N = a.N;
        const double *Aval = arr(B_AVAL);
        if (s.pad0 == 0) {
            const int rc = block_refine<float, NB>(N, b, Aval, rhs, ycur, x, arr(B_RES), arr(B_D),
                                                   slot, lds, s.st);
            if (rc != 1) return rc;
            s.pad0 = 1;
            if (block_factor_zone<double, NB>(b, Aval, slot, lds, s.st)) return -4;
        }
        // a non-finite FP64 solution is returned as it is: the step then ends in GOT_NAN, as
        // the reference's does when its solve overflows
        (void) block_refine<double, NB>(N, b, Aval, rhs, ycur, x, arr(B_RES), arr(B_D),
                                        slot, lds, s.st);
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
    BlockBDSolver<NB> sv{a, b, slot, lds, s, zrow(Z.v, Z.nre_stride, z), arr(B_YK)};
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
    for (int i = blockIdx.x; i < a.Z.nact; i += gridDim.x)
        zone_newton_block<NB>(a, b, slot, lds, a.Z.active[i]);
}

// every zone in PH_NEWTON under integration_method 1: one Bader-Deuflhard step each (a kernel
// of its own: the step's code would otherwise raise the Newton kernel's register use)
template <int NB>
__global__ void __launch_bounds__(64) zone_bd_block_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    for (int i = blockIdx.x; i < a.Z.nact; i += gridDim.x)
        zone_bd_block<NB>(a, b, slot, lds, a.Z.active[i]);
}

}  // namespace nuppn
