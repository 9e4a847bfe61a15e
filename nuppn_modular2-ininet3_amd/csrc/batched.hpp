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
// 1e-30 floor ignores species far below any reported abundance).
// When the corrections stop shrinking (a contraction worse than 1/2 per step) with FP32 factors,
// the solve is accepted if an earlier step contracted strongly (the stall is then the FP64
// residual's rounding floor); otherwise cond(A) * eps(T) is not small and the zone is marked for
// FP64 factors (for the rest of the call) and left in the Newton phase: the FP64 kernel,
// launched after the FP32 one in the same pass, redoes the iteration.  FP64 factors give the direct solve plus one
// refinement step, accepted as the reference accepts its single GEPP solve.
#pragma once
#include "block_lu.hpp"
#include "zone.hpp"

namespace nuppn {

struct BatchOpts {
    int fp64_factors;   // 1: FP64 factors from the start (no FP32 stage)
    int max_refine;     // refinement steps per solve with FP32 factors
    double refine_tol;  // stop when every |d_i| <= refine_tol * (|x_i| + |y_i| + 1e-30)
    double accept_tol;  // a stalled FP32 refinement is accepted below this (same measure)
};

// per-slot scratch of the block Newton kernels (one slot per resident wavefront)
enum BSlotArr { B_XN, B_VX, B_ST, B_AVAL, B_RHS, B_X, B_RES, B_D, B_PART, B_COUNT };

// per-zone factor storage: kept between Newton iterations, so an iteration can reuse the
// previous iteration's factors as its preconditioner
enum BFacArr { F_CS, F_SINV, F_W, F_BM, F_E, F_XF, F_G, F_GPIV, F_COUNT };

struct BLayout {
    size_t off[B_COUNT];
    size_t bytes;
    size_t foff[F_COUNT];
    size_t fbytes;
};

inline BLayout block_layout(int nsp, int nre, int n, int nnzA, int NB, int nblk, int nip,
                            int nbd, int npart) {
    const size_t bs = (size_t) nblk * NB * NB * 8;   // FP64-sized: FP32 factors use half
    const size_t len[B_COUNT] = {8ull * nsp, 8ull * nre, 16ull * nre, 8ull * nnzA, 8ull * n,
                                 8ull * n, 8ull * n, 8ull * n, 8ull * npart};
    const size_t flen[F_COUNT] = {8ull * n, bs, bs, bs, 8ull * nbd * nip + 8,
                                  8ull * nbd * nip + 8, 8ull * nbd * nbd + 8, 4ull * nbd + 4};
    BLayout L;
    size_t o = 0;
    for (int i = 0; i < B_COUNT; ++i) {
        L.off[i] = o;
        o += (len[i] + 255) & ~(size_t) 255;
    }
    L.bytes = o;
    o = 0;
    for (int i = 0; i < F_COUNT; ++i) {
        L.foff[i] = o;
        o += (flen[i] + 255) & ~(size_t) 255;
    }
    L.fbytes = o;
    return L;
}

struct BKArgs {
    DevBlock B;
    BatchOpts O;
    char *slot_base;
    size_t slot_bytes;
    size_t off[B_COUNT];
    char *fac_base;      // [nz] factor areas
    size_t fac_bytes;
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
    return F;
}

// Refined solve of the assembled system (A in CSR order) with factors in precision T.  fstate
// holds the zone's factor state (0 none, 1 FP32, 2 FP64): factors of the previous Newton
// iteration are reused as the preconditioner while they contract well (first correction below
// reuse_tol), and refreshed otherwise.  y: current abundances (species order) for the
// componentwise test.  Returns 0 (x holds the solution), 1 when FP32 factors cannot deliver the
// FP64 solution (the zone moves to FP64), or -4 when an FP64 factorisation fails.
template <class T, int NB>
__device__ int block_refined_solve(const DevNet &N, const BKArgs &b, const double *Aval,
                                   const double *rhs, const double *y, double *x, double *res,
                                   double *d, double *part, char *fac, NUPPN_LDS char *lds,
                                   int &fstate, ZoneStats &st) {
    constexpr bool fp64 = sizeof(T) == 8;
    constexpr int mine = fp64 ? 2 : 1;
    constexpr double reuse_tol = 0.05;
    const DevBlock &B = b.B;
    const int n = B.n;
    const BlockFactors<T> F = zone_factors<T>(fac, b.foff);
    const BlockLds<T> L = make_block_lds<T, NB>(lds, B.nip);
    bool fresh = false;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (fstate != mine) {
            const int fail = block_factor<T, NB>(B, Aval, F, L);
            st.n_factor++;
            if (fp64) st.n_factor_fp64++;
            if (fail) {
                fstate = 0;
                return fp64 ? -4 : 1;
            }
            fstate = mine;
            fresh = true;
        }
        for (int i = lane_id(); i < n; i += WAVE) {
            x[i] = 0.0;
            res[i] = rhs[i];
        }
        wave_sync();
        const int steps = (fp64 && fresh) ? 2 : b.O.max_refine;
        double eta_prev = 0.0, best_ratio = 1.0;
        bool refresh = false;
        for (int k = 0; k < steps; ++k) {
            block_solve<T, NB>(B, F, L, res, d);
            st.n_solve++;
            double eta = 0.0;
            bool fin = true;
            for (int i = lane_id(); i < n; i += WAVE) {
                const double xi = x[i] + d[i];
                x[i] = xi;
                const double scale = fabs(xi) + fabs(y[N.red2sp[i]]) + 1.0e-30;
                eta = fmax(eta, fabs(d[i]) / scale);
                fin = fin && isfinite(xi);
            }
            eta = wave_max(eta);
            wave_sync();
            if (ballot(!fin)) {
                if (!fresh) { refresh = true; break; }
                fstate = 0;
                return fp64 ? -4 : 1;
            }
            if (eta <= b.O.refine_tol || (fp64 && fresh && k == steps - 1)) return 0;
            // stale factors that no longer contract well are refreshed
            if (!fresh && k == 1 && eta > reuse_tol) { refresh = true; break; }
            // contraction from the second correction on (the first solve's eta is 1 by
            // construction: x starts at zero).  A stall after a strong contraction is the FP64
            // residual's own rounding floor and is accepted; a stall without one means
            // cond(A) * eps(T) is not small, and FP32 factors cannot deliver the FP64 solution.
            if (k > 1) {
                const double ratio = eta / eta_prev;
                if (ratio > 0.5) {
                    if (best_ratio <= 0.1) return 0;
                    if (!fresh) { refresh = true; break; }
                    fstate = 0;
                    return fp64 ? 0 : 1;
                }
                best_ratio = fmin(best_ratio, ratio);
            }
            if (k == steps - 1) {
                if (eta <= b.O.accept_tol) return 0;
                if (!fresh) { refresh = true; break; }
                fstate = 0;
                return fp64 ? 0 : 1;
            }
            eta_prev = eta;
            block_residual(B, Aval, rhs, x, res, part);
            st.n_refine++;
        }
        if (!refresh) return 0;
        fstate = 0;   // refactor with the current matrix and solve again
    }
    return 0;
}

// the batched path's assembly: fluxes and stems per reaction, then I - dt J (CSR order) and the
// Newton rhs y_old - y_cur + dt dX/dt as chunked sums over precomputed coefficient lists
__device__ inline void block_assemble(const DevNet &N, const DevBlock &B, double dt,
                                      const double *y_old, const double *y_cur, const double *v,
                                      double *xn, double *vx, double *stem, double *Aval,
                                      double *rhs, double *part) {
    molar(N, y_cur, xn);
    wave_sync();
    reaction_terms(N, xn, v, vx, stem, stem + N.nre);
    wave_sync();
    const DevSeg &J = B.jac;
    seg_chunks(J, part, [&](int e) { return J.e_coef[e] * stem[J.e_idx[e]]; });
    for (int k = lane_id(); k < B.nnzA; k += WAVE) {
        const double jv = 1.0e-99 + seg_total(J, part, k);
        double val = jv * (-dt);
        if (B.csr_isdiag[k]) val = val + 1.0;
        Aval[k] = val;
    }
    wave_sync();
    const DevSeg &D = B.dxdt;
    seg_chunks(D, part, [&](int e) { return D.e_coef[e] * vx[D.e_idx[e]]; });
    for (int r = lane_id(); r < N.n; r += WAVE) {
        const int sp = N.red2sp[r];
        rhs[r] = y_old[sp] - y_cur[sp] + dt * seg_total(D, part, r);
    }
    wave_sync();
}

// One Newton iteration (newton_iteration, backward_euler.F90:120-176) for zone z with the
// block solve in precision T; the zone must be in PH_NEWTON with factor precision T (state
// pad0: 0 FP32, 1 FP64).  Writes y_new and grd (or lib_err) and sets AFTER_NEWTON, or raises
// the zone to FP64 and leaves it in PH_NEWTON.
template <class T, int NB>
__device__ void zone_newton_block(const KArgs &a, const BKArgs &b, char *slot,
                                  NUPPN_LDS char *lds, int z) {
    // zone state: pad0 factor precision (0 FP32, 1 FP64), pad1 factor state (0 none, 1 FP32
    // factors, 2 FP64 factors) of the zone's factor area
    constexpr int want = sizeof(T) == 8 ? 1 : 0;
    const DevNet &N = a.N;
    const DevKnobs &K = a.K;
    const ZoneData &Z = a.Z;
    ZoneState s = Z.state[z];
    if (s.phase != PH_NEWTON || s.pad0 != want) return;
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
    double *part = (double *) (slot + b.off[B_PART]);
    const double dt = s.dttry;
    block_assemble(N, b.B, dt, y_old, y_cur, v, xn, vx, stem, Aval, rhs, part);
    char *fac = b.fac_base + (size_t) z * b.fac_bytes;
    const int rc = block_refined_solve<T, NB>(N, b, Aval, rhs, y_cur, x, res, d, part, fac, lds,
                                              s.pad1, s.st);
    if (rc == 1) {   // FP32 factors insufficient: the FP64 kernel redoes this iteration
        s.pad0 = 1;
        s.pad1 = 0;
        if (lane_id() == 0) Z.state[z] = s;
        wave_sync();
        return;
    }
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

// every zone in PH_NEWTON at precision T; each wavefront owns one scratch slot and strides over
// the zones
template <class T, int NB>
__global__ void __launch_bounds__(64) zone_newton_block_kernel(KArgs a, BKArgs b) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS char *lds = (NUPPN_LDS char *) smem;
    char *slot = b.slot_base + (size_t) blockIdx.x * b.slot_bytes;
    for (int z = blockIdx.x; z < a.Z.nz; z += gridDim.x) zone_newton_block<T, NB>(a, b, slot, lds, z);
}

}  // namespace nuppn
