// Per-zone integration on one wavefront: backward-Euler Newton (solver/source/backward_euler.F90),
// Bader-Deuflhard (solver/source/bader_deuflhard.F90) and the sub-step controller
// (solver/source/solver.F90:53-238), written against the device building blocks.
#pragma once
#include "device_common.hpp"

namespace nuppn {

// knobs (solver_knobs.F90) as the kernels read them
struct DevKnobs {
    int ittd, max_sub_steps, integration_method, interpolate_temp, interpolate_dens;
    double dgrd, grdthreshold, min_t9_for_slv, delta_state_max;
};

// per-zone diagnostics, layout of nuppn_zone_stats
struct ZoneStats {
    int nsubt, titers, iter, nvar1, nsucc, nfail;
    long long n_newton, n_factor, n_solve, lu_nnz;
    long long n_refine, n_factor_fp64;
};

// Per-slot global scratch (one slot per resident wavefront of the persistent kernel).
struct Slot {
    double *y_old, *y_cur, *y_new, *yps, *yps_old, *xn, *dxdt;  // [nsp]
    double *vx, *st1, *st2, *vbuf;                              // [nre]
    double *Aval;                                               // [nnzA]
    double *rhs;                                                // [n]
    double *sortbuf;                                            // [pow2 >= nsp]
    double *jac_cpy;                                            // [nnzA]  (BD)
    double *delta;                                              // [nsp]   (BD)
    double *qcol, *yerr;                                        // [8*nsp] (BD)
    double *yscal, *yext;                                       // [nsp]   (BD)
    LUStore lu;
};


// BD tables (bader_deuflhard.F90:66-124), computed on the host
struct BDTables {
    int nseq[9];
    double a[9];
    double alf[8][8];
    int kmax, kopt0;
};

constexpr double SUM_TOL = 1.0e-6;
constexpr double SMALL_C = 1.0e-30;

// -------------------------------------------------------------------------------------------
// bitonic sort (ascending) of a buffer (global or LDS) of length m (power of two) by one
// wavefront
template <class P>
__device__ inline void wave_bitonic_sort(P a, int m) {
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane_id(); i < m; i += WAVE) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const double ai = a[i], aj = a[ixj];
                    const bool up = (i & k) == 0;
                    if ((ai > aj) == up) { a[i] = aj; a[ixj] = ai; }
                }
            }
            wave_sync();
        }
    }
}

// sortedsum (utils.F90:481-501): sum in ascending order.  buf holds >= pow2(n) doubles.
template <class P>
__device__ inline double sorted_sum(const double *vals, int n, P buf) {
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = lane_id(); i < m; i += WAVE) buf[i] = (i < n) ? vals[i] : __builtin_inf();
    wave_sync();
    wave_bitonic_sort(buf, m);
    return seq_sum(buf, n);
}

// calculate_ye (utils.F90:537-553)
template <class P>
__device__ inline double calculate_ye(const DevNet &N, const double *x, double *eta, P buf) {
    for (int i = lane_id(); i < N.nsp; i += WAVE)
        eta[i] = (i != N.igamma) ? (N.aw[i] - 2.0 * N.zw[i]) * x[i] / N.aw[i] : 0.0;
    wave_sync();
    return (1.0 - sorted_sum(eta, N.nsp, buf)) * 0.5;
}

__device__ inline void copy_vec(double *dst, const double *src, int n) {
    for (int i = lane_id(); i < n; i += WAVE) dst[i] = src[i];
    wave_sync();
}

__device__ inline double wave_maxneg(const double *y, int n) {
    double m = -__builtin_huge_val();
    for (int i = lane_id(); i < n; i += WAVE) m = fmax(m, -y[i]);
    return wave_max(m);
}

// ypsplus (utils.F90:726-740) of one value
__device__ __forceinline__ double ypsplus_value(double t) {
    if (t < -1.0e-27) t = 1.0e-99;
    if (t < 1.0e-27) t = 1.0e-99;
    return t;
}


// rates of one state into v (one wavefront): the REACLIB formula, or the T9 cache when it is on
__device__ inline void zone_rates(const DevNet &N, double t9, double rho, double *v) {
    if (N.rcache) {
        const CachePos c = cache_pos(N, t9);
        for (int l = lane_id(); l < N.nre; l += WAVE) v[l] = cached_rate(N, l, N.rtype[l], c, rho);
    } else {
        const T9Pow p = t9_powers(t9);
        for (int l = lane_id(); l < N.nre; l += WAVE)
            v[l] = reaclib_rate(N.coef, N.nre, l, N.rtype[l], p, rho);
    }
    wave_sync();
}

// -------------------------------------------------------------------------------------------
// Bader-Deuflhard (bader_deuflhard.F90:139-488).  The reduced matrix of level k is
// I - h J with J taken once per step; every level factors once and solves nseq(k)+1 times.
// The reference keeps this state between calls: frst is integrate_network's SAVE'd local
// (solver.F90:75), kopt a module variable (bader_deuflhard.F90:32), reduce and the error terms
// locals and module variables that later steps read before setting (sumerr when abuerr < 1e-3,
// reduce when no branch sets it).
struct BDState {
    int frst, reduce, kopt, pad;
    double errmax, abuerr, negerr, sumerr;
};

// errors.F90:5-16
enum BDErr { E_SLV_CNVRG = 1, E_BAD_STEP = 3, E_NEG_ABUNDS = 4, E_HUGE_ABUNDS = 5, E_SUM_X = 6,
             E_GOT_NAN = 7 };

}  // namespace nuppn
