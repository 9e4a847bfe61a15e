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
    // the dynamic network (cyminformal >= 0): the mask and the reduced entries are recomputed
    // at every Jacobian evaluation (jac_rhs.F90:308-436)
    int dynamic;
    double cyminformal, min_rate_ov_dt;
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
    int *first;                                                 // [nsp]   (dynamic network)
    unsigned char *rflag, *eflag;                               // [nre], [nnzA]
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
    struct Eq {
        double a, zz, xx;
    };
    batched_for<4>(
        lane_id(), N.nsp, WAVE, [&](int i) { return Eq{N.aw[i], N.zw[i], x[i]}; },
        [&](int i, const Eq &q) {
            eta[i] = (i != N.igamma) ? (q.a - 2.0 * q.zz) * q.xx / q.a : 0.0;
        });
    wave_sync();
    return (1.0 - sorted_sum(eta, N.nsp, buf)) * 0.5;
}

__device__ inline void copy_vec(double *dst, const double *src, int n) {
    batched_for<4>(
        lane_id(), n, WAVE, [&](int i) { return src[i]; }, [&](int i, double v) { dst[i] = v; });
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


// physical constants, utils/source/constants.F90:11-45 (the compile-time values)
namespace phys {
constexpr double kEv2Erg = 1.602177e-12;
constexpr double kMev2Erg = kEv2Erg * 1.0e6;
constexpr double kErg2Mev = 1.0 / kMev2Erg;
constexpr double kPi = 3.141592653589793;
constexpr double kAvogadro = 6.022140857e23;
constexpr double kBoltzmann = 1.3806485279e-16;
constexpr double kPlanckMev = 4.135667662e-21;
constexpr double kPlanckErg = kPlanckMev * kMev2Erg;
constexpr double kHbarErg = kPlanckErg / 2.0 / kPi;
constexpr double kAmu = 1.6605402e-24;
constexpr double kQe = 4.80320425e-10;
constexpr double kQe2 = kQe * kQe;
}  // namespace phys

// calculate_reverse_rates + reverse_rate (reverse.F90:107-337; partition functions on, no
// Coulomb corrections): each pair's reverse rate from its forward rate when that exceeds 1e-97,
// G = partition function x spin factor (nse_swj.F90:613-625).  Pairs write distinct reverse
// rates and read forward ones only, so lanes take them in any order.
__device__ inline void zone_detailed_balance(const DevNet &N, double t9, double rho, double *v) {
    using namespace phys;
    const DevPhys &P = N.P;
    const int nre = N.nre;
    const PfPos q = pf_pos(P, t9);
    const double kTerg = kBoltzmann * t9 * 1.0e9;
    const double kTMeV = kTerg * kErg2Mev;
    auto G = [&](int s) { return pf_value(P, s, q) * P.g_spin[s]; };
    for (int d = lane_id(); d < P.ndb; d += WAVE) {
        const int i = P.db_fwd[d];
        const double vi = v[i];
        if (!(vi > 1.0e-97)) continue;
        const int a1 = N.k[i], na1 = N.k[nre + i], a2 = N.k[2 * nre + i];
        const int b1 = N.k[4 * nre + i], nb1 = N.k[5 * nre + i], b2 = N.k[6 * nre + i];
        const int nb2 = N.k[7 * nre + i];
        const double G_a1 = G(a1), m_a1 = P.mass_g[a1];
        double G_a2, m_a2, G_b1, m_b1;
        if (na1 == 1) {
            G_a2 = G(a2);
            m_a2 = P.mass_g[a2];
        } else if (na1 == 2) {
            G_a2 = G_a1;
            m_a2 = m_a1;
        } else {
            continue;
        }
        const double G_b2 = G(b2), m_b2 = P.mass_g[b2];
        if (nb2 == 1) {
            if (nb1 == 1) {
                G_b1 = G(b1);
                m_b1 = P.mass_g[b1];
            } else {
                G_b1 = 0.0;
                m_b1 = 0.0;
            }
        } else if (nb2 == 2) {
            G_b1 = G_b2;
            m_b1 = m_b2;
        } else {
            continue;
        }
        double pfac, mfac, Q, rhofac;
        if (b1 == N.igamma) {   // (n,g), (p,g), (a,g)
            pfac = G_a1 * G_a2 / G_b2;
            mfac = pow(m_a1 * m_a2 / m_b2 * kTerg * 2.0 * kPi / kPlanckErg / kPlanckErg, 1.5);
            Q = -P.q_mev[a1] - P.q_mev[a2] + P.q_mev[b2];
            rhofac = 1.0 / rho / kAvogadro;
        } else {
            pfac = G_a1 * G_a2 / G_b1 / G_b2;
            mfac = pow(m_a1 * m_a2 / m_b1 / m_b2, 1.5);
            Q = -P.q_mev[a1] - P.q_mev[a2] + P.q_mev[b1] + P.q_mev[b2];
            rhofac = 1.0;
        }
        const double x = fmin(-Q / kTMeV, 200.0);
        const double efac = exp(x);
        double fac = mfac * efac * rhofac;
        fac = fac * pfac;
        fac = fmin(fmax(fac, 1.0e-99), 1.0e99);
        v[P.db_rev[d]] = fmax(vi * fac, 1.0e-99);
    }
    wave_sync();
}

// Chugunov, DeWitt & Yakovlev (2007) screening as evaluate_all_rates applies it
// (evaluate_rates.F90:65-121; screening_chugunov.F90:86-235; calculate_azbar utils.F90:504-525):
// the zone's mean ion weight and charge from x (sequential sums in species order, staged in
// buf >= nsp doubles), the plasma temperature with its fade below 0.2 T_p, then each reaction's
// factor from its reactants' charges (an identical pair, k2 = 2, screens k1 with itself)
__device__ inline void zone_screening(const DevNet &N, double t9, double rho, const double *x,
                                      double *v, NUPPN_LDS double *buf) {
    using namespace phys;
    const int nsp = N.nsp, nre = N.nre;
    const double third = 1.0 / 3.0;
    double sums[3];
    for (int k = 0; k < 3; ++k) {
        for (int i = lane_id(); i < nsp; i += WAVE) {
            double t = 0.0;
            if (i != N.igamma) {
                const double y = x[i] / N.aw[i];
                t = k == 0 ? y : (k == 1 ? N.aw[i] * y : N.zw[i] * y);
            }
            buf[i] = t;
        }
        wave_sync();
        sums[k] = seq_sum(buf, nsp);
        wave_sync();
    }
    const double abar = sums[1] / sums[0], zbar = sums[2] / sums[0];
    const double fourov3pi = 4.0 / (3.0 * kPi * kPi);
    const double fourpiqe2 = 4.0 * kPi * kQe2;
    const double tkin = t9 * 1.0e9;
    const double ntot = rho / (abar * kAmu);
    const double denom = pow(3.0 / (4.0 * kPi * zbar * ntot), third);
    const double a_av = denom * pow(zbar, third);
    const double wp = pow(fourpiqe2 * zbar * zbar * ntot / (abar * kAmu), 0.5);
    const double tp = kHbarErg * wp / kBoltzmann;
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
    const double gamfac = kQe2 / (a_av * tk) / kBoltzmann;
    const double zeta = pow(fourov3pi * tp * tp / tk / tk, third);
    const double c_a1 = 2.7822, c_a2 = 98.34, c_a3 = sqrt(3.0) - c_a1 / sqrt(c_a2);
    const double c_b1 = -1.7476, c_b2 = 66.07, c_b3 = 1.12, c_b4 = 65.0, alfa = 0.022;
    // in place, each rate from itself: 8 of a lane's rates are read before any is written
    NUPPN_GLB double *vg = gptr(v);
    batched_for<8>(lane_id(), nre, WAVE, [&](int i) {
        const double vi = vg[i];
        if (vi <= 1.0e-99) return vi;
        const int r1 = N.k[i], r2 = N.k[nre + i] == 2 ? N.k[i] : N.k[2 * nre + i];
        const double z1 = N.zw[r1], z2 = N.zw[r2];
        double screen = 1.0;
        if (z1 > 0.0 && z2 > 0.0) {
            const double gam = z1 * z2 * gamfac;
            const double beta = 0.41 - 0.6 / gam;
            const double gama = 0.06 + 2.2 / gam;
            double gamtild = gam / pow(1.0 + alfa * zeta + beta * (zeta * zeta) +
                                       gama * (zeta * zeta * zeta), third);
            if (gamtild >= 650.0) {
                gamtild = 600.0;
            } else if (gamtild > 550.0) {
                const double sg = 1.0 / (1.0 + exp(-(gamtild - 599.0) / 1.0));
                gamtild = sg * 600.0 + (1.0 - sg) * gamtild;
            }
            const double gamtild2 = gamtild * gamtild;
            const double t1 = pow(gamtild, 3.0 * 0.5) *
                              (c_a1 / sqrt(c_a2 + gamtild) + c_a3 / (1.0 + gamtild));
            const double t2 = c_b1 * gamtild2 / (c_b2 + gamtild);
            const double t3 = c_b3 * gamtild2 / (c_b4 + gamtild2);
            screen = fmin(exp(t1 + t2 + t3), 1.0e30);
        }
        screen = fmin(screen, 1.0e30);
        return vi * screen;
    }, [&](int i, double r) { vg[i] = r; });
    wave_sync();
}

// evaluate_all_rates of one state into v (one wavefront; evaluate_rates.F90:14-172): the
// REACLIB formula (or the T9 cache when it is on) with the partition-function factors of the
// REACLIB reverse rates, then detailed balance and screening at abundances x (buf >= nsp
// doubles of LDS), then the clip
__device__ inline void zone_rates(const DevNet &N, double t9, double rho, double *v,
                                  const double *x, NUPPN_LDS double *buf) {
    const DevPhys &P = N.P;
    // the rates of 8 of a lane's reactions are formed before any is stored (the stores to v
    // could alias the coefficient loads, which would cost a memory round trip per reaction)
    NUPPN_GLB double *vg = gptr(v);
    const auto st = [&](int l, double r) { vg[l] = r; };
    if (N.rcache) {
        const CachePos c = cache_pos(N, t9);
        batched_for<8>(
            lane_id(), N.nre, WAVE, [&](int l) { return cached_rate(N, l, N.rtype[l], c, rho); },
            st);
    } else {
        const T9Pow p = t9_powers(t9);
        if (P.pf_n > 0) {
            const PfPos q = pf_pos(P, t9);
            batched_for<8>(
                lane_id(), N.nre, WAVE,
                [&](int l) {
                    const bool rev = P.rev_ch[l] > 0;
                    return reaclib_rate(N.coef, N.nre, l, N.rtype[l], p, rho,
                                        rev ? reverse_gfac(P, N.nre, l, q) : 1.0, rev);
                },
                st);
        } else {
            batched_for<8>(
                lane_id(), N.nre, WAVE,
                [&](int l) { return reaclib_rate(N.coef, N.nre, l, N.rtype[l], p, rho); }, st);
        }
    }
    wave_sync();
    if (P.pf_n > 0 && P.ndb > 0) zone_detailed_balance(N, t9, rho, v);
    if (P.screen) zone_screening(N, t9, rho, x, v, buf);
    if (P.ndb > 0 || P.screen) {
        batched_for<8>(
            lane_id(), N.nre, WAVE, [&](int l) { return fmax(vg[l], 1.0e-99); }, st);
        wave_sync();
    }
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
