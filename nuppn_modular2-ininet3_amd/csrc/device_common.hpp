// Device building blocks shared by the kernels: wavefront helpers, REACLIB rates, the pull-form
// dX/dt and Jacobian assembly, and the left-looking GEPP with SuperLU's pivot policy.
//
// One wavefront (64 lanes) owns one zone.  Every floating-point result that the reference
// computes sequentially is produced here with the same operations in the same order per
// element (the translation unit is compiled with -ffp-contract=off), so results are equal bit
// for bit to a sequential evaluation given equal inputs; only the libm calls of the rate
// formula (exp, log, cbrt) may differ in the last place from the host's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nuppn {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }

__device__ __forceinline__ uint64_t lanes_below() {
    return (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id()));
}

// value of lane l, for a lane index that is the same on every lane (through the scalar unit)
__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ double readlane_d(double v, int l) {
    const unsigned long long u = (unsigned long long) __double_as_longlong(v);
    const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) u, l);
    const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (u >> 32), l);
    return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}

// LDS address space: pointers into the workgroup's LDS compile to ds_* instructions
#define NUPPN_LDS __attribute__((address_space(3)))

// Full ordering point between dependent phases of one wavefront (global and LDS memory).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); __syncthreads(); }

// Ordering point between dependent LDS accesses of different lanes of one wavefront: the LDS
// executes a wavefront's DS instructions in issue order, so only compiler reordering has to be
// prevented.
__device__ __forceinline__ void lds_fence() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// Ordering point for one wavefront's global and LDS accesses without a workgroup barrier: the
// other wavefronts of a multi-wavefront workgroup may be anywhere else meanwhile (an s_barrier
// here would pair with theirs).  In a single-wavefront workgroup it is the same as wave_sync.
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// The threads that work on one zone: one wavefront (NW = 1: the plain path and the batched
// path's many-zone kernels) or the NW wavefronts of one workgroup (the batched path's kernel for
// passes with few zones).  sync() is then a workgroup barrier; the reductions go through red, a
// few LDS slots of the workgroup.
template <int NW>
struct Team {
    static constexpr int NWAVE = NW;
    static constexpr int NT = WAVE * NW;
    __device__ static int tid() { return NW == 1 ? lane_id() : (int) threadIdx.x; }
    __device__ static int wave() { return NW == 1 ? 0 : (int) threadIdx.x / WAVE; }
    __device__ static void sync() { wave_sync(); }
};

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}

// --- reductions over the threads of a zone (Team): every thread gets the result.  With one
// wavefront they are cross-lane operations; with several, each wavefront's value goes through
// the LDS slots red and the wavefronts' values are combined in wavefront order.
template <class TM>
__device__ __forceinline__ double team_max(double v, NUPPN_LDS double *red) {
    v = wave_max(v);
    if constexpr (TM::NWAVE == 1) {
        return v;
    } else {
        if (lane_id() == 0) red[TM::wave()] = v;
        TM::sync();
        double m = red[0];
#pragma unroll
        for (int w = 1; w < TM::NWAVE; ++w) m = fmax(m, red[w]);
        TM::sync();
        return m;
    }
}

template <class TM>
__device__ __forceinline__ double team_sum(double v, NUPPN_LDS double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    if constexpr (TM::NWAVE == 1) {
        return v;
    } else {
        if (lane_id() == 0) red[TM::wave()] = v;
        TM::sync();
        double m = red[0];
#pragma unroll
        for (int w = 1; w < TM::NWAVE; ++w) m = m + red[w];
        TM::sync();
        return m;
    }
}

// K independent team_sums with one pair of barriers (the same order of additions as K calls of
// team_sum, so the same bits): scr holds K * NWAVE doubles of LDS
template <class TM, int K>
__device__ __forceinline__ void team_sum_k(double (&v)[K], NUPPN_LDS double *scr) {
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, WAVE);
    if constexpr (TM::NWAVE > 1) {
        if (lane_id() == 0)
#pragma unroll
            for (int k = 0; k < K; ++k) scr[k * TM::NWAVE + TM::wave()] = v[k];
        TM::sync();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double m = scr[k * TM::NWAVE];
#pragma unroll
            for (int w = 1; w < TM::NWAVE; ++w) m = m + scr[k * TM::NWAVE + w];
            v[k] = m;
        }
        TM::sync();
    }
}

template <class TM>
__device__ __forceinline__ bool team_any(bool p, NUPPN_LDS double *red) {
    const bool w = ballot(p) != 0;
    if constexpr (TM::NWAVE == 1) {
        return w;
    } else {
        if (lane_id() == 0) red[TM::wave()] = w ? 1.0 : 0.0;
        TM::sync();
        bool any = false;
#pragma unroll
        for (int k = 0; k < TM::NWAVE; ++k) any = any || red[k] != 0.0;
        TM::sync();
        return any;
    }
}


// s = ((0 + a[0]) + a[1]) + ... : the sequential order of Fortran's SUM (every lane gets s)
template <class P>
__device__ inline double seq_sum(P a, int n) {
    double s = 0.0;
    for (int base = 0; base < n; base += WAVE) {
        int i = base + lane_id();
        double v = (i < n) ? a[i] : 0.0;
        int cnt = min(WAVE, n - base);
        for (int t = 0; t < cnt; ++t) s = s + readlane_d(v, t);
    }
    return s;
}

// The same sum over an LDS buffer: every lane folds the values itself from broadcast reads (the
// reads of a group are independent and issue together; only the adds form the chain), which
// avoids the cross-lane moves of the register form above.
__device__ inline double seq_sum(NUPPN_LDS double *a, int n) {
    double s = 0.0;
    int i = 0;
    for (; i + 16 <= n; i += 16) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = a[i + k];
#pragma unroll
        for (int k = 0; k < 16; ++k) s = s + v[k];
    }
    for (; i < n; ++i) s = s + a[i];
    return s;
}

// seq_sum of a global array through an LDS buffer of >= n doubles
__device__ inline double seq_sum_staged(const double *a, int n, NUPPN_LDS double *buf) {
    for (int i = lane_id(); i < n; i += WAVE) buf[i] = a[i];
    wave_sync();
    const double s = seq_sum(buf, n);
    wave_sync();
    return s;
}

// real ** integer as gfortran evaluates it (libgcc __powidf2: square-and-multiply)
__device__ __forceinline__ double powi(double x, int m) {
    unsigned int n = m < 0 ? -(unsigned int) m : (unsigned int) m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

__device__ __constant__ const double kFact[5] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0};

// T9 powers of reaclib_calculate_rates (physics/source/reaclib.F90:443,473-479)
struct T9Pow {
    double t9, t9inv, p13, p13inv, p53, lnt9;
};

__device__ __forceinline__ T9Pow t9_powers(double t9in) {
    T9Pow p;
    p.t9 = fmax(t9in, 0.01);
    p.t9inv = 1.0 / p.t9;
    p.p13 = pow(p.t9, 1.0 / 3.0);
    p.p13inv = 1.0 / p.p13;
    double p43 = p.t9 * p.p13;
    p.p53 = p.p13 * p43;
    p.lnt9 = log(p.t9);
    return p;
}

// one rate: REACLIB sum in the reference's order, exp, density scaling by reaction type,
// the 1e-99 floors of reaclib.F90:581 and evaluate_rates.F90:46,172
__device__ __forceinline__ double reaclib_rate(const double *__restrict__ coef, int nre, int l,
                                               int rtype, const T9Pow &p, double rho,
                                               double gfac = 1.0, bool reverse = false) {
    double r0 = coef[l] + coef[nre + l] * p.t9inv;
    r0 = r0 + coef[2 * nre + l] * p.p13inv;
    r0 = r0 + coef[3 * nre + l] * p.p13;
    r0 = r0 + coef[4 * nre + l] * p.t9;
    r0 = r0 + coef[5 * nre + l] * p.p53;
    r0 = r0 + coef[6 * nre + l] * p.lnt9;
    double reacv = 0.0 + exp(r0);
    if (reverse) reacv = reacv * gfac;   // a REACLIB reverse rate's partition functions
    double rv = (rtype == 0) ? reacv * rho : reacv;
    rv = fmax(rv, 1.0e-99);
    rv = fmax(rv * 1.0, 1.0e-99);
    return fmax(rv, 1.0e-99);
}

// Rate physics beyond the bare REACLIB sum (physics/source/evaluate_rates.F90:14-172), all off
// when pf_n and screen are 0: partition functions on REACLIB's T9 grid and the factors of its
// reverse ('v') rates (reaclib.F90:512-557, 586-605), detailed-balance reverse rates of the
// paired reactions (reverse.F90:107-337), Chugunov screening (screening_chugunov.F90).
struct DevPhys {
    int pf_n;                // grid points (0: no partition functions, no detailed balance)
    const double *pf_t9;     // [pf_n]
    const double *pf_tab;    // [nsp][pf_n]
    const int *rev_ch;       // [nre] chapter of a REACLIB reverse rate, 0 otherwise
    const int *rev_part;     // [6][nre] participants e1..e6 (species or -1)
    int ndb;                 // detailed-balance pairs
    const int *db_fwd;       // [ndb] forward reaction
    const int *db_rev;       // [ndb] its reverse partner
    const double *g_spin, *mass_g, *q_mev;   // [nsp]
    int screen;              // 1: Chugunov screening
};

// reaclib_interpolate_partition_functions (reaclib.F90:586-605) with bsearch_r8
// (utils.F90:687-722): the interval of the clipped temperature and the fraction within it
struct PfPos {
    int i0;
    double f;
};

__device__ inline PfPos pf_pos(const DevPhys &P, double t9_in) {
    const int n = P.pf_n;
    const double *rt9 = P.pf_t9;
    const double t9 = fmax(rt9[0], fmin(t9_in, rt9[n - 1]));
    int a = 1, b = n, r = b - a, m = (a + b) / 2;
    while (t9 != rt9[m - 1] && r > 1) {
        if (rt9[m - 1] > t9) b = m;
        else a = m;
        r = b - a;
        m = (a + b) / 2;
    }
    if (rt9[m - 1] > t9) m = m - 1;
    const double t90 = rt9[m - 1], t91 = rt9[m];
    return {m - 1, (t9 - t90) / (t91 - t90)};
}

__device__ __forceinline__ double pf_value(const DevPhys &P, int s, PfPos q) {
    const double *p = P.pf_tab + (size_t) s * P.pf_n;
    return p[q.i0] + (p[q.i0 + 1] - p[q.i0]) * q.f;
}

// Gfac of a REACLIB reverse rate by chapter (reaclib.F90:512-557); 1 for the others
__device__ inline double reverse_gfac(const DevPhys &P, int nre, int l, PfPos q) {
    const int ch = P.rev_ch[l];
    double g[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int sp = P.rev_part[e * nre + l];
        g[e] = sp >= 0 ? pf_value(P, sp, q) : 1.0;
    }
    switch (ch) {
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

// ------------------------------------------------------------------------------------------
// Static network description on the device (built once on the host).
struct DevNet {
    int nsp, nre, n, nnzA, igamma;
    const int *k;           // [8*nre]
    const int *rtype;       // [nre]
    const double *coef;     // [7*nre]
    const double *aw;       // [nsp]
    const double *zw;       // [nsp]
    const int *red2sp;      // [n] reduced index -> species
    // dX/dt pull map over reduced species: entries (l << 2 | role), role 0..3 = k1,k3,k5,k7
    const int *dx_ptr;      // [n+1]
    const int *dx_ent;
    // Jacobian pull map over reduced CSC entries q: entries (l << 3 | m), m = statement order
    // ice1..ice8 of calculate_jacobian (jac_rhs.F90:365-392)
    const int *jq_ptr;      // [nnzA+1]
    const int *jq_ent;
    const int *A_colptr;    // [n+1] reduced CSC pattern
    const int *A_rowind;    // [nnzA]
    const unsigned char *A_isdiag;  // [nnzA]
    const int *perm_c;      // [n] SuperLU convention
    const int *iperm_c;     // [n]
    // the dynamic network: reduced CSC entry of slot m (ice1..ice8) of reaction l at
    // [m * nre + l], -1 where the slot holds gamma
    const int *dyn_q;
    // the REACLIB T9 cache (null when the rates are evaluated directly): [kNT9][nre] rates at
    // rho = 1 on the grid rc_t9 [kNT9] of spacing rc_dt9
    const double *rcache;
    const double *rc_t9;
    double rc_dt9;
    DevPhys P;
};

// num_t9_preproc (physics/source/reaclib.F90:155): the cache's grid, T9 = 0.01 .. 10
constexpr int kNT9 = 5001;

// a rate from the T9 cache (reaclib.F90:443-470): the grid interval of max(t9, 0.01)
// (floor((t9 - 0.01) / dt9), clamped to the table), linear interpolation, x rho for the
// two-body rates, then the floors of reaclib.F90:581 and evaluate_rates.F90:46,172
struct CachePos {
    int i1;
    double f;
};

__device__ __forceinline__ CachePos cache_pos(const DevNet &N, double t9in) {
    const double t9 = fmax(t9in, 0.01);
    int i1 = (int) floor((t9 - 0.01) / N.rc_dt9);
    i1 = max(0, min(i1, kNT9 - 2));
    return {i1, (t9 - N.rc_t9[i1]) / N.rc_dt9};
}

__device__ __forceinline__ double cached_rate(const DevNet &N, int l, int rtype, CachePos p,
                                              double rho) {
    const double *c1 = N.rcache + (size_t) p.i1 * N.nre;
    const double a = c1[l], b = c1[N.nre + l];
    double rv = a + p.f * (b - a);
    if (rtype == 0) rv = rv * rho;
    rv = fmax(rv, 1.0e-99);
    rv = fmax(rv * 1.0, 1.0e-99);
    return fmax(rv, 1.0e-99);
}

// Pointers reach the device code as generic pointers (struct fields, arguments of functions
// that are not inlined), so the compiler emits flat memory instructions for them.  A flat load
// counts against the LDS counter as well: every LDS wait then also waits for the outstanding
// global loads, and prefetches issued ahead of LDS-bound work stop overlapping it.  The hot
// helpers therefore access memory through global-address-space pointers; every pointer given to
// a kernel here is device (global) memory.
#define NUPPN_GLB __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ NUPPN_GLB T *gptr(T *p) {
    return (NUPPN_GLB T *) p;
}
template <class T>
__device__ __forceinline__ const NUPPN_GLB T *gptr(const T *p) {
    return (const NUPPN_GLB T *) p;
}
// for (i = t0; i < n; i += nt) store(i, load(i)), with the loads of U of a thread's iterations
// issued before any of their stores.  A plain loop whose stores might alias its loads waits
// out a full memory round trip per iteration; here one round trip covers U iterations.  Each
// element's arithmetic is unchanged.
template <int U, class LD, class ST>
__device__ __forceinline__ void batched_for(int t0, int n, int nt, LD load, ST store) {
    for (int i0 = t0; i0 < n; i0 += U * nt) {
        decltype(load(0)) r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * nt;
            if (i < n) r[u] = load(i);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * nt;
            if (i < n) store(i, r[u]);
        }
    }
}

// xn = X/A for considered isotopes (A > 0), xn(gamma) = 1 (jac_rhs.F90:235-240, 312-317)
// (thread t0 of the nt that share the zone)
__device__ inline void molar(const DevNet &N, const double *xp, double *xnp, int t0 = lane_id(),
                             int nt = WAVE) {
    const NUPPN_GLB double *x = gptr(xp), *aw = gptr(N.aw);
    NUPPN_GLB double *xn = gptr(xnp);
    struct Mq {
        double x, a;
    };
    batched_for<4>(
        t0, N.nsp, nt, [&](int i) { return Mq{x[i], aw[i]}; },
        [&](int i, const Mq &q) { xn[i] = i == N.igamma ? 1.0 : (q.a > 0.0 ? q.x / q.a : 0.0); });
}

// per-reaction quantities: vx (calculate_dxdt flux, jac_rhs.F90:245) and the two Jacobian stems
// (jac_rhs.F90:355, 359)
__device__ inline void reaction_terms(const DevNet &N, const double *xnp, const double *vp,
                                      double *vxp, double *st1p, double *st2p,
                                      int t0 = lane_id(), int nt = WAVE) {
    const int nre = N.nre;
    const NUPPN_GLB int *k = gptr(N.k);
    const NUPPN_GLB double *xn = gptr(xnp), *v = gptr(vp);
    NUPPN_GLB double *vx = gptr(vxp), *st1 = gptr(st1p), *st2 = gptr(st2p);
    struct Rt {
        double y1, y2, vl;
        int n1, n2;
    };
    batched_for<4>(
        t0, nre, nt,
        [&](int l) {
            const int in1 = k[l], n1 = k[nre + l], in2 = k[2 * nre + l], n2 = k[3 * nre + l];
            return Rt{xn[in1], xn[in2], v[l], n1, n2};
        },
        [&](int l, const Rt &q) {
            const double y1 = q.y1, y2 = q.y2, vl = q.vl;
            const int n1 = q.n1, n2 = q.n2;
            vx[l] = kFact[n1] * vl * powi(y1, n1) * y2;
            st1[l] = (double) n1 * kFact[n1] * kFact[n2] * powi(y2, n2) * powi(y1, n1 - 1) * vl;
            st2[l] = (double) n2 * kFact[n1] * kFact[n2] * powi(y1, n1) * powi(y2, n2 - 1) * vl;
        });
}

// dX/dt of reduced species r from the per-reaction fluxes, accumulated in reaction order
// (calculate_dxdt's scatter order), times A (jac_rhs.F90:250-272)
__device__ __forceinline__ double dxdt_species(const DevNet &N, int r, const double *vx) {
    const int nre = N.nre;
    double acc = 0.0;
    for (int e = N.dx_ptr[r]; e < N.dx_ptr[r + 1]; ++e) {
        const int ent = N.dx_ent[e], l = ent >> 2, role = ent & 3;
        const double f = vx[l];
        double term;
        if (role == 0) term = -N.k[nre + l] * f;            // -k2 vx
        else if (role == 1) term = -f;                      // -vx
        else if (role == 2) term = N.k[5 * nre + l] * f;    // k6 vx
        else term = N.k[7 * nre + l] * f;                   // k8 vx
        acc = acc + term;
    }
    return acc * N.aw[N.red2sp[r]];
}

// Jacobian slot value before scaling: 1e-99 seed + contributions in reaction order
// (jac_rhs.F90:309, 365-392)
__device__ __forceinline__ double jac_entry(const DevNet &N, int q, const double *st1,
                                            const double *st2) {
    const int nre = N.nre;
    const double *an = N.aw;
    double acc = 1.0e-99;
    for (int e = N.jq_ptr[q]; e < N.jq_ptr[q + 1]; ++e) {
        const int ent = N.jq_ent[e], l = ent >> 3, m = ent & 7;
        const int in1 = N.k[l], n1 = N.k[nre + l], in2 = N.k[2 * nre + l], n2 = N.k[3 * nre + l];
        const int out1 = N.k[4 * nre + l], no1 = N.k[5 * nre + l], out2 = N.k[6 * nre + l];
        const int no2 = N.k[7 * nre + l];
        const double s1 = st1[l], s2 = st2[l];
        double t;
        switch (m) {
            case 0: t = (-n1 * s1) * an[in1] / an[in1]; break;
            case 1: t = (-n2 * s1) * an[in2] / an[in1]; break;
            case 2: t = (-n1 * s2) * an[in1] / an[in2]; break;
            case 3: t = (no1 * s1) * an[out1] / an[in1]; break;
            case 4: t = (no2 * s1) * an[out2] / an[in1]; break;
            case 5: t = (-n2 * s2) * an[in2] / an[in2]; break;
            case 6: t = (no1 * s2) * an[out1] / an[in2]; break;
            default: t = (no2 * s2) * an[out2] / an[in2]; break;
        }
        acc = acc + t;
    }
    return acc;
}

// Reduced matrix values of I - scale*J (newton_iteration val*(-dt) + reduce_jacobian,
// jac_rhs.F90:140-161).  reduce_jacobian drops every entry of -dt J that is exactly 0 or -1
// (138-146) before it adds the identity: a dropped off-diagonal entry is stored as an explicit
// zero here (the static pattern keeps its slot; with the column order fixed, an explicit zero
// changes no value of the factors or the solve).  The diagonal of a masked species is restored
// and takes val + 1 whatever it held (151-161).  Returns 1 where the reference stops instead:
// an off-diagonal Jacobian entry exactly zero (calculate_jacobian, 444-451) or a reduced
// diagonal exactly zero (177-182).
__device__ __forceinline__ double reduced_value(double jv, double scale, bool diag, int &stop) {
    double val = jv * (-scale);
    if (diag) {
        val = val + 1.0;
        if (val == 0.0) stop = 1;
    } else {
        if (jv == 0.0) stop = 1;
        if (fabs(val) == 0.0 || val == -1.0) val = 0.0;
    }
    return val;
}

__device__ inline int assemble_matrix(const DevNet &N, double scale, const double *st1,
                                      const double *st2, double *Aval) {
    int stop = 0;
    for (int q = lane_id(); q < N.nnzA; q += WAVE)
        Aval[q] = reduced_value(jac_entry(N, q, st1, st2), scale, N.A_isdiag[q], stop);
    return ballot(stop) ? 1 : 0;
}

// ------------------------------------------------------------------------------------------
// Left-looking GEPP on one wavefront.
struct LUWork {                    // LDS
    NUPPN_LDS double *x;           // [n] dense column / solve vector
    NUPPN_LDS unsigned char *flag; // [n]
    NUPPN_LDS int *perm_r;         // [n]
    NUPPN_LDS int *pivrow;         // [n]
    NUPPN_LDS int *touched;        // [n]
    NUPPN_LDS int *Lp;             // [n+1]
    NUPPN_LDS int *ulist;          // [n] pivot columns updating the current column, in order
};

struct LUStore {          // global, per zone slot
    int *Li;
    double *Lx;
    int *Ui;
    double *Ux;
    int *Up;              // [n+1]
};

// Factor the n x n CSC matrix (Acp, Ari, Aval; columns taken in the order iperm_c).  Pivot
// policy of SuperLU's dpivotL (SuperLU_5.0/SRC/dpivotL.c:114-156): the largest |value| among
// the not-yet-pivoted rows of the updated column (first in discovery order on ties) unless the
// original diagonal is nonzero and >= 1.0 * that maximum; L scaled by the reciprocal of the
// pivot (dpivotL.c:178-180).  Column j is updated by the pivot columns k < j in ascending k,
// so every entry receives its updates in the same order as a sequential left-looking sweep.
// Returns 0 or j+1 for the first column without a nonzero pivot (factorisation stops there).
__device__ inline int gepp_factor(int n, const int *__restrict__ Acp,
                                  const int *__restrict__ Ari, const double *__restrict__ Aval,
                                  const int *__restrict__ iperm_c, LUWork w, LUStore s) {
    const int lane = lane_id();
    const uint64_t below = lanes_below();
    for (int i = lane; i < n; i += WAVE) {
        w.x[i] = 0.0;
        w.flag[i] = 0;
        w.perm_r[i] = -1;
    }
    if (lane == 0) { w.Lp[0] = 0; s.Up[0] = 0; }
    wave_sync();
    int nl = 0, nu = 0;
    for (int j = 0; j < n; ++j) {
        const int c = iperm_c[j];
        const int a0 = Acp[c], a1 = Acp[c + 1];
        int nt = 0;
        for (int base = a0; base < a1; base += WAVE) {
            const int p = base + lane;
            const bool act = p < a1;
            const int r = act ? Ari[p] : 0;
            const uint64_t m = ballot(act);
            if (act) {
                w.x[r] = Aval[p];
                w.flag[r] = 1;
                w.touched[nt + popc(m & below)] = r;
            }
            nt += popc(m);
        }
        lds_fence();
        // updates from earlier pivot columns, ascending k
        int nk = 0;
        for (int k0 = 0; k0 < j; k0 += WAVE) {
            const int kl = k0 + lane;
            const int prl = (kl < j) ? w.pivrow[kl] : 0;
            int done = -1;
            for (;;) {
                const bool cand = (kl < j) && (lane > done) && w.flag[prl];
                const uint64_t m = ballot(cand);
                if (!m) break;
                const int t = __ffsll((unsigned long long) m) - 1;
                done = t;
                const int k = k0 + t;
                const int pr = readlane_i(prl, t);
                const double xk = w.x[pr];
                if (lane == 0) w.ulist[nk] = k;
                ++nk;
                const int l0 = w.Lp[k], l1 = w.Lp[k + 1];
                for (int base = l0; base < l1; base += WAVE) {
                    const int p = base + lane;
                    const bool act = p < l1;
                    const int i = act ? s.Li[p] : 0;
                    const double lv = act ? s.Lx[p] : 0.0;
                    const bool isnew = act && !w.flag[i];
                    const uint64_t nm = ballot(isnew);
                    if (isnew) {
                        w.touched[nt + popc(nm & below)] = i;
                        w.flag[i] = 1;
                        w.x[i] = 0.0;
                    }
                    nt += popc(nm);
                    if (act) w.x[i] = w.x[i] - lv * xk;
                    lds_fence();
                }
            }
        }
        // U(:,j): the updating pivot rows' final values, written once per column
        for (int q = lane; q < nk; q += WAVE) {
            const int k = w.ulist[q];
            s.Ui[nu + q] = k;
            s.Ux[nu + q] = w.x[w.pivrow[k]];
        }
        nu += nk;
        // pivot search over the not-yet-pivoted rows, in discovery order
        double best = 0.0;
        int bestt = 0x7fffffff, firstt = 0x7fffffff, diagt = -1;
        for (int t = lane; t < nt; t += WAVE) {
            const int r = w.touched[t];
            if (w.perm_r[r] < 0) {
                const double av = fabs(w.x[r]);
                if (av > best) { best = av; bestt = t; }
                if (t < firstt) firstt = t;
                if (r == c) diagt = t;
            }
        }
        const double gmax = wave_max(best);
        const int bt = wave_min_i((best == gmax && gmax > 0.0) ? bestt : 0x7fffffff);
        firstt = wave_min_i(firstt);
        diagt = wave_max_i(diagt);
        if (firstt == 0x7fffffff || gmax == 0.0) return j + 1;
        int pt = bt;
        if (diagt >= 0) {
            const double av = fabs(w.x[w.touched[diagt]]);
            if (av != 0.0 && av >= 1.0 * gmax) pt = diagt;
        }
        const int piv = w.touched[pt];
        const double pv = w.x[piv];
        if (lane == 0) {
            s.Ui[nu] = j;
            s.Ux[nu] = pv;
            s.Up[j + 1] = nu + 1;
            w.perm_r[piv] = j;
            w.pivrow[j] = piv;
        }
        ++nu;
        lds_fence();
        const double temp = 1.0 / pv;
        for (int base = 0; base < nt; base += WAVE) {
            const int t = base + lane;
            const bool act = t < nt;
            const int r = act ? w.touched[t] : 0;
            const bool inl = act && w.perm_r[r] < 0;
            const uint64_t m = ballot(inl);
            if (inl) {
                const int q = nl + popc(m & below);
                s.Li[q] = r;
                s.Lx[q] = w.x[r] * temp;
            }
            nl += popc(m);
        }
        if (lane == 0) w.Lp[j + 1] = nl;
        for (int t = lane; t < nt; t += WAVE) {
            const int r = w.touched[t];
            w.flag[r] = 0;
            w.x[r] = 0.0;
        }
        wave_sync();   // L(:,j) stores visible to the loads of the next columns
    }
    return 0;
}

// Solve with the factors: b <- Pc U^-1 L^-1 Pr b (dgstrs NOTRANS order, SuperLU_5.0/SRC/
// dgstrs.c:149-317).  b is global [n]; uses w.x as the work vector.
__device__ inline void gepp_solve(int n, const int *__restrict__ perm_c, LUWork w, LUStore s,
                                  double *b) {
    const int lane = lane_id();
    NUPPN_LDS double *y = w.x;
    for (int i = lane; i < n; i += WAVE) y[w.perm_r[i]] = b[i];
    lds_fence();
    for (int j = 0; j < n; ++j) {
        const double yj = y[j];
        const int l0 = w.Lp[j], l1 = w.Lp[j + 1];
        for (int p = l0 + lane; p < l1; p += WAVE) {
            const int i = w.perm_r[s.Li[p]];
            y[i] = y[i] - s.Lx[p] * yj;
        }
        lds_fence();
    }
    for (int j = n - 1; j >= 0; --j) {
        const int u0 = s.Up[j], pd = s.Up[j + 1] - 1;
        const double yj = y[j] / s.Ux[pd];
        lds_fence();   // every lane has read y(j) before it is overwritten
        if (lane == 0) y[j] = yj;
        for (int p = u0 + lane; p < pd; p += WAVE) {
            const int k = s.Ui[p];
            y[k] = y[k] - s.Ux[p] * yj;
        }
        lds_fence();
    }
    for (int i = lane; i < n; i += WAVE) b[i] = y[perm_c[i]];
    wave_sync();
}

}  // namespace nuppn
