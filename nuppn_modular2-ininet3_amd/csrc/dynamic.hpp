// The dynamic network of calculate_jacobian (solver/source/jac_rhs.F90:308-436) for one zone:
// which species take part in the solve (the mask) and which Jacobian entries the reduced matrix
// keeps, recomputed from the current abundances and rates at every Jacobian evaluation.
//
// The reference decides both in one pass over the reactions in network order: a reaction whose
// molar flux |dfdN| exceeds cyminformal / dt marks its four species (350-353); a reaction's
// entries then enter the reduced matrix only if its reactants and products are marked at that
// point of the loop (398-436), so a reaction early in the list whose species are marked only by
// later reactions is left out although its species end up in the solve.  A reaction with
// |v| / dt <= min_rate_ov_dt is skipped entirely (319).
//
// On the device the loop order is kept without a serial loop: with first(s) the index of the
// first reaction that marks species s (a scan of s's reaction list, which lists reactions in
// ascending order), "s is marked when reaction l is reached" is first(s) <= l.  So every
// reaction decides its entries independently, and an entry is kept when any of its reactions
// keeps it (the reference sets the same row / column for each).
//
// The reduced system is embedded in the static one (every species of every reaction): a species
// outside the mask keeps an identity row and a zero column and a zero right-hand side, so its
// correction is exactly zero, and a left-out entry is an explicit zero.  The solution of the
// masked species is then the solution of the reference's reduced system, and an unmasked
// species keeps y_cur, which is what ddecompac leaves there (y_new = y_cur at the start of every
// iteration, backward_euler.F90:54,104; x = 0 in bader_deuflhard_step, 420).
#pragma once
#include "device_common.hpp"
#include "integrate.hpp"

namespace nuppn {

constexpr int kNotMarked = 0x7fffffff;

// per-zone scratch of the dynamic network
struct DynWork {
    int *first;            // [nsp] first marking reaction of each species (kNotMarked: none)
    unsigned char *rflag;  // [nre] 1: marks its species, 2: skipped (min_rate_ov_dt), 0: neither
    unsigned char *eflag;  // [number of reduced entries] 1: the entry is in the reduced matrix
};

// Marks the species and the entries for the abundances behind xn and the rates v at time step
// dzeit, zeroes the stems of the skipped reactions (their Jacobian terms vanish, so the streamed
// and the pull-map sums leave them out exactly) and returns the number of marked species
// (nvar1).  emap[m * nre + l]: the reduced entry (in the caller's entry order) of slot m
// (ice1..ice8, statement order of 365-392) of reaction l, -1 where the slot holds gamma.
template <class TM>
__device__ __forceinline__ int dyn_network(const DevNet &N, const DevKnobs &K, double dzeit, const double *xnp,
                           const double *vp, double *st1p, double *st2p, const int *emapp,
                           int nent, const DynWork &W, NUPPN_LDS double *red) {
    const int nre = N.nre, g = N.igamma, t0 = TM::tid();
    const NUPPN_GLB int *k = gptr(N.k);
    const NUPPN_GLB double *xn = gptr(xnp), *v = gptr(vp);
    NUPPN_GLB double *st1 = gptr(st1p), *st2 = gptr(st2p);
    NUPPN_GLB unsigned char *rflag = gptr(W.rflag), *eflag = gptr(W.eflag);
    NUPPN_GLB int *first = gptr(W.first);
    const double dfdn_min = K.cyminformal / dzeit;
    int gfirst = kNotMarked;
    for (int l = t0; l < nre; l += TM::NT) {
        const double vl = v[l];
        unsigned char f = 0;
        if (fabs(vl) / dzeit <= K.min_rate_ov_dt) {
            f = 2;
            st1[l] = 0.0;
            st2[l] = 0.0;
        } else {
            const int in1 = k[l], n1 = k[nre + l], in2 = k[2 * nre + l], n2 = k[3 * nre + l];
            const double stem = kFact[n1] * kFact[n2] * vl;
            const double dfdn = stem * powi(xn[in1], n1) * xn[in2];
            if (fabs(dfdn) > dfdn_min) {
                f = 1;
                if (in1 == g || in2 == g || k[4 * nre + l] == g || k[6 * nre + l] == g)
                    gfirst = min(gfirst, l);
            }
        }
        rflag[l] = f;
    }
    for (int e = t0; e < nent; e += TM::NT) eflag[e] = 0;
    gfirst = (int) -team_max<TM>(-(double) gfirst, red);
    TM::sync();
    // first marking reaction of every species of the static system, and the count
    const NUPPN_GLB int *dx_ptr = gptr(N.dx_ptr), *dx_ent = gptr(N.dx_ent);
    const NUPPN_GLB int *red2sp = gptr(N.red2sp);
    int cnt = 0;
    for (int r = t0; r < N.n; r += TM::NT) {
        int fr = kNotMarked;
        for (int e = dx_ptr[r]; e < dx_ptr[r + 1]; ++e) {
            const int l = dx_ent[e] >> 2;
            if (rflag[l] == 1) {
                fr = l;
                break;
            }
        }
        first[red2sp[r]] = fr;
        cnt += fr != kNotMarked;
    }
    if (t0 == 0) first[g] = gfirst;
    const int nvar1 = (int) team_sum<TM>((double) cnt, red);
    TM::sync();
    // the entries of every reaction whose species are marked when the loop reaches it
    const NUPPN_GLB int *emap = gptr(emapp);
    for (int l = t0; l < nre; l += TM::NT) {
        if (rflag[l] == 2) continue;
        const int in1 = k[l], in2 = k[2 * nre + l], out1 = k[4 * nre + l], out2 = k[6 * nre + l];
        if (in2 != g && first[in2] > l) continue;
        if (out1 != g && first[out1] > l) continue;
        if (first[in1] > l || first[out2] > l) continue;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int e = emap[m * nre + l];
            if (e >= 0) eflag[e] = 1;
        }
    }
    TM::sync();
    return nvar1;
}

// whether species sp is in the solve (dyn_network's result)
__device__ __forceinline__ bool dyn_marked(const DynWork &W, int sp) {
    return gptr(W.first)[sp] != kNotMarked;
}

// assemble_matrix under the dynamic network: the reduced matrix's entries as reduced_value
// gives them (with its stops), explicit zeros for the entries it leaves out and identity rows
// for the species outside the mask
__device__ __forceinline__ int assemble_matrix_dyn(const DevNet &N, double scale, const double *st1,
                                          const double *st2, const DynWork &W, double *Aval) {
    const NUPPN_GLB unsigned char *eflag = gptr(W.eflag);
    int stop = 0;
    for (int q = lane_id(); q < N.nnzA; q += WAVE) {
        const bool diag = N.A_isdiag[q];
        double val;
        if (!dyn_marked(W, N.red2sp[N.A_rowind[q]]))
            val = diag ? 1.0 : 0.0;
        else if (!diag && !eflag[q])
            val = 0.0;
        else
            val = reduced_value(jac_entry(N, q, st1, st2), scale, diag, stop);
        Aval[q] = val;
    }
    return ballot(stop) ? 1 : 0;
}

}  // namespace nuppn
