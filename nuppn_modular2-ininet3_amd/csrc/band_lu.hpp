// Band form of the batched path's factorisation, for networks whose interior is too wide for the
// block-tridiagonal form (the 5,000-species network: half-bandwidth ~100).
//
// The reduced matrix in the block path's order (interior species in network order, the few hub
// species of the border last) is an arrowhead: a band of half-width bw plus dense border rows and
// columns.  It is factored pivot-free in that order, M = L U, with the same scalings as the block
// path (rows to molar units, columns by the reciprocal diagonal), by one workgroup of 256
// threads per zone:
//
// * the active part of the elimination is a window of W = bw + 1 interior indices plus the
//   border, WT = W + nbd <= 128 rows and columns; interior index i lives in window slot i mod W,
//   so the slot of the eliminated index k is reused for index k + W, whose entries enter the
//   window at that step (none of them has been updated before: an update of row or column r
//   needs a pivot within bw of r);
// * the window sits in registers: thread (tr, tc) of a 16 x 16 grid owns the slots (tr + 16 i,
//   tc + 16 j), up to 8 x 8 of them, whatever the rotation; each step the owners of the pivot row
//   and column put them in LDS, and after one barrier every thread updates its slots with a
//   rank-1 product, writes out its share of L's column and U's row, and loads the entering
//   entries (staged in LDS by the scatter of the same step) into the recycled slots;
// * after the interior the border slots hold the Schur complement, factored with partial
//   pivoting in FP64 by one thread.
//
// The solve runs the two band sweeps on the first wavefront from a vector in LDS, with each
// group of kBandChunk columns of L or U loaded into LDS by the whole workgroup (double buffered:
// the next group's loads are in flight while the wavefront works through the current one).
#pragma once
#include "block_lu.hpp"

namespace nuppn {

constexpr int kBandTG = 16;                        // thread grid side: 16 x 16 = 256 threads
constexpr int kBandR = 8;                          // window rows (and columns) per thread
constexpr int kBandMaxWT = kBandTG * kBandR;       // 128: window + border the grid holds
constexpr int kBandChunk = 32;                     // sweep columns staged per LDS buffer
constexpr int kBandEnterChunk = 16;                // entering lists prefetched per LDS bank
constexpr int kBandEnterCap = 512;                 // entries one bank holds
using BandTeam = Team<4>;

struct DevBand {
    int ni, nbd, bw, W, WT;
    // entering entries: list 0 the border-border block, list r + 1 those of interior index r
    const int *ent_ptr;    // [ni + 2]
    const int *ent_k;      // position in the A value array (CSR order)
    const int *ent_dst;    // window slot (row * WT + col)
    const int *ent_col;    // reduced column (for the column scaling)
    const double *ent_ys;  // mass(col) / mass(row)
};

template <class T>
struct BandFactors {
    T *Lb;        // [ni * bw]: L[k + 1 + t][k] at k * bw + t (zero past ni)
    T *Uc;        // [ni * W]:  U[j - d][j] at j * W + d, d = 0..bw (d = 0 the pivot)
    T *LE;        // [nbd * ni]: L[border b][k] at b * ni + k
    T *UF;        // [ni * nbd]: U[k][border b] at k * nbd + b
    double *G;    // [nbd * nbd] LU of the border Schur complement (column-major)
    int *gpiv;    // [nbd] its row interchanges
    double *cs;   // [n] 1 / diagonal of A: the column scaling (reduced order)
    double *colsc, *rowinv;   // [n] by position: mass * cs, 1 / mass
};

// LDS of one zone (bytes from base): the team reduction slots, the factorisation's pivot row /
// column and entering buffers (two parities each), the solve vector and the sweep buffers
template <class T>
struct BandLds {
    NUPPN_LDS double *red;   // [8]
    NUPPN_LDS T *urow;       // [2][kBandMaxWT]
    NUPPN_LDS T *lcol;       // [2][kBandMaxWT]
    NUPPN_LDS T *erow;       // [2][kBandMaxWT]
    NUPPN_LDS T *ecol;       // [2][kBandMaxWT]
    NUPPN_LDS T *v;          // [n]
    NUPPN_LDS T *stage;      // [2][kBandChunk * W]
    NUPPN_LDS double *tb;    // [nbd]
    NUPPN_LDS T *eval;       // [2][kBandEnterCap] prefetched entering values
    NUPPN_LDS int *edst;     // [2][kBandEnterCap] their window slots
    NUPPN_LDS int *eoff;     // [2][kBandEnterChunk + 1] list starts within the bank
};

__host__ __device__ inline size_t band_lds_bytes(int n, int W, int nbd, size_t tsz) {
    size_t o = 64;
    o += 8 * kBandMaxWT * tsz;
    o = (o + 15) & ~(size_t) 15;
    o += (size_t) n * tsz;
    o = (o + 15) & ~(size_t) 15;
    o += 2 * (size_t) kBandChunk * W * tsz;
    o += 8 * (size_t) (kBandMaxWT / kBandTG);   // nbd <= 8
    (void) nbd;
    o = (o + 15) & ~(size_t) 15;
    o += 2 * (size_t) kBandEnterCap * (tsz + 4) + 8 * (kBandEnterChunk + 1);
    return (o + 255) & ~(size_t) 255;
}

template <class T>
__device__ inline BandLds<T> make_band_lds(NUPPN_LDS char *base, int n, int W) {
    BandLds<T> L;
    size_t o = 0;
    L.red = (NUPPN_LDS double *) (base + o);
    o += 64;
    L.urow = (NUPPN_LDS T *) (base + o);
    L.lcol = L.urow + 2 * kBandMaxWT;
    L.erow = L.lcol + 2 * kBandMaxWT;
    L.ecol = L.erow + 2 * kBandMaxWT;
    o += 8 * kBandMaxWT * sizeof(T);
    o = (o + 15) & ~(size_t) 15;
    L.v = (NUPPN_LDS T *) (base + o);
    o += (size_t) n * sizeof(T);
    o = (o + 15) & ~(size_t) 15;
    L.stage = (NUPPN_LDS T *) (base + o);
    o += 2 * (size_t) kBandChunk * W * sizeof(T);
    L.tb = (NUPPN_LDS double *) (base + o);
    o += 8 * (size_t) (kBandMaxWT / kBandTG);   // nbd <= 8
    o = (o + 15) & ~(size_t) 15;
    L.eval = (NUPPN_LDS T *) (base + o);
    o += 2 * (size_t) kBandEnterCap * sizeof(T);
    L.edst = (NUPPN_LDS int *) (base + o);
    o += 2 * (size_t) kBandEnterCap * 4;
    L.eoff = (NUPPN_LDS int *) (base + o);
    return L;
}

__device__ __forceinline__ float readlane_t(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_t(double v, int l) { return readlane_d(v, l); }

// Factors of the assembled matrix (A in CSR order) in precision T.  Returns 0, or 1 when a pivot
// or the border Schur complement is not finite or zero.  All 256 threads of the zone call it.
template <class T>
__device__ __forceinline__ int band_factor(const DevBlock &B, const DevBand &D, const double *Aval,
                                           const BandFactors<T> &F, const BandLds<T> &L) {
    using TM = BandTeam;
    const int t = TM::tid(), tr = t / kBandTG, tc = t % kBandTG;
    const int n = B.n, ni = D.ni, nbd = D.nbd, bw = D.bw, W = D.W, WT = D.WT;
    const NUPPN_GLB double *A = gptr(Aval);
    NUPPN_GLB double *cs = gptr(F.cs), *colsc = gptr(F.colsc), *rowinv = gptr(F.rowinv);
    const NUPPN_GLB int *vpos = gptr(B.vpos), *diag = gptr(B.diag_a);
    const NUPPN_GLB double *mass = gptr(B.mass);
    int bad = 0;
    for (int i = t; i < n; i += TM::NT) {
        const double d = A[diag[i]];
        const double c = 1.0 / d;
        if (!(isfinite(c) && d != 0.0)) bad = 1;
        cs[i] = c;
        colsc[vpos[i]] = mass[i] * c;
        rowinv[vpos[i]] = 1.0 / mass[i];
    }
    for (int i = t; i < 8 * kBandMaxWT; i += TM::NT) L.urow[i] = T(0);
    if (team_any<TM>(bad != 0, L.red)) return 1;   // (its barriers also order the above)

    const NUPPN_GLB int *eptr = gptr(D.ent_ptr), *ek = gptr(D.ent_k), *edst = gptr(D.ent_dst);
    const NUPPN_GLB int *ecolx = gptr(D.ent_col);
    const NUPPN_GLB double *eys = gptr(D.ent_ys);
    // The entering entries, index after index (lists 1..ni), are prefetched into LDS by groups
    // of kBandEnterChunk lists, with their values already scaled, one group ahead: a step then
    // reads its entries from LDS instead of waiting for three dependent global loads.
    auto prefetch = [&](int grp) __attribute__((always_inline)) {
        const int l0 = 1 + grp * kBandEnterChunk, l1 = min(l0 + kBandEnterChunk, ni + 1);
        if (l0 > ni) return;
        const int bank = grp & 1, e0 = eptr[l0];
        NUPPN_LDS T *bv = L.eval + bank * kBandEnterCap;
        NUPPN_LDS int *bd = L.edst + bank * kBandEnterCap;
        for (int e = e0 + t; e < eptr[l1]; e += TM::NT) {
            bv[e - e0] = (T) (A[ek[e]] * eys[e] * cs[ecolx[e]]);
            bd[e - e0] = edst[e];
        }
        for (int q = t; q <= l1 - l0; q += TM::NT)
            L.eoff[bank * (kBandEnterChunk + 1) + q] = eptr[l0 + q] - e0;
    };
    // Every later group goes through a pipeline of four steps, one dependent load level per
    // step (the list bounds, the entries' fields, the matrix values, the LDS writes), so that no
    // step waits for a global load: group g + 1 starts at the first index of group g and is in
    // its bank three steps later, long before its first index.
    static_assert(kBandEnterCap <= 2 * TM::NT, "two entries per thread");
    int q_e0 = 0, q_e1 = 0, q_off = 0, q_k[2] = {0, 0}, q_col[2] = {0, 0}, q_dst[2] = {0, 0};
    double q_ys[2] = {0.0, 0.0}, q_a[2] = {0.0, 0.0}, q_c[2] = {0.0, 0.0};
    auto pipeline = [&](int rin) __attribute__((always_inline)) {
        const int ph = rin % kBandEnterChunk, grp = rin / kBandEnterChunk + 1;
        const int l0 = 1 + grp * kBandEnterChunk, l1 = min(l0 + kBandEnterChunk, ni + 1);
        if (ph > 3 || l0 > ni) return;
        if (ph == 0) {
            q_e0 = eptr[l0];
            q_e1 = eptr[l1];
            if (t <= l1 - l0) q_off = eptr[l0 + t];
        } else if (ph == 1) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e = q_e0 + t + TM::NT * i;
                if (e < q_e1) {
                    q_k[i] = ek[e];
                    q_ys[i] = eys[e];
                    q_col[i] = ecolx[e];
                    q_dst[i] = edst[e];
                }
            }
        } else if (ph == 2) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
                if (q_e0 + t + TM::NT * i < q_e1) {
                    q_a[i] = A[q_k[i]];
                    q_c[i] = cs[q_col[i]];
                }
        } else {
            const int bank = grp & 1;
            NUPPN_LDS T *bv = L.eval + bank * kBandEnterCap;
            NUPPN_LDS int *bd = L.edst + bank * kBandEnterCap;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int e = q_e0 + t + TM::NT * i;
                if (e < q_e1) {
                    bv[e - q_e0] = (T) (q_a[i] * q_ys[i] * q_c[i]);
                    bd[e - q_e0] = q_dst[i];
                }
            }
            if (t <= l1 - l0) L.eoff[bank * (kBandEnterChunk + 1) + t] = q_off - q_e0;
        }
    };
    // the entries of interior index rin into the entering buffers of parity p (row part: the
    // slot row sr)
    auto scatter = [&](int rin, int sr, int p) __attribute__((always_inline)) {
        const int grp = rin / kBandEnterChunk, q = rin - grp * kBandEnterChunk, bank = grp & 1;
        const NUPPN_LDS int *off = L.eoff + bank * (kBandEnterChunk + 1);
        const NUPPN_LDS T *bv = L.eval + bank * kBandEnterCap;
        const NUPPN_LDS int *bd = L.edst + bank * kBandEnterCap;
        for (int e = off[q] + t; e < off[q + 1]; e += TM::NT) {
            const int dst = bd[e], r = dst / WT, c = dst - r * WT;
            if (r == sr) L.erow[p * kBandMaxWT + c] = bv[e];
            else L.ecol[p * kBandMaxWT + r] = bv[e];
        }
    };
    T w[kBandR][kBandR];
#pragma unroll
    for (int i = 0; i < kBandR; ++i)
#pragma unroll
        for (int j = 0; j < kBandR; ++j) w[i][j] = T(0);
    // the border-border block: its entries through the row buffer, one border row at a time
    for (int b = 0; b < nbd; ++b) {
        for (int e = eptr[0] + t; e < eptr[1]; e += TM::NT) {
            const int dst = edst[e], r = dst / WT;
            if (r == W + b)
                L.erow[dst - r * WT] = (T) (A[ek[e]] * eys[e] * cs[ecolx[e]]);
        }
        TM::sync();
#pragma unroll
        for (int i = 0; i < kBandR; ++i)
            if (tr + kBandTG * i == W + b)
#pragma unroll
                for (int j = 0; j < kBandR; ++j) {
                    const int c = tc + kBandTG * j;
                    w[i][j] = L.erow[c];
                }
        TM::sync();
        for (int c = t; c < kBandMaxWT; c += TM::NT) L.erow[c] = T(0);
        TM::sync();
    }
    // the first W interior indices enter one at a time, then every step eliminates index k and
    // index k + W enters in its slot.  L's column and U's row of every step go to LDS first and
    // reach global memory in batches of CHF steps (a barrier waits for the stores before it, so
    // storing every step would make every step wait for its stores), from two banks in turn.
    const int nfill = min(W, ni);
    NUPPN_GLB T *Lb = gptr(F.Lb), *Uc = gptr(F.Uc), *LE = gptr(F.LE), *UF = gptr(F.UF);
    const int rowlen = bw + W + 2 * nbd;
    const int CHF = kBandChunk * W / rowlen;
    auto flush = [&](int k0, int cnt) __attribute__((always_inline)) {
        const NUPPN_LDS T *o = L.stage + (size_t) ((k0 / CHF) & 1) * CHF * rowlen;
        for (int rr = 0; rr < cnt; ++rr) {
            const int k = k0 + rr;
            const NUPPN_LDS T *orow = o + (size_t) rr * rowlen;
            for (int q = t; q < bw; q += TM::NT) Lb[(size_t) k * bw + q] = orow[q];
            for (int q = t; q <= bw; q += TM::NT)
                if (k + q < ni) Uc[(size_t) (k + q) * W + q] = orow[bw + q];
            for (int b = t; b < nbd; b += TM::NT) {
                LE[(size_t) b * ni + k] = orow[bw + W + b];
                UF[(size_t) k * nbd + b] = orow[bw + W + nbd + b];
            }
        }
    };
    prefetch(0);
    TM::sync();
    int fail = 0;
    for (int k = -nfill; k < ni; ++k) {
        const int p = k & 1;
        const int s = k < 0 ? k + nfill : k % W;         // slot of the pivot (or of the entry)
        const int rin = k < 0 ? k + nfill : k + W;      // index entering at this step
        const bool enter = rin < ni;
        const int si = s / kBandTG, sm = s % kBandTG;   // register index and owner of slot s
        NUPPN_LDS T *ur = L.urow + p * kBandMaxWT, *lc = L.lcol + p * kBandMaxWT;
        NUPPN_LDS T *er = L.erow + p * kBandMaxWT, *ec = L.ecol + p * kBandMaxWT;
        if (k >= 0) {
            if (tr == sm) {
#pragma unroll
                for (int i = 0; i < kBandR; ++i)
                    if (i == si)
#pragma unroll
                        for (int j = 0; j < kBandR; ++j) ur[tc + kBandTG * j] = w[i][j];
            }
            if (tc == sm) {
#pragma unroll
                for (int j = 0; j < kBandR; ++j)
                    if (j == si)
#pragma unroll
                        for (int i = 0; i < kBandR; ++i) lc[tr + kBandTG * i] = w[i][j];
            }
        }
        if (enter) scatter(rin, rin % W, p);
        TM::sync();
        // (group g + 1's bank was last read before this barrier, at index g * chunk - 1)
        if (enter) pipeline(rin);
        if (k > 0 && k % CHF == 0) flush(k - CHF, CHF);
        if (k >= 0) {
            const T piv = ur[s];
            if (!(isfinite((double) piv) && piv != T(0))) {   // uniform: every thread reads it
                fail = 1;
                break;
            }
            const T inv = T(1) / piv;
            // the rank-1 update of every slot (the pivot row and column are recycled below)
            T l[kBandR], u[kBandR];
#pragma unroll
            for (int i = 0; i < kBandR; ++i) l[i] = lc[tr + kBandTG * i] * inv;
#pragma unroll
            for (int j = 0; j < kBandR; ++j) u[j] = ur[tc + kBandTG * j];
#pragma unroll
            for (int i = 0; i < kBandR; ++i)
#pragma unroll
                for (int j = 0; j < kBandR; ++j) w[i][j] = fma_t(-l[i], u[j], w[i][j]);
            // L's column k, U's row k and the border parts to the output bank
            NUPPN_LDS T *orow = L.stage + (size_t) ((k / CHF) & 1) * CHF * rowlen +
                                (size_t) (k % CHF) * rowlen;
            for (int q = t; q < bw; q += TM::NT) {
                int sl = s + 1 + q;
                if (sl >= W) sl -= W;
                orow[q] = k + 1 + q < ni ? lc[sl] * inv : T(0);
            }
            for (int q = t; q <= bw; q += TM::NT) {
                int sl = s + q;
                if (sl >= W) sl -= W;
                orow[bw + q] = ur[sl];
            }
            for (int b = t; b < nbd; b += TM::NT) {
                orow[bw + W + b] = lc[W + b] * inv;
                orow[bw + W + nbd + b] = ur[W + b];
            }
        }
        // the recycled slot: zero, or the entering index's entries; each buffer entry has one
        // reader (the slot's owner), which clears it
        if (tr == sm) {
#pragma unroll
            for (int i = 0; i < kBandR; ++i)
                if (i == si)
#pragma unroll
                    for (int j = 0; j < kBandR; ++j) {
                        NUPPN_LDS T *e = er + tc + kBandTG * j;
                        w[i][j] = enter ? *e : T(0);
                        *e = T(0);
                    }
        }
        if (tc == sm) {
#pragma unroll
            for (int j = 0; j < kBandR; ++j)
                if (j == si)
#pragma unroll
                    for (int i = 0; i < kBandR; ++i)
                        if (!(tr == sm && i == si)) {
                            NUPPN_LDS T *e = ec + tr + kBandTG * i;
                            w[i][j] = enter ? *e : T(0);
                            *e = T(0);
                        }
        }
    }
    if (fail) {
        TM::sync();
        return 1;
    }
    TM::sync();
    if (ni > 0) {
        const int k0 = (ni - 1) / CHF * CHF;
        flush(k0, ni - k0);
    }
    // the border Schur complement to FP64, LU with partial pivoting (thread 0)
    NUPPN_GLB double *G = gptr(F.G);
#pragma unroll
    for (int i = 0; i < kBandR; ++i) {
        const int r = tr + kBandTG * i;
#pragma unroll
        for (int j = 0; j < kBandR; ++j) {
            const int c = tc + kBandTG * j;
            if (r >= W && r < WT && c >= W && c < WT) G[(r - W) + (c - W) * nbd] = (double) w[i][j];
        }
    }
    TM::sync();
    int gbad = 0;
    if (t == 0) {
        NUPPN_GLB int *gp = gptr(F.gpiv);
        for (int k = 0; k < nbd; ++k) {
            int pr = k;
            double best = fabs(G[k + k * nbd]);
            for (int r = k + 1; r < nbd; ++r)
                if (fabs(G[r + k * nbd]) > best) {
                    best = fabs(G[r + k * nbd]);
                    pr = r;
                }
            gp[k] = pr;
            if (!(isfinite(best) && best > 0.0)) {
                gbad = 1;
                break;
            }
            if (pr != k)
                for (int c = 0; c < nbd; ++c) {
                    const double x = G[k + c * nbd];
                    G[k + c * nbd] = G[pr + c * nbd];
                    G[pr + c * nbd] = x;
                }
            const double ip = 1.0 / G[k + k * nbd];
            for (int r = k + 1; r < nbd; ++r) {
                const double f = G[r + k * nbd] * ip;
                G[r + k * nbd] = f;
                for (int c = k + 1; c < nbd; ++c) G[r + c * nbd] -= f * G[k + c * nbd];
            }
        }
    }
    return team_any<TM>(gbad != 0, L.red) ? 1 : 0;
}

// x = A^-1 b with the band factors (b, x: FP64 [n], reduced order, mass-fraction units)
template <class T>
__device__ __forceinline__ void band_solve(const DevBlock &B, const DevBand &D,
                                           const BandFactors<T> &F, const BandLds<T> &L,
                                           const double *bp, double *xp) {
    using TM = BandTeam;
    constexpr int NT = TM::NT, CH = kBandChunk;
    const NUPPN_GLB double *b = gptr(bp), *rowinv = gptr(F.rowinv), *colsc = gptr(F.colsc);
    const NUPPN_GLB int *vpos = gptr(B.vpos);
    NUPPN_GLB double *x = gptr(xp);
    const NUPPN_GLB T *Lb = gptr(F.Lb), *Uc = gptr(F.Uc), *LE = gptr(F.LE), *UF = gptr(F.UF);
    const int t0 = TM::tid(), ni = D.ni, nbd = D.nbd, bw = D.bw, W = D.W;
    NUPPN_LDS T *v = L.v;
    // the power-of-two scaling of block_solve: the largest entry in [1, 2)
    double m = 0.0;
    for (int c = t0; c < B.n; c += NT) m = fmax(m, fabs(b[c] * rowinv[vpos[c]]));
    m = team_max<TM>(m, L.red);
    const int ex = (m > 0.0 && isfinite(m)) ? ilogb(m) : 0;
    const double sc = ldexp(1.0, -ex), unsc = ldexp(1.0, ex);
    for (int c = t0; c < B.n; c += NT) v[vpos[c]] = (T) (b[c] * rowinv[vpos[c]] * sc);
    // forward: y_k final at step k, then v[k + 1 + q] -= L[k + 1 + q][k] y_k; columns staged by
    // groups of CH (the group g + 1 loaded while the first wavefront works through group g)
    const int ng = (ni + CH - 1) / CH;
    // (a group is loaded by threads first, first + stride, ...: all of them for the first group,
    // the three wavefronts that do not sweep for the others)
    auto stage_l = [&](int g, int buf, int first, int stride) __attribute__((always_inline)) {
        const int k0 = g * CH, kn = min(CH, ni - k0);
        NUPPN_LDS T *st = L.stage + (size_t) buf * CH * W;
        for (int e = first; e < kn * bw; e += stride) st[e] = Lb[(size_t) k0 * bw + e];
    };
    if (ng > 0) stage_l(0, 0, t0, NT);
    TM::sync();
    // The sweeping wavefront keeps the W active entries in registers, lane l the window slots l
    // and l + 64 (interior index i in slot i mod W): each step reads y_k from its owner lane,
    // updates the other slots and recycles y_k's slot for index k + W (from LDS, still the
    // right-hand side there), so the step-to-step chain runs through registers, not LDS.
    const int l = lane_id();
    T r0 = T(0), r1 = T(0);
    if (TM::wave() == 0) {
        if (l < W && l < ni) r0 = v[l];
        if (l + WAVE < W && l + WAVE < ni) r1 = v[l + WAVE];
    }
    int s = 0;   // slot of index k
    for (int g = 0; g < ng; ++g) {
        if (g + 1 < ng && TM::wave() != 0) stage_l(g + 1, (g + 1) & 1, t0 - WAVE, NT - WAVE);
        if (TM::wave() == 0) {
            const NUPPN_LDS T *st = L.stage + (size_t) (g & 1) * CH * W;
            const int k0 = g * CH, kn = min(CH, ni - k0);
            for (int kk = 0; kk < kn; ++kk) {
                const int k = k0 + kk;
                int d0 = l - s, d1 = l + WAVE - s;   // slot distance ahead of the pivot
                if (d0 < 0) d0 += W;
                if (d1 < 0) d1 += W;
                const T a0 = (d0 > 0 && k + d0 < ni) ? st[kk * bw + d0 - 1] : T(0);
                const T a1 = (l + WAVE < W && d1 > 0 && k + d1 < ni) ? st[kk * bw + d1 - 1] : T(0);
                const T y = readlane_t(s < WAVE ? r0 : r1, s & (WAVE - 1));
                r0 = fma_t(-a0, y, r0);
                r1 = fma_t(-a1, y, r1);
                const T nv = k + W < ni ? v[k + W] : T(0);
                if (l == 0) v[k] = y;
                if (l == (s & (WAVE - 1))) {
                    if (s < WAVE) r0 = nv;
                    else r1 = nv;
                }
                if (++s == W) s = 0;
            }
        }
        TM::sync();
    }
    // the border: t = b_border - LE y, then G u = t with the row interchanges (thread 0)
    if (nbd > 0) {
        for (int bb = 0; bb < nbd; ++bb) {
            double s = 0.0;
            for (int k = t0; k < ni; k += NT) s += (double) LE[(size_t) bb * ni + k] * (double) v[k];
            s = team_sum<TM>(s, L.red);
            if (t0 == 0) L.tb[bb] = (double) v[ni + bb] - s;
        }
        TM::sync();
        if (t0 == 0) {
            NUPPN_LDS double *tb = L.tb;
            for (int k = 0; k < nbd; ++k) {
                const int p = F.gpiv[k];
                if (p != k) { const double u = tb[k]; tb[k] = tb[p]; tb[p] = u; }
            }
            for (int k = 0; k < nbd; ++k)
                for (int r = k + 1; r < nbd; ++r) tb[r] -= F.G[r + k * nbd] * tb[k];
            for (int k = nbd - 1; k >= 0; --k) {
                tb[k] /= F.G[k + k * nbd];
                for (int r = 0; r < k; ++r) tb[r] -= F.G[r + k * nbd] * tb[k];
            }
            for (int k = 0; k < nbd; ++k) v[ni + k] = (T) tb[k];
        }
        TM::sync();
        for (int k = t0; k < ni; k += NT) {
            T s = v[k];
            for (int bb = 0; bb < nbd; ++bb) s = fma_t(-UF[(size_t) k * nbd + bb], v[ni + bb], s);
            v[k] = s;
        }
    }
    TM::sync();
    // backward: y_j = v[j] / U[j][j], then v[j - d] -= U[j - d][j] y_j; U's columns by groups
    // from the last
    auto stage_u = [&](int g, int buf, int first, int stride) __attribute__((always_inline)) {
        const int j1 = ni - g * CH, jn = min(CH, j1);
        NUPPN_LDS T *st = L.stage + (size_t) buf * CH * W;
        for (int e = first; e < jn * W; e += stride) st[e] = Uc[(size_t) (j1 - jn) * W + e];
    };
    if (ng > 0) stage_u(0, 0, t0, NT);
    TM::sync();
    // the same register window, descending: slot i mod W holds index i of j - W + 1 .. j
    r0 = T(0);
    r1 = T(0);
    if (TM::wave() == 0 && ni > 0) {
        const int top = ni - 1, ts = top % W;
        int dd0 = ts - l, dd1 = ts - l - WAVE;   // distance below the top index
        if (dd0 < 0) dd0 += W;
        if (dd1 < 0) dd1 += W;
        if (l < W && top - dd0 >= 0) r0 = v[top - dd0];
        if (l + WAVE < W && top - dd1 >= 0) r1 = v[top - dd1];
    }
    s = ni > 0 ? (ni - 1) % W : 0;   // slot of index j
    for (int g = 0; g < ng; ++g) {
        if (g + 1 < ng && TM::wave() != 0) stage_u(g + 1, (g + 1) & 1, t0 - WAVE, NT - WAVE);
        if (TM::wave() == 0) {
            const NUPPN_LDS T *st = L.stage + (size_t) (g & 1) * CH * W;
            const int j1 = ni - g * CH, jn = min(CH, j1), j0 = j1 - jn;
            for (int jj = jn - 1; jj >= 0; --jj) {
                const int j = j0 + jj;
                int d0 = s - l, d1 = s - l - WAVE;   // slot distance below the pivot
                if (d0 < 0) d0 += W;
                if (d1 < 0) d1 += W;
                const T a0 = (d0 > 0 && j - d0 >= 0) ? st[jj * W + d0] : T(0);
                const T a1 = (l + WAVE < W && d1 > 0 && j - d1 >= 0) ? st[jj * W + d1] : T(0);
                const T y = readlane_t(s < WAVE ? r0 : r1, s & (WAVE - 1)) / st[jj * W];
                r0 = fma_t(-a0, y, r0);
                r1 = fma_t(-a1, y, r1);
                const T nv = j - W >= 0 ? v[j - W] : T(0);
                if (l == 0) v[j] = y;
                if (l == (s & (WAVE - 1))) {
                    if (s < WAVE) r0 = nv;
                    else r1 = nv;
                }
                if (--s < 0) s = W - 1;
            }
        }
        TM::sync();
    }
    for (int c = t0; c < B.n; c += NT) x[c] = (double) v[vpos[c]] * colsc[vpos[c]] * unsc;
    TM::sync();
}

}  // namespace nuppn
