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
};

__host__ __device__ inline size_t band_lds_bytes(int n, int W, int nbd, size_t tsz) {
    size_t o = 64;
    o += 8 * kBandMaxWT * tsz;
    o = (o + 15) & ~(size_t) 15;
    o += (size_t) n * tsz;
    o = (o + 15) & ~(size_t) 15;
    o += 2 * (size_t) kBandChunk * W * tsz;
    o += 8 * (size_t) (nbd + 1);
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
    return L;
}

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
    // the entries of list li into the entering buffers of parity p (row part: slot row sr)
    auto scatter = [&](int li, int sr, int p) __attribute__((always_inline)) {
        for (int e = eptr[li] + t; e < eptr[li + 1]; e += TM::NT) {
            const T val = (T) (A[ek[e]] * eys[e] * cs[ecolx[e]]);
            const int dst = edst[e], r = dst / WT, c = dst - r * WT;
            if (r == sr) L.erow[p * kBandMaxWT + c] = val;
            else L.ecol[p * kBandMaxWT + r] = val;
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
    // index k + W enters in its slot
    const int nfill = min(W, ni);
    NUPPN_GLB T *Lb = gptr(F.Lb), *Uc = gptr(F.Uc), *LE = gptr(F.LE), *UF = gptr(F.UF);
    int fail = 0;
    for (int k = -nfill; k < ni; ++k) {
        const int p = k & 1;
        const int s = k < 0 ? k + nfill : k % W;         // slot of the pivot (or of the entry)
        const int rin = k < 0 ? k + nfill : k + W;      // index entering at this step
        const bool enter = rin < ni;
        NUPPN_LDS T *ur = L.urow + p * kBandMaxWT, *lc = L.lcol + p * kBandMaxWT;
        NUPPN_LDS T *er = L.erow + p * kBandMaxWT, *ec = L.ecol + p * kBandMaxWT;
        if (k >= 0) {
#pragma unroll
            for (int i = 0; i < kBandR; ++i)
                if (tr + kBandTG * i == s)
#pragma unroll
                    for (int j = 0; j < kBandR; ++j) ur[tc + kBandTG * j] = w[i][j];
#pragma unroll
            for (int j = 0; j < kBandR; ++j)
                if (tc + kBandTG * j == s)
#pragma unroll
                    for (int i = 0; i < kBandR; ++i) lc[tr + kBandTG * i] = w[i][j];
        }
        if (enter) scatter(rin + 1, rin % W, p);
        TM::sync();
        if (k >= 0) {
            const T piv = ur[s];
            if (!(isfinite((double) piv) && piv != T(0))) {   // uniform: every thread reads it
                fail = 1;
                break;
            }
            const T inv = T(1) / piv;
            T l[kBandR], u[kBandR];
#pragma unroll
            for (int i = 0; i < kBandR; ++i) {
                const int r = tr + kBandTG * i;
                l[i] = r == s ? T(0) : lc[r] * inv;
            }
#pragma unroll
            for (int j = 0; j < kBandR; ++j) {
                const int c = tc + kBandTG * j;
                u[j] = c == s ? T(0) : ur[c];
            }
#pragma unroll
            for (int i = 0; i < kBandR; ++i)
#pragma unroll
                for (int j = 0; j < kBandR; ++j) w[i][j] = fma_t(-l[i], u[j], w[i][j]);
            // L's column k, U's row k (column-band storage), the border parts
            for (int q = t; q < bw; q += TM::NT) {
                const int r = k + 1 + q;
                Lb[(size_t) k * bw + q] = r < ni ? lc[r % W] * inv : T(0);
            }
            for (int q = t; q <= bw; q += TM::NT) {
                const int j = k + q;
                if (j < ni) Uc[(size_t) j * W + q] = ur[j % W];
            }
            for (int b = t; b < nbd; b += TM::NT) {
                LE[(size_t) b * ni + k] = lc[W + b] * inv;
                UF[(size_t) k * nbd + b] = ur[W + b];
            }
        }
        // the recycled slot: zero, or the entering index's entries
#pragma unroll
        for (int i = 0; i < kBandR; ++i) {
            const int r = tr + kBandTG * i;
#pragma unroll
            for (int j = 0; j < kBandR; ++j) {
                const int c = tc + kBandTG * j;
                if (r == s) {
                    w[i][j] = enter ? er[c] : T(0);
                } else if (c == s) {
                    w[i][j] = enter ? ec[r] : T(0);
                }
            }
        }
        // each buffer entry has one reader (the slot's owner): it clears what it read
        if (enter) {
#pragma unroll
            for (int i = 0; i < kBandR; ++i)
                if (tr + kBandTG * i == s)
#pragma unroll
                    for (int j = 0; j < kBandR; ++j) er[tc + kBandTG * j] = T(0);
#pragma unroll
            for (int j = 0; j < kBandR; ++j)
                if (tc + kBandTG * j == s)
#pragma unroll
                    for (int i = 0; i < kBandR; ++i)
                        if (tr + kBandTG * i != s) ec[tr + kBandTG * i] = T(0);
        }
    }
    if (fail) {
        TM::sync();
        return 1;
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
    for (int g = 0; g < ng; ++g) {
        if (g + 1 < ng && TM::wave() != 0) stage_l(g + 1, (g + 1) & 1, t0 - WAVE, NT - WAVE);
        if (TM::wave() == 0) {
            const NUPPN_LDS T *st = L.stage + (size_t) (g & 1) * CH * W;
            const int k0 = g * CH, kn = min(CH, ni - k0);
            const int l = lane_id();
            for (int kk = 0; kk < kn; ++kk) {
                const int k = k0 + kk;
                const T y = v[k];
                lds_fence();
                for (int q = l; q < bw; q += WAVE)
                    if (k + 1 + q < ni) v[k + 1 + q] = fma_t(-st[kk * bw + q], y, v[k + 1 + q]);
                lds_fence();
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
    for (int g = 0; g < ng; ++g) {
        if (g + 1 < ng && TM::wave() != 0) stage_u(g + 1, (g + 1) & 1, t0 - WAVE, NT - WAVE);
        if (TM::wave() == 0) {
            const NUPPN_LDS T *st = L.stage + (size_t) (g & 1) * CH * W;
            const int j1 = ni - g * CH, jn = min(CH, j1), j0 = j1 - jn;
            const int l = lane_id();
            for (int jj = jn - 1; jj >= 0; --jj) {
                const int j = j0 + jj;
                const T y = v[j] / st[jj * W];
                lds_fence();
                for (int d = 1 + l; d <= bw; d += WAVE)
                    if (j - d >= 0) v[j - d] = fma_t(-st[jj * W + d], y, v[j - d]);
                if (l == 0) v[j] = y;
                lds_fence();
            }
        }
        TM::sync();
    }
    for (int c = t0; c < B.n; c += NT) x[c] = (double) v[vpos[c]] * colsc[vpos[c]] * unsc;
    TM::sync();
}

}  // namespace nuppn
