// Pivot-free block-tridiagonal factorisation of the reduced system I - dt J on one wavefront,
// used by the batched path (FP32 factors with FP64 refinement, or FP64 factors).
//
// Why no pivoting: in molar units (Y = X/A) the interior part of I - dt J (every species but the
// dense light projectiles) is a column diagonally dominant Z-matrix: each column holds 1 + dt
// times the destruction rate on the diagonal and minus dt times the production rates elsewhere,
// and mass conservation bounds the latter by the former.  Gaussian elimination without pivoting
// is backward stable on such matrices (growth factor <= 2).  The reference's threshold pivoting
// (SuperLU_5.0/SRC/dpivotL.c:114-156) instead works in mass-fraction units, where the column
// ratios are off by A_row / A_col and off-diagonal pivots are common; both give a backward-stable
// solve, and this path's FP64 residual refinement compares them on equal footing.
//
// Layout (host-built, BlockSym): interior species in network order, split into nblk blocks of
// NB; border species (nbd) last.  The matrix is scaled to molar units and its columns by the
// reciprocal diagonal (values then lie in [-2, 2] with unit diagonal, inside FP32 range), and
//     A = [ T  F ]      T block tridiagonal: diagonal D_i, super-diagonal C_i, sub-diagonal B_i
//         [ E  G ]
// is factored as
//     S_0 = D_0,  W_i = S_i^-1 C_i,  S_{i+1} = D_{i+1} - B_i W_i          (block Thomas)
//     X_F = T^-1 F,  G' = G - E X_F                                       (border Schur complement)
// storing S_i^-1 (explicit, Gauss-Jordan), W_i and B_i per block, X_F, E and the LU of G'.
// A solve is then two sweeps of dense NB x NB mat-vecs plus the border correction.
#pragma once
#include "device_common.hpp"
#include "symbolic.hpp"

namespace nuppn {

// a SegStream on the device (the host builder's comment has the layout), physical order
struct DevSeg {
    int nseg;
    const int *e_idx;       // operand index, bit 31 on a segment's last term (J: with the count)
    const double *e_coef;   // coefficients (null for J)
    const int *round_base;  // [kSegRounds + 1]
    const int *lane_len;    // [64 kSegRounds]
    const int *lane_s0;     // [64 kSegRounds]
};

struct DevBlock {
    int NB, nbd, ni, nblk, nip, n, nnzA;
    const int *vpos;       // [n] reduced index -> padded vector slot
    const double *mass;    // [n]
    const int *diag_a;     // [n] position of each column's diagonal in the A value array
    const int *ent_ptr, *ent_a, *ent_loc, *ent_col;   // ent_a: position in the A value array
    const int *ent_dst;    // D/C/B entries: [area][block][NB*NB] destination (area 0 D, 1 C, 2 B)
    const double *ent_ys;
    const int *csr_col;    // [nnzA] column of each A entry (A kept in CSR order)
    const int *csr_row;    // [nnzA] row of each A entry
    const int *dyn_k;      // [8 nre] the dynamic network's entry map (DevNet::dyn_q) in CSR order
    const unsigned char *csr_isdiag;
    const int *resid_pos;  // [nnzA] physical position of each A entry in the residual stream
    const double *jac_rhi, *jac_rlo;   // [nnzA] the mass-number ratio of each J entry, hi + lo
    DevSeg jac, dxdt, resid;
};

// factors of one zone slot; T is the factor precision
template <class T>
struct BlockFactors {
    T *Sinv, *W, *Bm;   // [nblk][NB*NB] column-major
    T *E, *XF;          // [nbd][nip]
    double *G;          // [nbd*nbd] LU of G' (column-major), unit lower part below the diagonal
    int *gpiv;          // [nbd] row interchanges of the LU of G'
    double *cs;         // [n] 1 / diagonal of A: the column scaling
    double *colsc;      // [nip + nbd] by padded position: mass * column scaling
    double *rowinv;     // [nip + nbd] by padded position: 1 / mass
    // the dense blocks of the scaled matrix (D_i, C_i, B_i: [nblk][NB*NB] column-major, one set
    // per precision): written at the pattern's positions before every factorisation and read by
    // it and by the solves (B_i, C_i); the other positions hold zeros from the allocation
    T *Din, *Cin, *Bin;
};

template <class T, int NB>
struct BlockCfg {
    static constexpr int H = WAVE / NB;   // lanes per block column
    static constexpr int R = NB / H;      // rows of a block column held by one lane
    static constexpr int LD = NB + 4;     // LDS leading dimension (16-byte aligned, conflict-free)
};

// LDS of one zone: the two block-product operands X1, X2 of the factorisation loop (X2 also
// holds the Gauss-Jordan pivot columns and rows between the products), which the solve vector
// vec overlays (the loop and the solves never run at the same time), the border scratch, the
// team reduction slots and, for a zone of several wavefronts, a second operand area X1b, X2b
// (the bottom-up sweep of the twisted factorisation; later the border columns of the second
// wavefront pair, from vecb).  About 18.9 KB with FP64 factors and one wavefront (NB = 32,
// 1,100 species), so eight single-wavefront zones share a CU's 160 KB.
template <class T>
struct BlockLds {
    NUPPN_LDS T *X1, *X2, *vec;
    NUPPN_LDS double *bscr;   // [16] border solve scratch
    NUPPN_LDS double *red;    // [16] team reductions
    NUPPN_LDS T *X1b, *X2b, *vecb;
};

template <class T, int NB>
__host__ __device__ constexpr size_t block_lds_span(int nip) {
    return 2 * (size_t) NB * BlockCfg<T, NB>::LD > (size_t) nip + 8
               ? 2 * (size_t) NB * BlockCfg<T, NB>::LD
               : (size_t) nip + 8;
}

template <class T, int NB>
__host__ __device__ constexpr size_t block_lds_bytes(int nip, int nbd, int nw = 1) {
    (void) nbd;   // nbd <= 8: the border entries fit the 8 slots after nip
    return sizeof(T) * block_lds_span<T, NB>(nip) * (nw > 1 ? 2 : 1) + 16 + 32 * sizeof(double);
}

template <class T, int NB>
__device__ inline BlockLds<T> make_block_lds(NUPPN_LDS char *base, int nip) {
    using C = BlockCfg<T, NB>;
    BlockLds<T> L;
    NUPPN_LDS T *p = (NUPPN_LDS T *) base;
    L.X1 = p;
    L.X2 = p + NB * C::LD;
    L.vec = p;
    p += block_lds_span<T, NB>(nip);
    NUPPN_LDS char *q = (NUPPN_LDS char *) p;
    q = (NUPPN_LDS char *) (((uintptr_t) q + 15) & ~(uintptr_t) 15);
    L.bscr = (NUPPN_LDS double *) q;
    L.red = L.bscr + 16;
    L.X1b = (NUPPN_LDS T *) (L.red + 16);
    L.X2b = L.X1b + NB * C::LD;
    L.vecb = L.X1b;
    return L;
}

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c) { return fma(a, b, c); }

// 1/x for a pivot of the column-scaled matrix (magnitude near 1, never denormal): the hardware
// reciprocal refined by Newton-Raphson steps to the precision of T, without the IEEE division's
// range fix-ups
__device__ __forceinline__ double pivot_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ __forceinline__ float pivot_rcp(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// the R elements of a block column held in registers by one lane
template <class T, int NB>
struct Regs {
    T v[BlockCfg<T, NB>::R];
};

// --- lane-layout helpers: lane (c = lane % NB, h = lane / NB) holds rows h*R .. h*R+R-1 of
//     block column c
template <class T, int NB>
__device__ __forceinline__ void lds_load_cm(const NUPPN_LDS T *X, Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) a.v[t] = X[c * C::LD + h * C::R + t];
}

template <class T, int NB>
__device__ __forceinline__ void lds_store_cm(NUPPN_LDS T *X, const Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) X[c * C::LD + h * C::R + t] = a.v[t];
}

template <class T, int NB>
__device__ __forceinline__ void lds_store_rm(NUPPN_LDS T *X, const Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) X[(h * C::R + t) * C::LD + c] = a.v[t];
}

template <class T, int NB>
__device__ __forceinline__ void glb_load_cm(const T *Gp, Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    const NUPPN_GLB T *G = gptr(Gp);
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) a.v[t] = G[c * NB + h * C::R + t];
}

template <class T, int NB>
__device__ __forceinline__ void glb_store_cm(T *Gp, const Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    NUPPN_GLB T *G = gptr(Gp);
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) G[c * NB + h * C::R + t] = a.v[t];
}

// In-place inverse of the NB x NB block held in lane layout: Gauss-Jordan without pivoting, on
// 2 x 2 pivot blocks K = {k, k+1} (the leading 2 x 2 minors of these diagonally dominant blocks
// are far from singular).  One step, with P = a[K][K]:
//     a[i][j] -= a[i][K] P^-1 a[K][j]  (i, j not in K),  a[i][K] = -a[i][K] P^-1,
//     a[K][j] = P^-1 a[K][j],  a[K][K] = P^-1,
// the same as two scalar steps in exact arithmetic, with half the LDS round trips.  buf: 4 NB
// entries of LDS (the pivot columns and rows).
template <class T, int NB>
__device__ __forceinline__ void gj_inverse(Regs<T, NB> &a, NUPPN_LDS T *buf) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
    NUPPN_LDS T *col0 = buf, *col1 = buf + NB, *row0 = buf + 2 * NB, *row1 = buf + 3 * NB;
    constexpr int KU = NB <= 32 ? NB / 2 : 2;
#pragma unroll KU
    for (int k = 0; k < NB; k += 2) {
        const int kh = k / C::R, kt = k % C::R;   // rows k, k + 1 lie in the same lane half
        if (c == k || c == k + 1) {
            NUPPN_LDS T *cb = c == k ? col0 : col1;
#pragma unroll
            for (int t = 0; t < C::R; ++t) cb[h * C::R + t] = a.v[t];
        }
        if (h == kh) {
            T r0 = a.v[0], r1 = a.v[1];
#pragma unroll
            for (int t = 2; t < C::R; t += 2)
                if (t == kt) {
                    r0 = a.v[t];
                    r1 = a.v[t + 1];
                }
            row0[c] = r0;
            row1[c] = r1;
        }
        lds_fence();
        // the pivot columns' entries of this lane's rows, all loads issued at once and before the
        // pivot block's inverse, whose latency they overlap (one LDS round trip per step)
        T x0[C::R], x1[C::R];
#pragma unroll
        for (int t = 0; t < C::R; ++t) {
            x0[t] = col0[h * C::R + t];
            x1[t] = col1[h * C::R + t];
        }
        // P^-1 = [[p11, -p01], [-p10, p00]] / det
        const T p00 = col0[k], p10 = col0[k + 1], p01 = col1[k], p11 = col1[k + 1];
        const T dinv = pivot_rcp(p00 * p11 - p01 * p10);
        const T q00 = p11 * dinv, q01 = -p01 * dinv, q10 = -p10 * dinv, q11 = p00 * dinv;
        const T u0 = row0[c], u1 = row1[c];
        // (f0, f1): the new rows K of column c; columns outside K also take the rank-2 update
        const bool inK = c == k || c == k + 1;
        T f0, f1;
        if (c == k) {
            f0 = q00;
            f1 = q10;
        } else if (c == k + 1) {
            f0 = q01;
            f1 = q11;
        } else {
            f0 = fma_t(q00, u0, q01 * u1);
            f1 = fma_t(q10, u0, q11 * u1);
        }
#pragma unroll
        for (int t = 0; t < C::R; ++t)
            a.v[t] = fma_t(-x0[t], f0, fma_t(-x1[t], f1, inK ? T(0) : a.v[t]));
        if (h == kh) {
#pragma unroll
            for (int t = 0; t < C::R; t += 2)
                if (t == kt) {
                    a.v[t] = f0;
                    a.v[t + 1] = f1;
                }
        }
        lds_fence();   // every read of this step precedes the next step's writes
    }
}

// acc = A * B with A column-major and B row-major in LDS (NB x NB), result in lane layout
template <class T, int NB>
__device__ __forceinline__ void gemm_lds(const NUPPN_LDS T *Acm, const NUPPN_LDS T *Brm,
                                Regs<T, NB> &acc) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int t = 0; t < C::R; ++t) acc.v[t] = T(0);
#pragma unroll 2
    for (int k = 0; k < NB; ++k) {
        const T bk = Brm[k * C::LD + c];
        const NUPPN_LDS T *ap = Acm + k * C::LD + h * C::R;
#pragma unroll
        for (int t = 0; t < C::R; ++t) acc.v[t] = fma_t(ap[t], bk, acc.v[t]);
    }
}

// acc -= A * B (same layouts as gemm_lds)
template <class T, int NB>
__device__ __forceinline__ void gemm_lds_sub(const NUPPN_LDS T *Acm, const NUPPN_LDS T *Brm,
                                    Regs<T, NB> &acc) {
    using C = BlockCfg<T, NB>;
    const int c = lane_id() % NB, h = lane_id() / NB;
#pragma unroll 2
    for (int k = 0; k < NB; ++k) {
        const T bk = Brm[k * C::LD + c];
        const NUPPN_LDS T *ap = Acm + k * C::LD + h * C::R;
#pragma unroll
        for (int t = 0; t < C::R; ++t) acc.v[t] = fma_t(-ap[t], bk, acc.v[t]);
    }
}

// --- 32 x 32 block products on the matrix cores (NB = 32): 2 x 2 tiles of 16 x 16, k in steps of
// 4 (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32).  Lane l supplies A[16 ti + l % 16][4 kk + l / 16]
// and B[4 kk + l / 16][16 tj + l % 16]; result register r of lane l is column l % 16 of the tile,
// row l / 16 + 4 r with FP64 and 4 (l / 16) + r with FP32 (the two instructions' layouts, checked
// on the MI355X).
typedef double mf64x4 __attribute__((ext_vector_type(4)));
typedef float mf32x4 __attribute__((ext_vector_type(4)));
template <class T>
struct MVec;
template <>
struct MVec<double> {
    typedef mf64x4 type;
};
template <>
struct MVec<float> {
    typedef mf32x4 type;
};

// the four result tiles (ti, tj) of a 32 x 32 product, tile 2 ti + tj
template <class T>
struct MTiles {
    typename MVec<T>::type t[4];
};

__device__ __forceinline__ mf64x4 mfma16(double a, double b, mf64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mf32x4 mfma16(float a, float b, mf32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T>
__device__ __forceinline__ int mrow(int l, int r) {
    return sizeof(T) == 8 ? (l >> 4) + 4 * r : 4 * (l >> 4) + r;
}

// f(tile, r, row, col) for every result register of the lane
template <class T, class F>
__device__ __forceinline__ void mtile_for(F f) {
    const int l = lane_id();
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) f(2 * ti + tj, r, 16 * ti + mrow<T>(l, r), 16 * tj + (l & 15));
}

// acc += sgn A B with A column-major and B row-major in LDS (leading dimension LD), sgn = -1
// when negate
template <class T, int LD>
__device__ __forceinline__ void mfma_gemm32(const NUPPN_LDS T *Acm, const NUPPN_LDS T *Brm,
                                            MTiles<T> &acc, bool negate) {
    const int l = lane_id(), i = l & 15, k = l >> 4;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const int row = (4 * kk + k) * LD;
        T a0 = Acm[row + i], a1 = Acm[row + 16 + i];
        const T b0 = Brm[row + i], b1 = Brm[row + 16 + i];
        if (negate) {
            a0 = -a0;
            a1 = -a1;
        }
        acc.t[0] = mfma16(a0, b0, acc.t[0]);
        acc.t[1] = mfma16(a0, b1, acc.t[1]);
        acc.t[2] = mfma16(a1, b0, acc.t[2]);
        acc.t[3] = mfma16(a1, b1, acc.t[3]);
    }
}

// a 32 x 32 column-major block in global memory to / from the tile layout
template <class T>
__device__ __forceinline__ void mtile_load_glb(const T *Gp, MTiles<T> &M) {
    const NUPPN_GLB T *G = gptr(Gp);
    mtile_for<T>([&](int q, int r, int row, int col) { M.t[q][r] = G[col * 32 + row]; });
}
template <class T, int LD>
__device__ __forceinline__ void mtile_store_lds_cm(NUPPN_LDS T *X, const MTiles<T> &M) {
    mtile_for<T>([&](int q, int r, int row, int col) { X[col * LD + row] = M.t[q][r]; });
}
template <class T, int LD>
__device__ __forceinline__ void mtile_store_lds_rm(NUPPN_LDS T *X, const MTiles<T> &M) {
    mtile_for<T>([&](int q, int r, int row, int col) { X[row * LD + col] = M.t[q][r]; });
}

// scaled value of CSC entry e of a block list: molar units, column-scaled
__device__ __forceinline__ double block_entry(const DevBlock &B, const double *Aval,
                                              const double *cs, int e) {
    return gptr(Aval)[gptr(B.ent_a)[e]] * gptr(B.ent_ys)[e] * gptr(cs)[gptr(B.ent_col)[e]];
}

// lane (rr = lane % NB, h = lane / NB) part of an NB x NB column-major block for matvec_regs
template <class T, int NB>
__device__ __forceinline__ void load_mv(const T *__restrict__ Mp, T (&m)[NB / BlockCfg<T, NB>::H]) {
    constexpr int CPL = NB / BlockCfg<T, NB>::H;
    const NUPPN_GLB T *M = gptr(Mp);
    const int rr = lane_id() % NB, h = lane_id() / NB;
#pragma unroll
    for (int j = 0; j < CPL; ++j) m[j] = M[(h * CPL + j) * NB + rr];
}

// row rr of M x with M's lane part in registers and x in LDS (full sum in both halves)
template <class T, int NB>
__device__ __forceinline__ T matvec_regs(const T (&m)[NB / BlockCfg<T, NB>::H],
                                         const NUPPN_LDS T *x) {
    constexpr int CPL = NB / BlockCfg<T, NB>::H;
    const int h = lane_id() / NB;
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc = fma_t(m[j], x[h * CPL + j], acc);
    if (BlockCfg<T, NB>::H == 2) acc += __shfl_xor(acc, 32, WAVE);
    return acc;
}

// v_q <- T^-1 v_q (q < NV) for the interior block-tridiagonal part (v_q in LDS at v + q vs, nip
// entries each).  The block matrices do not depend on v, so each sweep loads the next block's
// matrices before it works on the current block: the global-memory latency overlaps the
// dependent chain of mat-vecs; with NV = 2 each block's matrices serve both vectors.
template <class T, int NB, int NV = 1>
__device__ __forceinline__ void thomas_solve(const DevBlock &B, const BlockFactors<T> &F,
                                             NUPPN_LDS T *v, int vs = 0) {
    constexpr int CPL = NB / BlockCfg<T, NB>::H;
    const int rr = lane_id() % NB, h = lane_id() / NB;
    constexpr size_t BS = (size_t) NB * NB;
    const int nblk = B.nblk;
    T mS[CPL], mB[CPL], nS[CPL], nB[CPL];
    load_mv<T, NB>(F.Sinv, mS);
#pragma unroll
    for (int j = 0; j < CPL; ++j) mB[j] = T(0);
    for (int i = 0; i < nblk; ++i) {
        if (i + 1 < nblk) {
            load_mv<T, NB>(F.Bin + (size_t) i * BS, nB);
            load_mv<T, NB>(F.Sinv + (size_t) (i + 1) * BS, nS);
        }
        if (i > 0) {
            T acc[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) acc[q] = matvec_regs<T, NB>(mB, v + q * vs + (i - 1) * NB);
            lds_fence();
            if (h == 0)
#pragma unroll
                for (int q = 0; q < NV; ++q) v[q * vs + i * NB + rr] = v[q * vs + i * NB + rr] - acc[q];
            lds_fence();
        }
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = matvec_regs<T, NB>(mS, v + q * vs + i * NB);
        lds_fence();
        if (h == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) v[q * vs + i * NB + rr] = acc[q];
        lds_fence();
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            mS[j] = nS[j];
            mB[j] = nB[j];
        }
    }
    if (nblk < 2) return;
    load_mv<T, NB>(F.W + (size_t) (nblk - 2) * BS, mS);
    for (int i = nblk - 2; i >= 0; --i) {
        if (i > 0) load_mv<T, NB>(F.W + (size_t) (i - 1) * BS, nS);
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = matvec_regs<T, NB>(mS, v + q * vs + (i + 1) * NB);
        lds_fence();
        if (h == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) v[q * vs + i * NB + rr] = v[q * vs + i * NB + rr] - acc[q];
        lds_fence();
#pragma unroll
        for (int j = 0; j < CPL; ++j) mS[j] = nS[j];
    }
}

// the middle block of the twisted factorisation: top-down above it, bottom-up below it
__device__ __forceinline__ int twist_block(int nblk) { return nblk / 2; }

// v_q <- T^-1 v_q (q < nv <= NV) with the twisted factors of block_factor (several
// wavefronts): a pair of wavefronts, role 0 the rows above the middle block m (forward
// v_i = S_i^-1 (v_i - B_{i-1} v_{i-1}), back v_i -= W_i v_{i+1}), role 1 the rows below it
// (forward v_i = R_i^-1 (v_i - C_i v_{i+1}), back v_i -= V_i v_{i-1}), the middle block
// v_m = Z_m^-1 (v_m - B_{m-1} v_{m-1} - C_m v_{m+1}) by role 0 in between.  Every wavefront of
// the zone calls it (the three barriers pair across the workgroup); active false: a wavefront
// that only takes part in the barriers.
template <class T, int NB, int NV, class TM>
__device__ __forceinline__ void twisted_solve(const DevBlock &B, const BlockFactors<T> &F,
                                              NUPPN_LDS T *v, int vs, int nv, int role,
                                              bool active) {
    constexpr int CPL = NB / BlockCfg<T, NB>::H;
    constexpr size_t BS = (size_t) NB * NB;
    const int rr = lane_id() % NB, h = lane_id() / NB;
    const int N = B.nblk, m = twist_block(N);
    // v_i <- v_i - M x (x = block j of each vector) or v_i <- M v_i
    auto sub = [&](const T (&M)[CPL], int i, int j) {
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q)
            acc[q] = q < nv ? matvec_regs<T, NB>(M, v + q * vs + j * NB) : T(0);
        lds_fence();
        if (h == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q)
                if (q < nv) v[q * vs + i * NB + rr] = v[q * vs + i * NB + rr] - acc[q];
        lds_fence();
    };
    auto mul = [&](const T (&M)[CPL], int i) {
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q)
            acc[q] = q < nv ? matvec_regs<T, NB>(M, v + q * vs + i * NB) : T(0);
        lds_fence();
        if (h == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q)
                if (q < nv) v[q * vs + i * NB + rr] = acc[q];
        lds_fence();
    };
    auto copy = [&](T (&d)[CPL], const T (&src)[CPL]) {
#pragma unroll
        for (int j = 0; j < CPL; ++j) d[j] = src[j];
    };
    T mS[CPL], mO[CPL], nS[CPL], nO[CPL];
    if (active && role == 0 && m > 0) {
        load_mv<T, NB>(F.Sinv, mS);
        for (int i = 0; i < m; ++i) {
            if (i + 1 < m) load_mv<T, NB>(F.Sinv + (size_t) (i + 1) * BS, nS);
            load_mv<T, NB>(F.Bin + (size_t) i * BS, nO);                    // B_i, next step
            if (i > 0) sub(mO, i, i - 1);
            mul(mS, i);
            copy(mS, nS);
            copy(mO, nO);
        }
    } else if (active && role == 1 && N - 1 > m) {
        load_mv<T, NB>(F.Sinv + (size_t) (N - 1) * BS, mS);
        for (int i = N - 1; i > m; --i) {
            if (i - 1 > m) load_mv<T, NB>(F.Sinv + (size_t) (i - 1) * BS, nS);
            load_mv<T, NB>(F.Cin + (size_t) (i - 1) * BS, nO);              // C_{i-1}, next
            if (i < N - 1) sub(mO, i, i + 1);
            mul(mS, i);
            copy(mS, nS);
            copy(mO, nO);
        }
    }
    wave_fence();
    TM::sync();
    if (active && role == 0) {
        if (m > 0) {
            load_mv<T, NB>(F.Bin + (size_t) (m - 1) * BS, mO);
            sub(mO, m, m - 1);
        }
        if (m < N - 1) {
            load_mv<T, NB>(F.Cin + (size_t) m * BS, mO);
            sub(mO, m, m + 1);
        }
        load_mv<T, NB>(F.Sinv + (size_t) m * BS, mS);
        mul(mS, m);
        wave_fence();
    }
    TM::sync();
    if (active && role == 0 && m > 0) {
        load_mv<T, NB>(F.W + (size_t) (m - 1) * BS, mO);
        for (int i = m - 1; i >= 0; --i) {
            if (i > 0) load_mv<T, NB>(F.W + (size_t) (i - 1) * BS, nO);
            sub(mO, i, i + 1);
            copy(mO, nO);
        }
    } else if (active && role == 1 && N - 1 > m) {
        load_mv<T, NB>(F.Bm + (size_t) m * BS, mO);                            // V_{m+1}
        for (int i = m + 1; i < N; ++i) {
            if (i + 1 < N) load_mv<T, NB>(F.Bm + (size_t) i * BS, nO);         // V_{i+1}
            sub(mO, i, i - 1);
            copy(mO, nO);
        }
    }
    wave_fence();
    TM::sync();
}

// mass fractions far below any reported abundance (1e-20) and below the reference's clipping
// of abundances to 1e-99 (y < 1e-27, utils.F90:726-740): the refinement tests measure every
// species against this floor at least
constexpr double kAbundanceFloor = 1.0e-30;

// the refinement's absolute floor: new abundances below it (far below the 1e-20 of any
// reported mass fraction and the 1e-27 under which ypsplus clips) need not converge relative
// to themselves
constexpr double kRefineFloor = 1.0e-25;

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// nsteps elimination steps of a block sweep on one wavefront, the dense blocks read from the
// input areas (D_i at Din[i], C_i at Cin[i], B_i at Bin[i]).  Step on block i, with a = the
// block's updated diagonal:
//   top-down (UP false):  a <- S_i^-1 -> Sinv[i];  W_i = S_i^-1 C_i -> W[i];
//                         a <- D_{i+1} - B_i W_i, next block i + 1;
//   bottom-up (UP true):  a <- R_i^-1 -> Sinv[i];  V_i = R_i^-1 B_{i-1} -> Bm[i-1];
//                         a <- D_{i-1} - C_{i-1} V_i, next block i - 1
// (on the last step with zero_last, a <- -C_{i-1} V_i: the middle block's share).  a holds the
// next block's updated diagonal on return.  The operands of a step's products load before its
// inversion, so their latency overlaps it.  Only wavefront-local ordering points, so in a
// multi-wavefront zone the other wavefronts may meanwhile run their own sweep or wait at their
// next barrier.  Returns 1 when an inverse is not finite.
template <class T, int NB, bool UP>
__device__ __forceinline__ int block_sweep(const BlockFactors<T> &F, NUPPN_LDS T *X1,
                                           NUPPN_LDS T *X2, int i0, int nsteps, bool zero_last,
                                           Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    constexpr size_t BS = (size_t) NB * NB;
    Regs<T, NB> cb, dn, bb;
    MTiles<T> dnt;   // NB = 32: the next diagonal block in the matrix-core tile layout
    for (int k = 0; k < nsteps; ++k) {
        const int i = UP ? i0 - k : i0 + k;
        const int nx = UP ? i - 1 : i + 1;                    // the next block
        T *m1 = UP ? F.Bm + (size_t) (i - 1) * BS : F.W + (size_t) i * BS;    // V_i / W_i out
        const T *i1 = UP ? F.Bin + (size_t) (i - 1) * BS : F.Cin + (size_t) i * BS;   // B_{i-1} / C_i
        const T *i2 = UP ? F.Cin + (size_t) (i - 1) * BS : F.Bin + (size_t) i * BS;   // C_{i-1} / B_i
        const bool zero = zero_last && k == nsteps - 1;
        glb_load_cm<T, NB>(i1, cb);
        glb_load_cm<T, NB>(i2, bb);
        if constexpr (NB == 32) {
            if (zero) {
#pragma unroll
                for (int q = 0; q < 4; ++q) dnt.t[q] = typename MVec<T>::type{0, 0, 0, 0};
            } else {
                mtile_load_glb<T>(F.Din + (size_t) nx * BS, dnt);
            }
        } else {
            if (zero) {
#pragma unroll
                for (int t = 0; t < C::R; ++t) dn.v[t] = T(0);
            } else {
                glb_load_cm<T, NB>(F.Din + (size_t) nx * BS, dn);
            }
        }
        gj_inverse<T, NB>(a, X2);   // X2 is free between the products
        glb_store_cm<T, NB>(F.Sinv + (size_t) i * BS, a);
        bool fin = true;
#pragma unroll
        for (int t = 0; t < C::R; ++t) fin = fin && isfinite((double) a.v[t]);
        if (ballot(!fin)) return 1;
        lds_fence();
        lds_store_cm<T, NB>(X1, a);
        lds_store_rm<T, NB>(X2, cb);
        lds_fence();
        if constexpr (NB == 32) {
            MTiles<T> w;   // the first product
#pragma unroll
            for (int q = 0; q < 4; ++q) w.t[q] = typename MVec<T>::type{0, 0, 0, 0};
            mfma_gemm32<T, C::LD>(X1, X2, w, false);
            // to global column-major through LDS (the tile layout scatters a lane's rows), and
            // row-major into X2 for the second product
            lds_fence();
            mtile_store_lds_cm<T, C::LD>(X1, w);
            mtile_store_lds_rm<T, C::LD>(X2, w);
            lds_fence();
            lds_load_cm<T, NB>(X1, a);
            glb_store_cm<T, NB>(m1, a);
            lds_fence();
            lds_store_cm<T, NB>(X1, bb);
            lds_fence();
            mfma_gemm32<T, C::LD>(X1, X2, dnt, true);                  // D - (second) (first)
            lds_fence();
            mtile_store_lds_cm<T, C::LD>(X1, dnt);
            lds_fence();
            lds_load_cm<T, NB>(X1, a);
        } else {
            gemm_lds<T, NB>(X1, X2, a);                              // the first product
            glb_store_cm<T, NB>(m1, a);
            lds_fence();
            lds_store_rm<T, NB>(X2, a);
            lds_store_cm<T, NB>(X1, bb);
            lds_fence();
#pragma unroll
            for (int t = 0; t < C::R; ++t) a.v[t] = dn.v[t];
            gemm_lds_sub<T, NB>(X1, X2, a);                          // D - (second) (first)
        }
    }
    wave_fence();
    return 0;
}

// The last (middle) block: a <- a^-1 -> Sinv[i]; 1 when not finite
template <class T, int NB>
__device__ __forceinline__ int block_last_inverse(const BlockFactors<T> &F, NUPPN_LDS T *X2,
                                                  int i, Regs<T, NB> &a) {
    using C = BlockCfg<T, NB>;
    gj_inverse<T, NB>(a, X2);
    glb_store_cm<T, NB>(F.Sinv + (size_t) i * NB * NB, a);
    bool fin = true;
#pragma unroll
    for (int t = 0; t < C::R; ++t) fin = fin && isfinite((double) a.v[t]);
    wave_fence();
    return ballot(!fin) ? 1 : 0;
}

// Factor the reduced matrix (FP64 CSC values Aval) into F with the zone's threads (TM).  Returns
// 0, or 1 when a pivot or the border Schur complement is not finite / zero (the caller falls back
// or reports).  The Thomas loop runs on the zone's first wavefront; with several wavefronts the
// dense-block set-up, the border columns' solves (one column per wavefront) and the Schur
// complement's sums are shared.
template <class T, int NB, class TM>
__device__ __forceinline__ int block_factor(const DevBlock &B, const double *Aval, const BlockFactors<T> &F,
                            const BlockLds<T> &L) {
    constexpr size_t BS = (size_t) NB * NB;
    constexpr int NT = TM::NT;
    const int t0 = TM::tid();
    const int nblk = B.nblk;
    bool bad = false;
    // (the loops below issue the loads of several of a thread's elements before their stores:
    // batched_for)
    batched_for<8>(
        t0, B.n, NT, [&](int i) { return gptr(Aval)[gptr(B.diag_a)[i]]; },
        [&](int i, double d) {
            if (!(fabs(d) > 0.0) || !isfinite(d)) bad = true;
            gptr(F.cs)[i] = 1.0 / d;
        });
    if (team_any<TM>(bad, L.red)) return 1;
    // scalings by padded position (interior blocks, then border): padding rows/columns stay 1
    for (int p = B.ni + t0; p < B.nip; p += NT) {
        gptr(F.colsc)[p] = 1.0;
        gptr(F.rowinv)[p] = 1.0;
    }
    struct Sc {
        double ms, c;
        int p;
    };
    batched_for<8>(
        t0, B.n, NT,
        [&](int i) { return Sc{gptr(B.mass)[i], gptr(F.cs)[i], gptr(B.vpos)[i]}; },
        [&](int i, const Sc &q) {
            (void) i;
            gptr(F.colsc)[q.p] = q.ms * q.c;
            gptr(F.rowinv)[q.p] = 1.0 / q.ms;
        });
    TM::sync();
    // dense blocks of the scaled matrix for all block steps at once (the loop then only streams
    // them, prefetched): the pattern's entries into the FP64 input areas, whose other positions
    // stay zero (nothing else writes these areas)
    {
        const size_t area = (size_t) nblk * BS;
        NUPPN_GLB T *D0 = gptr(F.Din), *C0 = gptr(F.Cin), *B0 = gptr(F.Bin);
        const int nent = gptr(B.ent_ptr)[3 * nblk - 2];
        struct En {
            double v;
            int dst;
        };
        batched_for<8>(
            t0, nent, NT, [&](int e) { return En{block_entry(B, Aval, F.cs, e), gptr(B.ent_dst)[e]}; },
            [&](int e, const En &q) {
                (void) e;
                const int ar = (int) (q.dst / area);
                NUPPN_GLB T *base = ar == 0 ? D0 : (ar == 1 ? C0 : B0);
                base[q.dst - ar * area] = (T) q.v;
            });
        const int pr = t0, last = nblk - 1;   // identity on the padded diagonal positions
        if (pr < NB && last * NB + pr >= B.ni) D0[(size_t) last * BS + pr * NB + pr] = T(1);
        TM::sync();
    }
    int fail = 0;
    if constexpr (TM::NWAVE == 1) {
        // block Thomas: S_0 = D_0, W_i = S_i^-1 C_i, S_{i+1} = D_{i+1} - B_i W_i
        Regs<T, NB> a;
        glb_load_cm<T, NB>(F.Din, a);
        fail = block_sweep<T, NB, false>(F, L.X1, L.X2, 0, nblk - 1, false, a);
        if (!fail) fail = block_last_inverse<T, NB>(F, L.X2, nblk - 1, a);
    } else {
        // twisted (two-ended) block factorisation: wavefront 0 eliminates top-down to the middle
        // block m, wavefront 1 bottom-up to it (R_{N-1} = D_{N-1}, V_i = R_i^-1 B_{i-1},
        // R_{i-1} = D_{i-1} - C_{i-1} V_i), and the middle block is
        // Z_m = D_m - B_{m-1} W_{m-1} - C_m V_{m+1}: half the dependent chain of block Thomas
        const int m = twist_block(nblk);
        Regs<T, NB> a;
        if (TM::wave() == 0) {
            glb_load_cm<T, NB>(F.Din, a);
            fail = block_sweep<T, NB, false>(F, L.X1, L.X2, 0, m, false, a);
        } else if (TM::wave() == 1 && nblk - 1 > m) {
            glb_load_cm<T, NB>(F.Din + (size_t) (nblk - 1) * BS, a);
            fail = block_sweep<T, NB, true>(F, L.X1b, L.X2b, nblk - 1, nblk - 1 - m, true, a);
            lds_store_cm<T, NB>(L.X1b, a);   // -C_m V_{m+1}
            wave_fence();
        }
        TM::sync();
        if (TM::wave() == 0 && !fail) {
            if (nblk - 1 > m) {
                Regs<T, NB> p;
                lds_load_cm<T, NB>(L.X1b, p);
#pragma unroll
                for (int t = 0; t < BlockCfg<T, NB>::R; ++t) a.v[t] = a.v[t] + p.v[t];
            }
            fail = block_last_inverse<T, NB>(F, L.X2, m, a);
        }
    }
    if (team_any<TM>(fail != 0, L.red)) return 1;
    // border: E, X_F = T^-1 F, G' = G - E X_F and its LU with partial pivoting
    const int nbd = B.nbd, nip = B.nip;
    if (nbd > 0) {
        const int LE = 3 * nblk - 2, LF = LE + 1, LG = LE + 2;
        NUPPN_GLB T *Eg = gptr(F.E), *XFg = gptr(F.XF);
        NUPPN_GLB double *Gg = gptr(F.G);
        const NUPPN_GLB int *eptr = gptr(B.ent_ptr), *eloc = gptr(B.ent_loc);
        for (int i = t0; i < nbd * nip; i += NT) {
            Eg[i] = T(0);
            XFg[i] = T(0);
        }
        for (int i = t0; i < nbd * nbd; i += NT) Gg[i] = 0.0;
        TM::sync();
        struct Ev {
            double v;
            int loc;
        };
        auto entries = [&](int list, auto put) {
            const int e0 = eptr[list];
            batched_for<8>(
                t0, eptr[list + 1] - e0, NT,
                [&](int j) { return Ev{block_entry(B, Aval, F.cs, e0 + j), eloc[e0 + j]}; },
                [&](int j, const Ev &q) {
                    (void) j;
                    put(q.loc, q.v);
                });
        };
        entries(LE, [&](int loc, double v) { Eg[loc] = (T) v; });
        entries(LF, [&](int loc, double v) { XFg[loc] = (T) v; });
        entries(LG, [&](int loc, double v) { Gg[loc] = v; });
        TM::sync();
        const int lane = lane_id();
        const int vs = nip + 8;
        if constexpr (TM::NWAVE == 1) {
            // X_F = T^-1 F, two columns per sweep where two vectors fit the LDS span
            const bool two = 2 * (size_t) vs <= block_lds_span<T, NB>(nip);
            for (int b = 0; b < nbd;) {
                const int nv = (two && b + 1 < nbd) ? 2 : 1;
                for (int q = 0; q < nv; ++q) {
                    NUPPN_GLB T *xf = XFg + (size_t) (b + q) * nip;
                    for (int j = lane; j < nip; j += WAVE) L.vec[q * vs + j] = xf[j];
                }
                lds_fence();
                if (nv == 2) thomas_solve<T, NB, 2>(B, F, L.vec, vs);
                else thomas_solve<T, NB>(B, F, L.vec);
                for (int q = 0; q < nv; ++q) {
                    NUPPN_GLB T *xf = XFg + (size_t) (b + q) * nip;
                    for (int j = lane; j < nip; j += WAVE) xf[j] = L.vec[q * vs + j];
                }
                lds_fence();
                b += nv;
            }
        } else {
            // X_F = T^-1 F with the twisted factors: wavefronts 0, 1 take two columns (one if two
            // vectors do not fit the LDS span), wavefronts 2, 3 the next column, round by round
            const bool two = 2 * (size_t) vs <= block_lds_span<T, NB>(nip);
            const int w = TM::wave(), pair = w >> 1, role = w & 1;
            const int nv0 = two ? 2 : 1, per_round = nv0 + (TM::NWAVE >= 4 ? 1 : 0);
            for (int b = 0; b < nbd; b += per_round) {
                const int c0 = pair == 0 ? b : b + nv0;
                const int nv = pair == 0 ? min(nv0, nbd - b) : (c0 < nbd ? 1 : 0);
                const bool act = pair < 2 && nv > 0;
                NUPPN_LDS T *v = pair == 0 ? L.vec : L.vecb;
                if (act && role == 0)
                    for (int q = 0; q < nv; ++q) {
                        NUPPN_GLB T *xf = XFg + (size_t) (c0 + q) * nip;
                        for (int j = lane; j < nip; j += WAVE) v[q * vs + j] = xf[j];
                    }
                TM::sync();
                twisted_solve<T, NB, 2, TM>(B, F, v, vs, act ? nv : 1, role, act);
                if (act && role == 0)
                    for (int q = 0; q < nv; ++q) {
                        NUPPN_GLB T *xf = XFg + (size_t) (c0 + q) * nip;
                        for (int j = lane; j < nip; j += WAVE) xf[j] = v[q * vs + j];
                    }
                TM::sync();
            }
        }
        TM::sync();
        if (nbd == 3) {
            // the nine sums of E X_F in one reduction (LDS scratch: the second operand area,
            // free once the border columns are solved; a single wavefront needs none)
            double s9[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) s9[q] = 0.0;
            for (int j = t0; j < nip; j += NT) {
                double e[3], x[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    e[q] = (double) Eg[(size_t) q * nip + j];
                    x[q] = (double) XFg[(size_t) q * nip + j];
                }
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
                    for (int b1 = 0; b1 < 3; ++b1) s9[3 * b2 + b1] += e[b1] * x[b2];
            }
            team_sum_k<TM, 9>(s9, (NUPPN_LDS double *) L.X1b);
            if (t0 == 0)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int b1 = 0; b1 < 3; ++b1) Gg[b1 + b2 * 3] -= s9[3 * b2 + b1];
        } else {
            for (int b2 = 0; b2 < nbd; ++b2)
                for (int b1 = 0; b1 < nbd; ++b1) {
                    double sacc = 0.0;
                    for (int j = t0; j < nip; j += NT)
                        sacc += (double) Eg[(size_t) b1 * nip + j] * (double) XFg[(size_t) b2 * nip + j];
                    sacc = team_sum<TM>(sacc, L.red);
                    if (t0 == 0) Gg[b1 + b2 * nbd] -= sacc;
                }
        }
        TM::sync();
        int gbad = 0;
        if (t0 == 0) {
            double *G = F.G;
            for (int k = 0; k < nbd; ++k) {
                int p = k;
                double best = fabs(G[k + k * nbd]);
                for (int r = k + 1; r < nbd; ++r)
                    if (fabs(G[r + k * nbd]) > best) { best = fabs(G[r + k * nbd]); p = r; }
                F.gpiv[k] = p;
                if (!(best > 0.0) || !isfinite(best)) { gbad = 1; break; }
                if (p != k)
                    for (int c = 0; c < nbd; ++c) {
                        const double t = G[k + c * nbd];
                        G[k + c * nbd] = G[p + c * nbd];
                        G[p + c * nbd] = t;
                    }
                const double inv = 1.0 / G[k + k * nbd];
                for (int r = k + 1; r < nbd; ++r) G[r + k * nbd] *= inv;
                for (int c = k + 1; c < nbd; ++c) {
                    const double u = G[k + c * nbd];
                    for (int r = k + 1; r < nbd; ++r) G[r + c * nbd] -= G[r + k * nbd] * u;
                }
            }
        }
        if (team_any<TM>(gbad != 0, L.red)) return 1;
    }
    return 0;
}

// x = A^-1 b with the factors F (b, x: FP64 [n], reduced order, mass-fraction units).  The
// Thomas sweeps run on the zone's first wavefront; the scalings and the border correction are
// shared by the zone's threads.
template <class T, int NB, class TM>
__device__ __forceinline__ void block_solve(const DevBlock &B, const BlockFactors<T> &F, const BlockLds<T> &L,
                            const double *bp, double *xp) {
    constexpr int NT = TM::NT;
    const NUPPN_GLB double *b = gptr(bp), *rowinv = gptr(F.rowinv), *colsc = gptr(F.colsc);
    const NUPPN_GLB int *vpos = gptr(B.vpos);
    NUPPN_GLB double *x = gptr(xp);
    const NUPPN_GLB T *Eg = gptr(F.E), *XFg = gptr(F.XF);
    const int t0 = TM::tid(), nip = B.nip, nbd = B.nbd;
    NUPPN_LDS T *v = L.vec;
    // an exact power of two brings the largest entry to [1, 2): FP32 then resolves entries down
    // to 2^-126 of the largest one rather than 2^-126 absolute (refinement residuals shrink far
    // below the FP32 range while the abundances they correct stay meaningful)
    double m = 0.0;
    for (int c = t0; c < B.n; c += NT) m = fmax(m, fabs(b[c] * rowinv[vpos[c]]));
    m = team_max<TM>(m, L.red);
    const int ex = (m > 0.0 && isfinite(m)) ? ilogb(m) : 0;
    const double sc = ldexp(1.0, -ex), unsc = ldexp(1.0, ex);
    for (int j = B.ni + t0; j < nip; j += NT) v[j] = T(0);
    for (int c = t0; c < B.n; c += NT) v[vpos[c]] = (T) (b[c] * rowinv[vpos[c]] * sc);
    TM::sync();
    if constexpr (TM::NWAVE == 1) thomas_solve<T, NB>(B, F, v);
    else twisted_solve<T, NB, 1, TM>(B, F, v, 0, 1, TM::wave(), TM::wave() < 2);
    TM::sync();
    if (nbd > 0) {
        NUPPN_LDS double *tb = L.bscr;
        if (nbd == 3) {   // the usual border (n, p, alpha): the three sums in one reduction
            double s3[3] = {0.0, 0.0, 0.0};
            for (int j = t0; j < nip; j += NT) {
                const double vj = (double) v[j];
#pragma unroll
                for (int bb = 0; bb < 3; ++bb) s3[bb] += (double) Eg[(size_t) bb * nip + j] * vj;
            }
            team_sum_k<TM, 3>(s3, L.red);
            if (t0 == 0)
                for (int bb = 0; bb < 3; ++bb) tb[bb] = (double) v[nip + bb] - s3[bb];
        } else {
            for (int bb = 0; bb < nbd; ++bb) {
                double s = 0.0;
                for (int j = t0; j < nip; j += NT) s += (double) Eg[(size_t) bb * nip + j] * (double) v[j];
                s = team_sum<TM>(s, L.red);
                if (t0 == 0) tb[bb] = (double) v[nip + bb] - s;
            }
        }
        TM::sync();
        // G' u = t with the row interchanges, unit-lower and upper triangular solves (thread 0)
        if (t0 == 0) {
            for (int k = 0; k < nbd; ++k) {
                const int p = F.gpiv[k];
                if (p != k) { const double t = tb[k]; tb[k] = tb[p]; tb[p] = t; }
            }
            for (int k = 0; k < nbd; ++k)
                for (int r = k + 1; r < nbd; ++r) tb[r] -= F.G[r + k * nbd] * tb[k];
            for (int k = nbd - 1; k >= 0; --k) {
                tb[k] /= F.G[k + k * nbd];
                for (int r = 0; r < k; ++r) tb[r] -= F.G[r + k * nbd] * tb[k];
            }
            for (int k = 0; k < nbd; ++k) v[nip + k] = (T) tb[k];
        }
        TM::sync();
        for (int j = t0; j < nip; j += NT) {
            T s = v[j];
            for (int bb = 0; bb < nbd; ++bb) s = fma_t(-XFg[(size_t) bb * nip + j], v[nip + bb], s);
            v[j] = s;
        }
        TM::sync();
    }
    batched_for<4>(
        t0, B.n, NT,
        [&](int c) {
            const int p = vpos[c];
            return (double) v[p] * colsc[p];
        },
        [&](int c, double t) { x[c] = t * unsc; });
    TM::sync();
}

// one or two running sums of a segmented sum
struct Sum1 {
    double s;
    __device__ void add(const Sum1 &o) { s = s + o.s; }
    __device__ Sum1 lane(int k) const { return {readlane_d(s, k)}; }
};
struct Sum2 {
    double s, a;
    __device__ void add(const Sum2 &o) {
        s = s + o.s;
        a = a + o.a;
    }
    __device__ Sum2 lane(int k) const { return {readlane_d(s, k), readlane_d(a, k)}; }
};

// Segmented sums of val(P, operand) (P: a term's physical position) over the terms of round r
// of S by one wavefront.  Lane l sums its range of terms in order, sixteen independent loads at
// a time, so that a lane keeps many gathers in flight instead of one short segment after
// another; the lanes' j-th terms are adjacent in memory, so each of those loads is one
// contiguous access.  A segment that ends inside a lane's range and did not start before it is
// finished there; the pieces of the others (the first segment of every lane, and the one still
// open at its end) are added in lane order afterwards.  Every segment is summed in term order,
// its pieces in lane order, so the result does not depend on timing.  emit(s, total) is called
// once per segment, from any lane.  A round starts and ends at segment boundaries, so the
// rounds are independent.
template <class A, class V, class E>
__device__ __forceinline__ void seg_round(const DevSeg &S, int r, V val, E emit) {
    constexpr int BATCH = 16;
    const int l = lane_id();
    const NUPPN_GLB int *eidx = gptr(S.e_idx);
    const int len = gptr(S.lane_len)[64 * r + l];
    const int base = gptr(S.round_base)[r] + l;
    const int s0 = gptr(S.lane_s0)[64 * r + l];
    int s = s0;
    A acc{}, first{};
    bool has_first = false, open = false;
    // the operand words of the next batch are loaded while this batch's operands are gathered
    // (one memory round trip per batch instead of two)
    int rawn[BATCH];
    auto load_raw = [&](int j0) {
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
            rawn[j] = eidx[base + WAVE * min(j0 + j, len - 1)];   // past the range: a repeat
    };
    if (len > 0) load_raw(0);
    for (int j0 = 0; j0 < len; j0 += BATCH) {
        A v[BATCH];
        bool end[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            const int P = base + WAVE * min(j0 + j, len - 1);
            const int raw = rawn[j];
            end[j] = raw < 0;
            v[j] = val(P, raw & 0x7fffffff);
        }
        if (j0 + BATCH < len) load_raw(j0 + BATCH);
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            if (j0 + j < len) {
                acc.add(v[j]);
                open = !end[j];
                if (end[j]) {
                    if (s == s0) {
                        first = acc;
                        has_first = true;
                    } else {
                        emit(s, acc);
                    }
                    acc = A{};
                    ++s;
                }
            }
        }
    }
    // the pieces, lane by lane (uniform across the wave; lane 0 emits)
    int cur = -1;
    A tot{};
    for (int k = 0; k < WAVE; ++k) {
        if (readlane_i((int) has_first, k)) {
            const int sk = readlane_i(s0, k);
            const A p = first.lane(k);
            if (sk == cur) {
                tot.add(p);
            } else {
                if (cur >= 0 && l == 0) emit(cur, tot);
                cur = sk;
                tot = p;
            }
        }
        if (readlane_i((int) open, k)) {
            const int sk = readlane_i(s, k);
            const A p = acc.lane(k);
            if (sk == cur) {
                tot.add(p);
            } else {
                if (cur >= 0 && l == 0) emit(cur, tot);
                cur = sk;
                tot = p;
            }
        }
    }
    if (cur >= 0 && l == 0) emit(cur, tot);
}

// All rounds of S by the zone's threads: one wavefront takes them one after another, several
// wavefronts one round each (the same pieces either way, so the same bits).
template <class A, class TM, class V, class E>
__device__ __forceinline__ void seg_stream(const DevSeg &S, V val, E emit) {
    for (int r = TM::wave(); r < kSegRounds; r += TM::NWAVE) seg_round<A>(S, r, val, emit);
    TM::sync();
}

// A running sum carried in twice the working precision (Ogita, Rump and Oishi, "Accurate sum
// and dot product", SIAM J. Sci. Comput. 26 (2005), algorithm Dot2): hi + lo, where every
// addition into hi is split exactly by TwoSum (Knuth) and its rounding error collects in lo;
// a term arrives as an exact product split p + e by one FMA.  The result is as accurate as a
// sum computed in double-double and rounded once, error <= eps |sum| + O(n eps^2) sum |terms|.
// Beside it, the plain sum of the terms' magnitudes.
struct SumDD {
    double hi, lo, a;
    __device__ void add(const SumDD &o) {
        const double s = hi + o.hi;
        const double bb = s - hi;
        const double err = (hi - (s - bb)) + (o.hi - bb);
        hi = s;
        lo = lo + (o.lo + err);
        a = a + o.a;
    }
    __device__ SumDD lane(int k) const {
        return {readlane_d(hi, k), readlane_d(lo, k), readlane_d(a, k)};
    }
};

// the same without the magnitudes
struct SumDD2 {
    double hi, lo;
    __device__ void add(const SumDD2 &o) {
        const double s = hi + o.hi;
        const double bb = s - hi;
        const double err = (hi - (s - bb)) + (o.hi - bb);
        hi = s;
        lo = lo + (o.lo + err);
    }
    __device__ SumDD2 lane(int k) const { return {readlane_d(hi, k), readlane_d(lo, k)}; }
};

// r = rhs - A (x + xlo) (A's entries in the residual stream's physical order, Aphys; x + xlo
// the solution carried in twice the FP64 precision) with the products and sums carried in
// twice the FP64 precision, rounded once: the residual of the FP64 system to FP64 accuracy even where the
// terms cancel by many orders of magnitude, so iterative refinement converges to the solution
// of the FP64 system itself (its forward error, not only its backward error, reaches the FP64
// level: Demmel, Hida, Kahan, Li, Mukherjee and Riedy, "Error bounds from extra-precise
// iterative refinement", ACM TOMS 32 (2006)).  Returns the componentwise backward error
// max_i |r_i| / (|A| |x| + |rhs| + kAbundanceFloor)_i (Oettli-Prager).  lo, absum: [n] scratch.
template <class TM>
__device__ __forceinline__ double block_residual(const DevBlock &B, const double *Aphysp,
                                                 const double *rhsp, const double *xp,
                                                 const double *xlop, double *rp,
                                                 double *absump, double *lop,
                                                 NUPPN_LDS double *red) {
    const NUPPN_GLB double *Aphys = gptr(Aphysp), *rhs = gptr(rhsp), *x = gptr(xp);
    const NUPPN_GLB double *xlo = gptr(xlop);
    NUPPN_GLB double *r = gptr(rp), *absum = gptr(absump), *lo = gptr(lop);
    // the residual's term at physical position P: A's entry stored there (Aphys) times x of the
    // entry's column (the operand)
    seg_stream<SumDD, TM>(
        B.resid,
        [&](int P, int c) {
            const double av = Aphys[P], xv = x[c];
            const double t = av * xv;
            return SumDD{t, fma(av, xlo[c], fma(av, xv, -t)), fabs(t)};
        },
        [&](int i, const SumDD &t) {
            r[i] = t.hi;
            lo[i] = t.lo;
            absum[i] = t.a;
        });
    double om = 0.0;
    struct Rr {
        double b, hi, lo, a;
    };
    batched_for<4>(
        TM::tid(), B.n, TM::NT, [&](int i) { return Rr{rhs[i], r[i], lo[i], absum[i]}; },
        [&](int i, const Rr &q) {
            // rhs - (hi + lo): the difference rhs - hi split exactly, then the low parts
            const double b = q.b, hi = q.hi;
            const double s = b - hi;
            const double bb = s - b;
            const double err = (b - (s - bb)) + (-hi - bb);
            const double ri = s + (err - q.lo);
            r[i] = ri;
            om = fmax(om, fabs(ri) / (q.a + fabs(b) + kAbundanceFloor));
        });
    om = team_max<TM>(om, red);
    TM::sync();
    return om;
}

}  // namespace nuppn
