// Host-side symbolic phase, done once per network: the full Jacobian pattern (makemap), the
// static reduced pattern of I - dt J, the pull maps used by the device assembly, and the
// column order of the factorisation (COLAMD + column elimination tree post-order).
#pragma once
#include <vector>

namespace nuppn {

struct HostNetwork {
    int nsp = 0, nre = 0, igamma = 0;
    std::vector<int> k;         // [8*nre]
    std::vector<int> rtype;     // [nre]
    std::vector<double> coef;   // [7*nre]
    std::vector<double> aw, zw; // [nsp]
};

struct Symbolic {
    // makemap (solver/source/jac_rhs.F90:461-540): full pattern, row-major CSR, and the
    // per-reaction slot of each of the 8 positions (dns2csr_map(8,nre) -> map[m*nre+l])
    int nnz_full = 0;
    std::vector<int> full_rowptr, full_colind, map;
    // static reduced system (all species touched by a reaction, gamma excluded)
    int n = 0, nnzA = 0;
    std::vector<int> red2sp, sp2red;           // [n], [nsp]
    std::vector<int> A_colptr, A_rowind;       // CSC of the reduced pattern, rows sorted
    std::vector<unsigned char> A_isdiag;
    std::vector<int> jq_ptr, jq_ent;           // per CSC entry: (l<<3 | m) in reaction order
    std::vector<int> dyn_q;                    // [8*nre] CSC entry of slot m of reaction l, -1: gamma
    std::vector<int> dx_ptr, dx_ent;           // per reduced species: (l<<2 | role)
    std::vector<int> perm_c;                   // column order, SuperLU convention
};

// builds everything but the column order; returns 0 or a negative error
int build_symbolic(const HostNetwork &net, Symbolic &s);

// Segmented sums as the device streams them.  In logical order the terms of all segments lie
// back to back, in segment order and, inside a segment, in summation order; the last term of
// every segment carries bit 31 in its operand index, and an empty segment gets one zero term.
// The logical terms are cut into kSegRounds rounds at segment boundaries (near-even shares), and
// every round into 64 lane ranges of near-even length: lane l of round r sums lane_len[64 r + l]
// consecutive logical terms, starting in segment lane_s0[64 r + l].  A round is the work of one
// wavefront pass; a zone's single wavefront takes the rounds one after another, the wavefronts
// of a multi-wavefront zone share them.  All sum every segment in the same pieces and order, so
// they give the same bits.
// Storage is interleaved by lane: the j-th term of lane l of round r sits at physical position
// round_base[r] + 64 j + l, so the 64 lanes' j-th terms are adjacent and every load of the
// stream's operand indices and coefficients is one contiguous access (padding positions hold
// operand 0, coefficient 0 and are never summed).  phys[t] is the physical position of logical
// term t (host side only).
constexpr int kSegRounds = 4;
struct SegStream {
    int nseg = 0;
    std::vector<int> e_idx;         // operand index, | (1 << 31) on a segment's last term (J: the
                                    // count packed in bits 20..23, BlockSym::jac_rhi)
    std::vector<double> e_coef;     // coefficient (emptied for J)
    std::vector<int> round_base;    // [kSegRounds + 1]
    std::vector<int> lane_len;      // [64 kSegRounds]
    std::vector<int> lane_s0;       // [64 kSegRounds]
    std::vector<int> phys;          // [logical terms]
};

// Static block structure of the reduced matrix for the pivot-free batched factorisation.
//
// Species that couple to far more species than the median (the light projectiles n, p, alpha of
// a capture network) form the border and are ordered last; the other (interior) species keep the
// network order.  In that order the interior part of I - dt J is banded; with blocks of NB >= the
// half-bandwidth it is block tridiagonal:
//     A = [ D_0  C_0              F_0 ]
//         [ B_0  D_1  C_1         F_1 ]
//         [      B_1  D_2  ...    ... ]
//         [ E_0  E_1  ...         G   ]
// Every CSC entry of the reduced pattern is assigned to one dense block (or to E, F, G) with its
// position inside it; the device builds the dense blocks from these lists.
struct BlockSym {
    int NB = 0, nbd = 0, ni = 0, nblk = 0, nip = 0, bandwidth = 0;
    std::vector<int> pos;       // [n] reduced index -> ordered position (interior, then border)
    std::vector<int> ord;       // [n] ordered position -> reduced index
    std::vector<int> vpos;      // [n] reduced index -> slot in a padded vector (interior
                                //     positions 0..ni-1, border nip..nip+nbd-1)
    std::vector<double> mass;   // [n] mass number of each reduced species
    std::vector<int> diagq;     // [n] CSC index of the diagonal entry of each reduced column
    // entry lists: lists 0..nblk-1 = D_i, nblk..2nblk-2 = C_i (block row i, block col i+1),
    // 2nblk-1..3nblk-3 = B_i (block row i+1, block col i), then E, F, G (3 lists)
    std::vector<int> ent_ptr;   // [3*nblk - 2 + 3 + 1]
    std::vector<int> ent_q;     // CSC entry index
    std::vector<int> ent_loc;   // D/C/B: r_local + c_local*NB; E: b*nip + j; F: b*nip + j; G: r + c*nbd
    std::vector<int> ent_col;   // reduced column index
    std::vector<double> ent_ys; // mass(col) / mass(row): X units -> molar (Y) units
    // CSR view of the CSC pattern for the residual: per reduced row, (CSC entry, column)
    std::vector<int> csr_ptr, csr_q, csr_col;
    // the batched path keeps A in CSR order: ent_k / diagk are ent_q / diagq as CSR positions
    std::vector<int> ent_k, diagk;
    // the dynamic network's entry map (Symbolic::dyn_q) in CSR order, and each CSR entry's row
    std::vector<int> dyn_k, csr_row;
    // the band form (NB = 0: an interior too wide for blocks, band_lu.hpp): window W =
    // bandwidth + 1; the entries entering the factorisation window, list 0 the border-border
    // block, list r + 1 those of interior index r (the larger of the entry's interior positions,
    // or its interior position when the other index is on the border)
    int band_w = 0;
    std::vector<int> bent_ptr;    // [ni + 2]
    std::vector<int> bent_q;      // CSC entry index
    std::vector<int> bent_k;      // the same as a CSR position
    std::vector<int> bent_dst;    // window slot, row * (W + nbd) + col
    std::vector<int> bent_col;    // reduced column
    std::vector<double> bent_ys;  // mass(col) / mass(row)
    // D/C/B entries: destination in the dense block storage [area][block][NB*NB] (area 0 D_i,
    // 1 C_i, 2 B_i; column-major inside a block)
    std::vector<int> ent_dst;
    std::vector<unsigned char> csr_isdiag;
    // chunked sums: J entries (CSR order) from the per-reaction stems [st1 | st2] (2 nre),
    // dX/dt of reduced species from the fluxes vx (nre, coefficients include the mass number),
    // the residual rows from the CSR entries (operand = column; the coefficient of the term at
    // physical position P is A's entry stored there, see resid_pos)
    SegStream jac, dxdt, resid;
    // J: per CSR entry the ratio of mass numbers an[row] / an[col] as a double-double hi + lo,
    // the factor every contribution of the entry shares
    std::vector<double> jac_rhi, jac_rlo;
    std::vector<int> resid_pos;   // [nnzA] physical position of each CSR entry in the residual stream
};

// Builds the block structure for block size NB (32 or 64), or the band form for NB = 0;
// returns 0, or -1 when the interior half-bandwidth exceeds NB (the block-tridiagonal form does
// not hold) or, for the band form, the window and border exceed kBandWindowMax.
int build_block_symbolic(const HostNetwork &net, const Symbolic &s, int NB, BlockSym &b);

// smallest supported block size (32, 64) for which the network's interior is block tridiagonal;
// 0 if none
// the band form's limit on window + border (its factorisation's 16 x 16 thread grid holds
// 8 x 8 slots per thread) and choose_block_size's answer for that form
constexpr int kBandWindowMax = 128;
constexpr int kBandForm = -1;
// the factorisation prefetches the entering lists by groups of kBandListGroup consecutive
// interior indices into LDS banks of kBandListCap entries (band_lu.hpp)
constexpr int kBandListGroup = 16;
constexpr int kBandListCap = 512;

// the block size of the batched path: 32 or 64 where the interior is block tridiagonal with
// that size, else kBandForm where the band form fits, else 0 (no batched path)
int choose_block_size(const HostNetwork &net, const Symbolic &s);

// COLAMD (v2.3 semantics, SuperLU_5.0/SRC/colamd.c) on an m x n CSC pattern, default knobs;
// p receives the column ordering: p[k] = column eliminated k-th.  Returns 0 or an error.
int colamd_order(int m, int n, const int *colptr, const int *rowind, std::vector<int> &p);

// get_perm_c(COLAMD) + sp_preorder post-order (SuperLU_5.0/SRC/get_perm_c.c:405,
// sp_preorder.c:140-150): perm_c[i] = final position of column i; etree in that order.
int superlu_column_order(int n, const int *colptr, const int *rowind, std::vector<int> &perm_c,
                         std::vector<int> *etree_out = nullptr);

// column elimination tree of A*Pc (sp_coletree.c) and its post-order (TreePostorder)
void column_etree(int nr, int nc, const int *colbeg, const int *colend, const int *rowind,
                  std::vector<int> &parent);
std::vector<int> tree_postorder(int n, const std::vector<int> &parent);

}  // namespace nuppn
