// Host-side symbolic phase (see symbolic.hpp).
#include "symbolic.hpp"

#include <cmath>

#include <algorithm>
#include <cstdint>
#include <tuple>

namespace nuppn {

namespace {

inline int find_in_row(const std::vector<int> &rowptr, const std::vector<int> &colind, int r,
                       int c) {
    auto b = colind.begin() + rowptr[r], e = colind.begin() + rowptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? int(it - colind.begin()) : -1;
}

}  // namespace

int build_symbolic(const HostNetwork &net, Symbolic &s) {
    const int nsp = net.nsp, nre = net.nre, g = net.igamma;
    if (nsp <= 1 || nre <= 0 || g < 0 || g >= nsp) return -1;
    for (int m : {0, 2, 4, 6})
        for (int l = 0; l < nre; ++l)
            if (net.k[m * nre + l] < 0 || net.k[m * nre + l] >= nsp) return -1;

    // --- makemap: 8 positions per reaction + diagonal of each considered isotope, numbered
    //     row-major (jac_rhs.F90:481-511), slot lookup by binary search (getelm)
    std::vector<int64_t> keys;
    keys.reserve(8 * (size_t) nre + nsp);
    auto pos_rc = [&](int l, int m, int &r, int &c) {
        const int in1 = net.k[l], in2 = net.k[2 * nre + l];
        const int out1 = net.k[4 * nre + l], out2 = net.k[6 * nre + l];
        // dns2csr_map order: in1in1, in2in1, out1in1, out2in1, in1in2, in2in2, out1in2, out2in2
        const int rr[8] = {in1, in2, out1, out2, in1, in2, out1, out2};
        const int cc[8] = {in1, in1, in1, in1, in2, in2, in2, in2};
        r = rr[m];
        c = cc[m];
    };
    for (int l = 0; l < nre; ++l)
        for (int m = 0; m < 8; ++m) {
            int r, c;
            pos_rc(l, m, r, c);
            keys.push_back((int64_t) r * nsp + c);
        }
    for (int i = 0; i < nsp; ++i)
        if (i != g) keys.push_back((int64_t) i * nsp + i);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    s.nnz_full = (int) keys.size();
    s.full_rowptr.assign(nsp + 1, 0);
    s.full_colind.resize(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {
        s.full_rowptr[keys[i] / nsp + 1]++;
        s.full_colind[i] = (int) (keys[i] % nsp);
    }
    for (int i = 0; i < nsp; ++i) s.full_rowptr[i + 1] += s.full_rowptr[i];
    s.map.resize(8 * (size_t) nre);
    for (int l = 0; l < nre; ++l)
        for (int m = 0; m < 8; ++m) {
            int r, c;
            pos_rc(l, m, r, c);
            s.map[m * nre + l] = find_in_row(s.full_rowptr, s.full_colind, r, c);
        }

    // --- static mask: every species of every reaction (cyminformal < 0 marks all of them,
    //     jac_rhs.F90:350-353), gamma cleared (line 442)
    std::vector<unsigned char> mask(nsp, 0);
    for (int l = 0; l < nre; ++l)
        for (int m : {0, 2, 4, 6}) mask[net.k[m * nre + l]] = 1;
    mask[g] = 0;
    s.sp2red.assign(nsp, -1);
    s.red2sp.clear();
    for (int i = 0; i < nsp; ++i)
        if (mask[i]) {
            s.sp2red[i] = (int) s.red2sp.size();
            s.red2sp.push_back(i);
        }
    s.n = (int) s.red2sp.size();

    // --- reduced entries with their contributions.  Statement order m of calculate_jacobian
    //     (jac_rhs.F90:365-392): ice1 (in1,in1), ice2 (in2,in1), ice3 (in1,in2), ice4
    //     (out1,in1), ice5 (out2,in1), ice6 (in2,in2), ice7 (out1,in2), ice8 (out2,in2).
    struct Contrib { int col, row, l, m; };
    std::vector<Contrib> con;
    con.reserve(8 * (size_t) nre);
    for (int l = 0; l < nre; ++l) {
        const int in1 = net.k[l], in2 = net.k[2 * nre + l];
        const int out1 = net.k[4 * nre + l], out2 = net.k[6 * nre + l];
        const int rr[8] = {in1, in2, in1, out1, out2, in2, out1, out2};
        const int cc[8] = {in1, in1, in2, in1, in1, in2, in2, in2};
        const bool ok[8] = {in1 != g,
                            in1 != g && in2 != g,
                            in1 != g && in2 != g,
                            in1 != g && out1 != g,
                            in1 != g && out2 != g,
                            in2 != g,
                            in2 != g && out1 != g,
                            in2 != g && out2 != g};
        for (int m = 0; m < 8; ++m)
            if (ok[m]) con.push_back({s.sp2red[cc[m]], s.sp2red[rr[m]], l, m});
    }
    for (int r = 0; r < s.n; ++r) con.push_back({r, r, -1, -1});   // diagonal of masked species
    std::sort(con.begin(), con.end(), [](const Contrib &a, const Contrib &b) {
        return std::tie(a.col, a.row, a.l, a.m) < std::tie(b.col, b.row, b.l, b.m);
    });
    s.A_colptr.assign(s.n + 1, 0);
    s.A_rowind.clear();
    s.A_isdiag.clear();
    s.jq_ptr.assign(1, 0);
    s.jq_ent.clear();
    for (size_t i = 0; i < con.size();) {
        size_t j = i;
        const int col = con[i].col, row = con[i].row;
        while (j < con.size() && con[j].col == col && con[j].row == row) {
            if (con[j].l >= 0) s.jq_ent.push_back((con[j].l << 3) | con[j].m);
            ++j;
        }
        s.A_rowind.push_back(row);
        s.A_isdiag.push_back(row == col ? 1 : 0);
        s.A_colptr[col + 1]++;
        s.jq_ptr.push_back((int) s.jq_ent.size());
        i = j;
    }
    for (int c = 0; c < s.n; ++c) s.A_colptr[c + 1] += s.A_colptr[c];
    s.nnzA = (int) s.A_rowind.size();
    // the dynamic network's map: the entry each (reaction, slot) contributes to
    s.dyn_q.assign(8 * (size_t) nre, -1);
    for (int q = 0; q < s.nnzA; ++q)
        for (int e = s.jq_ptr[q]; e < s.jq_ptr[q + 1]; ++e)
            s.dyn_q[(s.jq_ent[e] & 7) * (size_t) nre + (s.jq_ent[e] >> 3)] = q;

    // --- dX/dt pull map: contributions of reaction l to species k1, k3, k5, k7 in that order
    //     (jac_rhs.F90:252-255), reactions ascending
    std::vector<std::vector<int>> dx(s.n);
    for (int l = 0; l < nre; ++l) {
        const int sp[4] = {net.k[l], net.k[2 * nre + l], net.k[4 * nre + l], net.k[6 * nre + l]};
        for (int role = 0; role < 4; ++role)
            if (sp[role] != g && mask[sp[role]]) dx[s.sp2red[sp[role]]].push_back((l << 2) | role);
    }
    s.dx_ptr.assign(1, 0);
    s.dx_ent.clear();
    for (int r = 0; r < s.n; ++r) {
        s.dx_ent.insert(s.dx_ent.end(), dx[r].begin(), dx[r].end());
        s.dx_ptr.push_back((int) s.dx_ent.size());
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// Column elimination tree (Liu's algorithm on the row cliques of A'A replaced by stars at each
// row's first column; disjoint sets with path halving) - SuperLU_5.0/SRC/sp_coletree.c, of which
// column_etree and dset_find below are a modified version (rewritten in C++).  That file's notice:
//   Copyright (c) 1994 by Xerox Corporation.  All rights reserved.
//   THIS MATERIAL IS PROVIDED AS IS, WITH ABSOLUTELY NO WARRANTY EXPRESSED OR IMPLIED.  ANY USE
//   IS AT YOUR OWN RISK.
//   Permission is hereby granted to use or copy this program for any purpose, provided the
//   above notices are retained on all copies.  Permission to modify the code and to distribute
//   modified code is granted, provided the above notices are retained, and a notice that the
//   code was modified is included with the above copyright notice.
//   (SuperLU routine, version 3.1: Univ. of California Berkeley, Xerox Palo Alto Research
//   Center, and Lawrence Berkeley National Lab.)  This code was modified.

namespace {
inline int dset_find(std::vector<int> &pp, int i) {
    int p = pp[i], gp = pp[p];
    while (gp != p) {
        pp[i] = gp;
        i = gp;
        p = pp[i];
        gp = pp[p];
    }
    return p;
}
}  // namespace

void column_etree(int nr, int nc, const int *colbeg, const int *colend, const int *rowind,
                  std::vector<int> &parent) {
    std::vector<int> root(nc), pp(nc), firstcol(nr, nc);
    parent.assign(nc, nc);
    for (int col = 0; col < nc; ++col)
        for (int p = colbeg[col]; p < colend[col]; ++p)
            firstcol[rowind[p]] = std::min(firstcol[rowind[p]], col);
    for (int col = 0; col < nc; ++col) {
        pp[col] = col;
        int cset = col;
        root[cset] = col;
        parent[col] = nc;
        for (int p = colbeg[col]; p < colend[col]; ++p) {
            const int row = firstcol[rowind[p]];
            if (row >= col) continue;
            const int rset = dset_find(pp, row);
            const int rroot = root[rset];
            if (rroot != col) {
                parent[rroot] = col;
                pp[cset] = rset;
                cset = rset;
                root[cset] = col;
            }
        }
    }
}

// post-order numbering of a forest given by parent[] (root parent = n), children visited in
// ascending order; returns post[0..n] with post[n] = n
std::vector<int> tree_postorder(int n, const std::vector<int> &parent) {
    std::vector<int> first_kid(n + 1, -1), next_kid(n + 1, -1), post(n + 1, 0);
    for (int v = n - 1; v >= 0; --v) {
        const int dad = parent[v];
        next_kid[v] = first_kid[dad];
        first_kid[dad] = v;
    }
    // iterative DFS from the dummy root n
    std::vector<int> stack;
    stack.reserve(n + 1);
    int postnum = 0;
    std::vector<int> it(n + 1);
    for (int v = 0; v <= n; ++v) it[v] = first_kid[v];
    stack.push_back(n);
    while (!stack.empty()) {
        const int v = stack.back();
        if (it[v] != -1) {
            const int w = it[v];
            it[v] = next_kid[w];
            stack.push_back(w);
        } else {
            post[v] = postnum++;
            stack.pop_back();
        }
    }
    return post;
}

int superlu_column_order(int n, const int *colptr, const int *rowind, std::vector<int> &perm_c,
                         std::vector<int> *etree_out) {
    std::vector<int> p;
    const int rc = colamd_order(n, n, colptr, rowind, p);
    if (rc != 0) return rc;
    // get_colamd: perm_c[p[i]] = i
    perm_c.assign(n, 0);
    for (int i = 0; i < n; ++i) perm_c[p[i]] = i;
    // sp_preorder: AC = A Pc, column etree of AC, post-order, perm_c <- post[perm_c]
    std::vector<int> colbeg(n), colend(n);
    for (int i = 0; i < n; ++i) {
        colbeg[perm_c[i]] = colptr[i];
        colend[perm_c[i]] = colptr[i + 1];
    }
    std::vector<int> etree;
    column_etree(n, n, colbeg.data(), colend.data(), rowind, etree);
    const std::vector<int> post = tree_postorder(n, etree);
    if (etree_out) {
        etree_out->assign(n, 0);
        for (int i = 0; i < n; ++i) (*etree_out)[post[i]] = post[etree[i]];
    }
    for (int i = 0; i < n; ++i) perm_c[i] = post[perm_c[i]];
    return 0;
}

}  // namespace nuppn

// ------------------------------------------------------------------------------------------
// Block structure for the pivot-free batched factorisation (see symbolic.hpp).

namespace nuppn {

namespace {

// border: reduced species whose row + column entry count exceeds 8x the median count
std::vector<char> border_species(const Symbolic &s) {
    const int n = s.n;
    std::vector<int> deg(n, 0);
    for (int c = 0; c < n; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) {
            deg[c]++;
            deg[s.A_rowind[p]]++;
        }
    std::vector<int> sorted(deg);
    std::nth_element(sorted.begin(), sorted.begin() + n / 2, sorted.end());
    const int med = std::max(1, sorted[n / 2]);
    std::vector<char> isb(n, 0);
    for (int i = 0; i < n; ++i) isb[i] = deg[i] > 8 * med;
    return isb;
}

int interior_bandwidth(const Symbolic &s, const std::vector<int> &pos, int ni) {
    int bw = 0;
    for (int c = 0; c < s.n; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) {
            const int pr = pos[s.A_rowind[p]], pc = pos[c];
            if (pr < ni && pc < ni) bw = std::max(bw, std::abs(pr - pc));
        }
    return bw;
}

}  // namespace

static int build_batched_sums(const HostNetwork &net, const Symbolic &s, BlockSym &b);

int build_block_symbolic(const HostNetwork &net, const Symbolic &s, int NB, BlockSym &b) {
    const int n = s.n;
    const std::vector<char> isb = border_species(s);
    b.NB = NB;
    b.ord.clear();
    for (int i = 0; i < n; ++i)
        if (!isb[i]) b.ord.push_back(i);
    b.ni = (int) b.ord.size();
    for (int i = 0; i < n; ++i)
        if (isb[i]) b.ord.push_back(i);
    b.nbd = n - b.ni;
    b.pos.assign(n, 0);
    for (int p = 0; p < n; ++p) b.pos[b.ord[p]] = p;
    b.bandwidth = interior_bandwidth(s, b.pos, b.ni);
    if (b.ni == 0) return -1;
    if (NB == 0) {
        // the band form: window and border within the factorisation's thread grid
        b.band_w = b.bandwidth + 1;
        if (b.band_w + b.nbd > kBandWindowMax) return -1;
        b.nblk = 0;
        b.nip = b.ni;
    } else {
        if (b.bandwidth > NB) return -1;
        b.nblk = (b.ni + NB - 1) / NB;
        b.nip = b.nblk * NB;
    }
    b.mass.resize(n);
    for (int i = 0; i < n; ++i) b.mass[i] = net.aw[s.red2sp[i]];
    b.diagq.assign(n, -1);
    for (int c = 0; c < n; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p)
            if (s.A_rowind[p] == c) b.diagq[c] = p;
    for (int c = 0; c < n; ++c)
        if (b.diagq[c] < 0 || b.mass[c] <= 0.0) return -2;

    if (NB == 0) {
        // entering lists of the band window (band_lu.hpp): slot of interior position i is
        // i mod W, of border b W + b
        const int W = b.band_w, WT = W + b.nbd, ni = b.ni;
        std::vector<std::vector<int>> lq(ni + 1);
        for (int c = 0; c < n; ++c)
            for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) {
                const int pr = b.pos[s.A_rowind[p]], pc = b.pos[c];
                int li;
                if (pr < ni && pc < ni) li = std::max(pr, pc) + 1;
                else if (pr < ni) li = pr + 1;
                else if (pc < ni) li = pc + 1;
                else li = 0;
                lq[li].push_back(p);
            }
        auto slot = [&](int ps) { return ps < ni ? ps % W : W + (ps - ni); };
        b.bent_ptr.assign(1, 0);
        b.bent_q.clear();
        b.bent_dst.clear();
        b.bent_col.clear();
        b.bent_ys.clear();
        std::vector<int> colof(s.nnzA);
        for (int c = 0; c < n; ++c)
            for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) colof[p] = c;
        for (int li = 0; li <= ni; ++li) {
            for (int p : lq[li]) {
                const int r = s.A_rowind[p], c = colof[p];
                b.bent_q.push_back(p);
                b.bent_dst.push_back(slot(b.pos[r]) * WT + slot(b.pos[c]));
                b.bent_col.push_back(c);
                b.bent_ys.push_back(b.mass[c] / b.mass[r]);
            }
            b.bent_ptr.push_back((int) b.bent_q.size());
        }
        for (int l0 = 1; l0 <= ni; l0 += kBandListGroup)
            if (b.bent_ptr[std::min(l0 + kBandListGroup, ni + 1)] - b.bent_ptr[l0] > kBandListCap)
                return -1;
        b.ent_ptr.assign(1, 0);
        b.ent_q.clear();
        b.ent_loc.clear();
        b.ent_col.clear();
        b.ent_ys.clear();
    }
    const int nblk = b.nblk, nlists = NB == 0 ? 0 : 3 * nblk - 2 + 3;
    const int LE = 3 * nblk - 2, LF = LE + 1, LG = LE + 2;
    std::vector<std::vector<int>> q(nlists), loc(nlists);
    for (int c = 0; c < n && NB > 0; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) {
            const int r = s.A_rowind[p];
            const int pr = b.pos[r], pc = b.pos[c];
            int list, l;
            if (pr < b.ni && pc < b.ni) {
                const int bi = pr / NB, bj = pc / NB, lr = pr % NB, lc = pc % NB;
                if (bi == bj) list = bi;
                else if (bj == bi + 1) list = nblk + bi;
                else if (bi == bj + 1) list = 2 * nblk - 1 + bj;
                else return -1;
                l = lr + lc * NB;
            } else if (pr >= b.ni && pc < b.ni) {
                list = LE;
                l = (pr - b.ni) * b.nip + pc;
            } else if (pr < b.ni) {
                list = LF;
                l = (pc - b.ni) * b.nip + pr;
            } else {
                list = LG;
                l = (pr - b.ni) + (pc - b.ni) * b.nbd;
            }
            q[list].push_back(p);
            loc[list].push_back(l);
        }
    if (NB > 0) {
    b.ent_ptr.assign(1, 0);
    b.ent_q.clear();
    b.ent_loc.clear();
    b.ent_col.clear();
    b.ent_ys.clear();
    std::vector<int> colof(s.nnzA);
    for (int c = 0; c < n; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) colof[p] = c;
    for (int L = 0; L < nlists; ++L) {
        for (size_t e = 0; e < q[L].size(); ++e) {
            const int p = q[L][e], c = colof[p], r = s.A_rowind[p];
            b.ent_q.push_back(p);
            b.ent_loc.push_back(loc[L][e]);
            b.ent_col.push_back(c);
            b.ent_ys.push_back(b.mass[c] / b.mass[r]);
        }
        b.ent_ptr.push_back((int) b.ent_q.size());
    }
    }
    b.vpos.assign(n, 0);
    for (int i = 0; i < n; ++i) b.vpos[i] = b.pos[i] < b.ni ? b.pos[i] : b.nip + b.pos[i] - b.ni;
    // CSR view for the residual r = rhs - A x
    b.csr_ptr.assign(n + 1, 0);
    for (int p = 0; p < s.nnzA; ++p) b.csr_ptr[s.A_rowind[p] + 1]++;
    for (int i = 0; i < n; ++i) b.csr_ptr[i + 1] += b.csr_ptr[i];
    b.csr_q.assign(s.nnzA, 0);
    b.csr_col.assign(s.nnzA, 0);
    std::vector<int> fill(b.csr_ptr.begin(), b.csr_ptr.end() - 1);
    for (int c = 0; c < n; ++c)
        for (int p = s.A_colptr[c]; p < s.A_colptr[c + 1]; ++p) {
            const int k = fill[s.A_rowind[p]]++;
            b.csr_q[k] = p;
            b.csr_col[k] = c;
        }
    return build_batched_sums(net, s, b);
}

namespace {

// segments given as per-segment element lists -> the streamed layout (SegStream)
void build_stream(const std::vector<std::vector<std::pair<int, double>>> &segs, SegStream &out) {
    out.nseg = (int) segs.size();
    std::vector<int> idx;       // logical order
    std::vector<double> cf;
    std::vector<int> seg_of;    // segment of every term
    for (int sg = 0; sg < out.nseg; ++sg) {
        const auto &seg = segs[sg];
        if (seg.empty()) {
            idx.push_back(0);
            cf.push_back(0.0);
            seg_of.push_back(sg);
        }
        for (size_t t = 0; t < seg.size(); ++t) {
            idx.push_back(seg[t].first);
            cf.push_back(seg[t].second);
            seg_of.push_back(sg);
        }
        idx.back() |= (int) (1u << 31);
    }
    const long long nterm = (long long) idx.size();
    // round boundaries: the first segment start at or after an even share of the terms
    std::vector<long long> rb(kSegRounds + 1, nterm);
    rb[0] = 0;
    for (int r = 1; r < kSegRounds; ++r) {
        long long t = std::max(rb[r - 1], (r * nterm) / kSegRounds);
        while (t < nterm && t > 0 && seg_of[t] == seg_of[t - 1]) ++t;
        rb[r] = t;
    }
    const int NL = 64 * kSegRounds;
    out.lane_len.assign(NL, 0);
    out.lane_s0.assign(NL, out.nseg);
    out.round_base.assign(kSegRounds + 1, 0);
    out.phys.assign(nterm, 0);
    std::vector<long long> t0(NL + 1);
    for (int r = 0; r < kSegRounds; ++r) {
        const long long len = rb[r + 1] - rb[r];
        int maxlen = 0;
        for (int l = 0; l < 64; ++l) {
            const long long a = rb[r] + (l * len) / 64, e = rb[r] + ((l + 1) * len) / 64;
            t0[64 * r + l] = a;
            out.lane_len[64 * r + l] = (int) (e - a);
            out.lane_s0[64 * r + l] = a < nterm ? seg_of[a] : out.nseg;
            maxlen = std::max(maxlen, (int) (e - a));
        }
        out.round_base[r + 1] = out.round_base[r] + 64 * maxlen;
    }
    const int nphys = out.round_base[kSegRounds];
    out.e_idx.assign(nphys, 0);
    out.e_coef.assign(nphys, 0.0);
    for (int r = 0; r < kSegRounds; ++r)
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < out.lane_len[64 * r + l]; ++j) {
                const long long t = t0[64 * r + l] + j;
                const int P = out.round_base[r] + 64 * j + l;
                out.phys[t] = P;
                out.e_idx[P] = idx[t];
                out.e_coef[P] = cf[t];
            }
}

}  // namespace

// chunked J / dX/dt / residual lists of the batched path (after build_block_symbolic's CSR view)
static int build_batched_sums(const HostNetwork &net, const Symbolic &s, BlockSym &b) {
    const int n = s.n, nre = net.nre, nnz = s.nnzA;
    std::vector<int> q2k(nnz);
    for (int k = 0; k < nnz; ++k) q2k[b.csr_q[k]] = k;
    b.ent_k.resize(b.ent_q.size());
    for (size_t e = 0; e < b.ent_q.size(); ++e) b.ent_k[e] = q2k[b.ent_q[e]];
    b.bent_k.resize(b.bent_q.size());
    for (size_t e = 0; e < b.bent_q.size(); ++e) b.bent_k[e] = q2k[b.bent_q[e]];
    const int nblk = b.nblk, BS = b.NB * b.NB;
    b.ent_dst.assign(nblk > 0 ? b.ent_ptr[3 * nblk - 2] : 0, 0);
    for (int L = 0; L < 3 * nblk - 2; ++L) {
        const int area = L < nblk ? 0 : (L < 2 * nblk - 1 ? 1 : 2);
        const int blk = L < nblk ? L : (L < 2 * nblk - 1 ? L - nblk : L - (2 * nblk - 1));
        for (int e = b.ent_ptr[L]; e < b.ent_ptr[L + 1]; ++e)
            b.ent_dst[e] = (area * nblk + blk) * BS + b.ent_loc[e];
    }
    b.diagk.resize(n);
    for (int c = 0; c < n; ++c) b.diagk[c] = q2k[b.diagq[c]];
    b.csr_isdiag.resize(nnz);
    for (int k = 0; k < nnz; ++k) b.csr_isdiag[k] = s.A_isdiag[b.csr_q[k]];
    b.dyn_k.resize(s.dyn_q.size());
    for (size_t e = 0; e < s.dyn_q.size(); ++e) b.dyn_k[e] = s.dyn_q[e] < 0 ? -1 : q2k[s.dyn_q[e]];
    b.csr_row.resize(nnz);
    for (int r = 0; r < n; ++r)
        for (int k = b.csr_ptr[r]; k < b.csr_ptr[r + 1]; ++k) b.csr_row[k] = r;
    const double *an = net.aw.data();
    // J contributions (statement order of calculate_jacobian, jac_rhs.F90:365-392):
    // value = (+-count) * an[row] / an[col] * stem, stem = st1[l] (m in 0,1,3,4) or st2[l].
    // Every contribution of one entry shares the ratio an[row] / an[col], so the stream carries
    // only the count, packed into the term's operand word (bits 20..23: count + 8; bits 0..19:
    // the stem index), and the ratio is applied once per entry: a double-double hi + lo, with hi
    // the rounded quotient and lo = (an[row] - hi an[col]) / an[col] (the division's remainder is
    // exact with an FMA).  The device sums count * stem in double-double (each product split
    // exactly) and multiplies the sum by the ratio, rounding once.
    b.jac_rhi.assign(nnz, 0.0);
    b.jac_rlo.assign(nnz, 0.0);
    std::vector<std::vector<std::pair<int, double>>> segs(nnz);
    for (int k = 0; k < nnz; ++k) {
        const int q = b.csr_q[k];
        const int srow = s.red2sp[b.csr_row[k]], scol = s.red2sp[b.csr_col[k]];
        const double ar = an[srow], ac = an[scol];
        b.jac_rhi[k] = ar / ac;
        b.jac_rlo[k] = std::fma(-b.jac_rhi[k], ac, ar) / ac;
        for (int e = s.jq_ptr[q]; e < s.jq_ptr[q + 1]; ++e) {
            const int l = s.jq_ent[e] >> 3, m = s.jq_ent[e] & 7;
            const int in1 = net.k[l], n1 = net.k[nre + l], in2 = net.k[2 * nre + l];
            const int n2 = net.k[3 * nre + l], out1 = net.k[4 * nre + l], no1 = net.k[5 * nre + l];
            const int out2 = net.k[6 * nre + l], no2 = net.k[7 * nre + l];
            const int rows[8] = {in1, in2, in1, out1, out2, in2, out1, out2};
            const int cols[8] = {in1, in1, in2, in1, in1, in2, in2, in2};
            const int cnt[8] = {-n1, -n2, -n1, no1, no2, -n2, no1, no2};
            const bool stem2[8] = {false, false, true, false, false, true, true, true};
            const int op = stem2[m] ? nre + l : l;
            if (rows[m] != srow || cols[m] != scol || cnt[m] < -8 || cnt[m] > 7 || op >= (1 << 20))
                return -1;
            segs[k].push_back({op, (double) cnt[m]});
        }
    }
    build_stream(segs, b.jac);
    // pack the counts into the operand words (padding and empty-segment terms: count 0)
    for (size_t P = 0; P < b.jac.e_idx.size(); ++P) {
        const int raw = b.jac.e_idx[P];
        const int cnt = (int) b.jac.e_coef[P];
        b.jac.e_idx[P] = (raw & (int) (1u << 31)) | (raw & 0xFFFFF) | ((cnt + 8) << 20);
    }
    b.jac.e_coef.clear();
    // dX/dt (calculate_dxdt, jac_rhs.F90:250-272): roles k1 -k2, k3 -1, k5 +k6, k7 +k8, times A
    std::vector<std::vector<std::pair<int, double>>> dsegs(n);
    for (int r = 0; r < n; ++r) {
        const double a = an[s.red2sp[r]];
        for (int e = s.dx_ptr[r]; e < s.dx_ptr[r + 1]; ++e) {
            const int l = s.dx_ent[e] >> 2, role = s.dx_ent[e] & 3;
            double c;
            if (role == 0) c = -net.k[nre + l];
            else if (role == 1) c = -1.0;
            else if (role == 2) c = net.k[5 * nre + l];
            else c = net.k[7 * nre + l];
            dsegs[r].push_back({l, c * a});
        }
    }
    build_stream(dsegs, b.dxdt);
    // residual rows: the CSR entries of each row, operand = column
    std::vector<std::vector<std::pair<int, double>>> rsegs(n);
    for (int i = 0; i < n; ++i)
        for (int k = b.csr_ptr[i]; k < b.csr_ptr[i + 1]; ++k) rsegs[i].push_back({b.csr_col[k], 1.0});
    build_stream(rsegs, b.resid);
    // every row holds its diagonal, so no dummy terms: logical term k is CSR entry k, and A's
    // entry k is also stored at the term's physical position for the residual
    if ((int) b.resid.phys.size() != nnz) return -1;
    b.resid_pos = b.resid.phys;
    return 0;
}

int choose_block_size(const HostNetwork &net, const Symbolic &s) {
    for (int NB : {32, 64, 0}) {
        BlockSym b;
        if (build_block_symbolic(net, s, NB, b) == 0) return NB == 0 ? kBandForm : NB;
    }
    return 0;
}

}  // namespace nuppn
