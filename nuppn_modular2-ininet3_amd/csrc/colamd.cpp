// This file is a modified version of COLAMD, rewritten in C++ with its own data layout from
// the COLAMD 2.3 sources vendored by the reference (SuperLU_5.0/SRC/colamd.c).  COLAMD's notice:
//
//   COLAMD version 2.3.  The authors of the code itself are Stefan I. Larimore and Timothy A.
//   Davis (davis@cise.ufl.edu), University of Florida.  The algorithm was developed in
//   collaboration with John Gilbert, Xerox PARC, and Esmond Ng, Oak Ridge National Laboratory.
//
//   Copyright and License:
//   Copyright (c) 1998-2003 by the University of Florida.  All Rights Reserved.
//   THIS MATERIAL IS PROVIDED AS IS, WITH ABSOLUTELY NO WARRANTY EXPRESSED OR IMPLIED.  ANY USE
//   IS AT YOUR OWN RISK.
//   Permission is hereby granted to use, copy, modify, and/or distribute this program,
//   provided that the Copyright, this License, and the Availability of the original version is
//   retained on all copies and made accessible to the end-user of any code or package that
//   includes COLAMD or any modified version of COLAMD.
//
//   Availability:
//   The colamd/symamd library is available at http://www.cise.ufl.edu/research/sparse/colamd/
//
// Column approximate minimum degree ordering with the exact step and tie-break semantics of
// COLAMD 2.3 as vendored by the reference (SuperLU_5.0/SRC/colamd.c, colamd.h:19-21), so that
// the column order - and therefore SuperLU's pivot sequence - is reproduced bit for bit.
//
// Same algorithm, own data layout: the per-column and per-row records are kept as parallel
// arrays, and each pair of fields that the reference stores in one union is kept in ONE array
// here too (their aliasing is part of the behaviour: e.g. a column's degree-list "prev" link,
// its hash-bucket head and its hash value share storage).  The work array A has the usable
// length of colamd_recommended (2 nnz + n_col + nnz/5) so garbage collection happens at the
// same steps.  Defaults: dense-row and dense-column knobs 0.5 (colamd_set_defaults).
#include <algorithm>
#include <climits>
#include <vector>

#include "symbolic.hpp"

namespace nuppn {

namespace {

constexpr int EMPTY = -1;
constexpr int ALIVE = 0;
constexpr int DEAD = -1;
constexpr int DEAD_PRINCIPAL = -1;
constexpr int DEAD_NON_PRINCIPAL = -2;

inline int ones_complement(int r) { return -r - 1; }

struct Cols {
    std::vector<int> start, length;
    std::vector<int> thick_parent;   // thickness (alive) | parent (dead)
    std::vector<int> score_order;    // score (alive) | order (dead)
    std::vector<int> prev_hash;      // headhash | hash | prev
    std::vector<int> next;           // degree_next | hash_next
    explicit Cols(int n)
        : start(n + 1), length(n + 1), thick_parent(n + 1), score_order(n + 1),
          prev_hash(n + 1), next(n + 1) {}
    bool dead(int c) const { return start[c] < ALIVE; }
    bool alive(int c) const { return start[c] >= ALIVE; }
    bool dead_principal(int c) const { return start[c] == DEAD_PRINCIPAL; }
    void kill_principal(int c) { start[c] = DEAD_PRINCIPAL; }
    void kill_non_principal(int c) { start[c] = DEAD_NON_PRINCIPAL; }
};

struct Rows {
    std::vector<int> start, length;
    std::vector<int> degree_p;       // degree | p
    std::vector<int> mark_first;     // mark | first_column
    explicit Rows(int m) : start(m + 1), length(m + 1), degree_p(m + 1), mark_first(m + 1) {}
    bool dead(int r) const { return mark_first[r] < ALIVE; }
    bool alive(int r) const { return mark_first[r] >= ALIVE; }
    void kill(int r) { mark_first[r] = DEAD; }
};

int clear_mark(int n_row, Rows &R) {
    for (int r = 0; r < n_row; ++r)
        if (R.alive(r)) R.mark_first[r] = 0;
    return 1;
}

// build the row form after the column form in A; returns false on invalid input
bool init_rows_cols(int n_row, int n_col, Rows &R, Cols &C, std::vector<int> &A,
                    std::vector<int> &p) {
    for (int col = 0; col < n_col; ++col) {
        C.start[col] = p[col];
        C.length[col] = p[col + 1] - p[col];
        if (C.length[col] < 0) return false;
        C.thick_parent[col] = 1;
        C.score_order[col] = 0;
        C.prev_hash[col] = EMPTY;
        C.next[col] = EMPTY;
    }
    bool jumbled = false;
    for (int row = 0; row < n_row; ++row) {
        R.length[row] = 0;
        R.mark_first[row] = -1;
    }
    for (int col = 0; col < n_col; ++col) {
        int last_row = -1;
        for (int q = p[col]; q < p[col + 1]; ++q) {
            const int row = A[q];
            if (row < 0 || row >= n_row) return false;
            if (row <= last_row || R.mark_first[row] == col) jumbled = true;
            if (R.mark_first[row] != col) R.length[row]++;
            else C.length[col]--;
            R.mark_first[row] = col;
            last_row = row;
        }
    }
    R.start[0] = p[n_col];
    R.degree_p[0] = R.start[0];
    R.mark_first[0] = -1;
    for (int row = 1; row < n_row; ++row) {
        R.start[row] = R.start[row - 1] + R.length[row - 1];
        R.degree_p[row] = R.start[row];
        R.mark_first[row] = -1;
    }
    if (jumbled) {
        for (int col = 0; col < n_col; ++col)
            for (int q = p[col]; q < p[col + 1]; ++q) {
                const int row = A[q];
                if (R.mark_first[row] != col) {
                    A[R.degree_p[row]++] = col;
                    R.mark_first[row] = col;
                }
            }
    } else {
        for (int col = 0; col < n_col; ++col)
            for (int q = p[col]; q < p[col + 1]; ++q) A[R.degree_p[A[q]]++] = col;
    }
    for (int row = 0; row < n_row; ++row) {
        R.mark_first[row] = 0;
        R.degree_p[row] = R.length[row];
    }
    if (jumbled) {
        for (int col = 0; col < n_col; ++col) p[col] = C.length[col];
        for (int row = 0; row < n_row; ++row)
            for (int q = R.start[row]; q < R.start[row] + R.length[row]; ++q) p[A[q]]--;
        C.start[0] = 0;
        p[0] = C.start[0];
        for (int col = 1; col < n_col; ++col) {
            C.start[col] = C.start[col - 1] + C.length[col - 1];
            p[col] = C.start[col];
        }
        for (int row = 0; row < n_row; ++row)
            for (int q = R.start[row]; q < R.start[row] + R.length[row]; ++q) A[p[A[q]]++] = row;
    }
    return true;
}

void init_scoring(int n_row, int n_col, Rows &R, Cols &C, std::vector<int> &A,
                  std::vector<int> &head, double knob_row, double knob_col, int *p_n_row2,
                  int *p_n_col2, int *p_max_deg) {
    // MAX (0, MIN (knob * n, n)) evaluated in double, converted to int (truncation)
    const int dense_row_count = (int) std::max(0.0, std::min(knob_row * n_col, (double) n_col));
    const int dense_col_count = (int) std::max(0.0, std::min(knob_col * n_row, (double) n_row));
    int max_deg = 0, n_col2 = n_col, n_row2 = n_row;
    // empty columns last, natural order
    for (int c = n_col - 1; c >= 0; --c)
        if (C.length[c] == 0) {
            C.score_order[c] = --n_col2;
            C.kill_principal(c);
        }
    // dense columns last, natural order
    for (int c = n_col - 1; c >= 0; --c) {
        if (C.dead(c)) continue;
        if (C.length[c] > dense_col_count) {
            C.score_order[c] = --n_col2;
            for (int q = C.start[c]; q < C.start[c] + C.length[c]; ++q) R.degree_p[A[q]]--;
            C.kill_principal(c);
        }
    }
    // dense and empty rows
    for (int r = 0; r < n_row; ++r) {
        const int deg = R.degree_p[r];
        if (deg > dense_row_count || deg == 0) {
            R.kill(r);
            --n_row2;
        } else {
            max_deg = std::max(max_deg, deg);
        }
    }
    // initial scores, pruning dead rows from the columns
    for (int c = n_col - 1; c >= 0; --c) {
        if (C.dead(c)) continue;
        int score = 0;
        int w = C.start[c];
        for (int q = C.start[c]; q < C.start[c] + C.length[c]; ++q) {
            const int row = A[q];
            if (R.dead(row)) continue;
            A[w++] = row;
            score += R.degree_p[row] - 1;
            score = std::min(score, n_col);
        }
        const int col_length = w - C.start[c];
        if (col_length == 0) {
            C.score_order[c] = --n_col2;
            C.kill_principal(c);
        } else {
            C.length[c] = col_length;
            C.score_order[c] = score;
        }
    }
    for (int c = 0; c <= n_col; ++c) head[c] = EMPTY;
    // degree lists, inserted in reverse so low column indices come first
    for (int c = n_col - 1; c >= 0; --c) {
        if (!C.alive(c)) continue;
        const int score = C.score_order[c];
        const int next_col = head[score];
        C.prev_hash[c] = EMPTY;
        C.next[c] = next_col;
        if (next_col != EMPTY) C.prev_hash[next_col] = c;
        head[score] = c;
    }
    *p_n_col2 = n_col2;
    *p_n_row2 = n_row2;
    *p_max_deg = max_deg;
}

int garbage_collection(int n_row, int n_col, Rows &R, Cols &C, std::vector<int> &A, int pfree) {
    int pdest = 0;
    for (int c = 0; c < n_col; ++c) {
        if (!C.alive(c)) continue;
        int psrc = C.start[c];
        C.start[c] = pdest;
        const int length = C.length[c];
        for (int j = 0; j < length; ++j) {
            const int r = A[psrc++];
            if (R.alive(r)) A[pdest++] = r;
        }
        C.length[c] = pdest - C.start[c];
    }
    for (int r = 0; r < n_row; ++r) {
        if (!R.alive(r)) continue;
        if (R.length[r] == 0) {
            R.kill(r);
        } else {
            const int psrc = R.start[r];
            R.mark_first[r] = A[psrc];
            A[psrc] = ones_complement(r);
        }
    }
    int psrc = pdest;
    while (psrc < pfree) {
        if (A[psrc++] < 0) {
            psrc--;
            const int r = ones_complement(A[psrc]);
            A[psrc] = R.mark_first[r];
            R.start[r] = pdest;
            const int length = R.length[r];
            for (int j = 0; j < length; ++j) {
                const int c = A[psrc++];
                if (C.alive(c)) A[pdest++] = c;
            }
            R.length[r] = pdest - R.start[r];
        }
    }
    return pdest;
}

void detect_super_cols(Cols &C, std::vector<int> &A, std::vector<int> &head, int row_start,
                       int row_length) {
    for (int rp = row_start; rp < row_start + row_length; ++rp) {
        const int col = A[rp];
        if (C.dead(col)) continue;
        const int hash = C.prev_hash[col];
        const int head_column = head[hash];
        const int first_col = (head_column > EMPTY) ? C.prev_hash[head_column] : -(head_column + 2);
        for (int super_c = first_col; super_c != EMPTY; super_c = C.next[super_c]) {
            const int length = C.length[super_c];
            int prev_c = super_c;
            for (int c = C.next[super_c]; c != EMPTY; c = C.next[c]) {
                if (C.length[c] != length || C.score_order[c] != C.score_order[super_c]) {
                    prev_c = c;
                    continue;
                }
                int i = 0;
                const int s1 = C.start[super_c], s2 = C.start[c];
                for (; i < length; ++i)
                    if (A[s1 + i] != A[s2 + i]) break;
                if (i != length) {
                    prev_c = c;
                    continue;
                }
                C.thick_parent[super_c] += C.thick_parent[c];
                C.thick_parent[c] = super_c;
                C.kill_non_principal(c);
                C.score_order[c] = EMPTY;
                C.next[prev_c] = C.next[c];
            }
        }
        if (head_column > EMPTY) C.prev_hash[head_column] = EMPTY;
        else head[hash] = EMPTY;
    }
}

int find_ordering(int n_row, int n_col, int Alen, Rows &R, Cols &C, std::vector<int> &A,
                  std::vector<int> &head, int n_col2, int max_deg, int pfree) {
    const int max_mark = INT_MAX - n_col;
    int tag_mark = clear_mark(n_row, R);
    int min_score = 0, ngarbage = 0;
    for (int k = 0; k < n_col2;) {
        // pivot column from the head of the minimum degree list
        while (head[min_score] == EMPTY && min_score < n_col) min_score++;
        const int pivot_col = head[min_score];
        int next_col = C.next[pivot_col];
        head[min_score] = next_col;
        if (next_col != EMPTY) C.prev_hash[next_col] = EMPTY;
        const int pivot_col_score = C.score_order[pivot_col];
        C.score_order[pivot_col] = k;
        const int pivot_col_thickness = C.thick_parent[pivot_col];
        k += pivot_col_thickness;
        const int needed_memory = std::min(pivot_col_score, n_col - k);
        if (pfree + needed_memory >= Alen) {
            pfree = garbage_collection(n_row, n_col, R, C, A, pfree);
            ngarbage++;
            tag_mark = clear_mark(n_row, R);
        }
        // pivot row: union of the rows of the pivot column
        const int pivot_row_start = pfree;
        int pivot_row_degree = 0;
        C.thick_parent[pivot_col] = -pivot_col_thickness;
        for (int cp = C.start[pivot_col]; cp < C.start[pivot_col] + C.length[pivot_col]; ++cp) {
            const int row = A[cp];
            if (R.dead(row)) continue;
            for (int rp = R.start[row]; rp < R.start[row] + R.length[row]; ++rp) {
                const int col = A[rp];
                const int col_thickness = C.thick_parent[col];
                if (col_thickness > 0 && C.alive(col)) {
                    C.thick_parent[col] = -col_thickness;
                    A[pfree++] = col;
                    pivot_row_degree += col_thickness;
                }
            }
        }
        C.thick_parent[pivot_col] = pivot_col_thickness;
        max_deg = std::max(max_deg, pivot_row_degree);
        for (int cp = C.start[pivot_col]; cp < C.start[pivot_col] + C.length[pivot_col]; ++cp)
            R.kill(A[cp]);
        const int pivot_row_length = pfree - pivot_row_start;
        const int pivot_row = (pivot_row_length > 0) ? A[C.start[pivot_col]] : EMPTY;

        // approximate degree: set differences of the rows in each pivot-row column
        for (int rp = pivot_row_start; rp < pivot_row_start + pivot_row_length; ++rp) {
            const int col = A[rp];
            const int col_thickness = -C.thick_parent[col];
            C.thick_parent[col] = col_thickness;
            const int cur_score = C.score_order[col];
            const int prev_col = C.prev_hash[col];
            next_col = C.next[col];
            if (prev_col == EMPTY) head[cur_score] = next_col;
            else C.next[prev_col] = next_col;
            if (next_col != EMPTY) C.prev_hash[next_col] = prev_col;
            for (int cp = C.start[col]; cp < C.start[col] + C.length[col]; ++cp) {
                const int row = A[cp];
                const int row_mark = R.mark_first[row];
                if (row_mark < ALIVE) continue;
                int set_difference = row_mark - tag_mark;
                if (set_difference < 0) set_difference = R.degree_p[row];
                set_difference -= col_thickness;
                if (set_difference == 0) R.kill(row);
                else R.mark_first[row] = set_difference + tag_mark;
            }
        }
        // scores, hash buckets
        for (int rp = pivot_row_start; rp < pivot_row_start + pivot_row_length; ++rp) {
            const int col = A[rp];
            unsigned int hash = 0;
            int cur_score = 0;
            int w = C.start[col];
            for (int cp = C.start[col]; cp < C.start[col] + C.length[col]; ++cp) {
                const int row = A[cp];
                const int row_mark = R.mark_first[row];
                if (row_mark < ALIVE) continue;
                A[w++] = row;
                hash += row;
                cur_score += row_mark - tag_mark;
                cur_score = std::min(cur_score, n_col);
            }
            C.length[col] = w - C.start[col];
            if (C.length[col] == 0) {
                C.kill_principal(col);
                pivot_row_degree -= C.thick_parent[col];
                C.score_order[col] = k;
                k += C.thick_parent[col];
            } else {
                C.score_order[col] = cur_score;
                hash %= (unsigned int) (n_col + 1);
                const int head_column = head[hash];
                int first_col;
                if (head_column > EMPTY) {
                    first_col = C.prev_hash[head_column];
                    C.prev_hash[head_column] = col;
                } else {
                    first_col = -(head_column + 2);
                    head[hash] = -(col + 2);
                }
                C.next[col] = first_col;
                C.prev_hash[col] = (int) hash;
            }
        }
        detect_super_cols(C, A, head, pivot_row_start, pivot_row_length);
        C.kill_principal(pivot_col);
        tag_mark += (max_deg + 1);
        if (tag_mark >= max_mark) tag_mark = clear_mark(n_row, R);
        // finalise the new pivot row and put its columns back in the degree lists
        int w = pivot_row_start;
        for (int rp = pivot_row_start; rp < pivot_row_start + pivot_row_length; ++rp) {
            const int col = A[rp];
            if (C.dead(col)) continue;
            A[w++] = col;
            A[C.start[col] + (C.length[col]++)] = pivot_row;
            int cur_score = C.score_order[col] + pivot_row_degree;
            const int max_score = n_col - k - C.thick_parent[col];
            cur_score -= C.thick_parent[col];
            cur_score = std::min(cur_score, max_score);
            C.score_order[col] = cur_score;
            next_col = head[cur_score];
            C.next[col] = next_col;
            C.prev_hash[col] = EMPTY;
            if (next_col != EMPTY) C.prev_hash[next_col] = col;
            head[cur_score] = col;
            min_score = std::min(min_score, cur_score);
        }
        if (pivot_row_degree > 0) {
            R.start[pivot_row] = pivot_row_start;
            R.length[pivot_row] = w - pivot_row_start;
            R.degree_p[pivot_row] = pivot_row_degree;
            R.mark_first[pivot_row] = 0;
        }
    }
    return ngarbage;
}

void order_children(int n_col, Cols &C, std::vector<int> &p) {
    for (int i = 0; i < n_col; ++i) {
        if (!C.dead_principal(i) && C.score_order[i] == EMPTY) {
            int parent = i;
            do {
                parent = C.thick_parent[parent];
            } while (!C.dead_principal(parent));
            int c = i;
            int order = C.score_order[parent];
            do {
                C.score_order[c] = order++;
                C.thick_parent[c] = parent;
                c = C.thick_parent[c];
            } while (C.score_order[c] == EMPTY);
            C.score_order[parent] = order;
        }
    }
    for (int c = 0; c < n_col; ++c) p[C.score_order[c]] = c;
}

}  // namespace

int colamd_order(int n_row, int n_col, const int *colptr, const int *rowind,
                 std::vector<int> &perm) {
    if (n_row < 0 || n_col < 0 || colptr[0] != 0) return -1;
    const int nnz = colptr[n_col];
    if (nnz < 0) return -1;
    const int Alen = 2 * nnz + n_col + nnz / 5;   // colamd_recommended minus Col/Row records
    std::vector<int> A(std::max(Alen, 1), 0);
    std::vector<int> p(n_col + 1);
    for (int i = 0; i <= n_col; ++i) p[i] = colptr[i];
    for (int i = 0; i < nnz; ++i) A[i] = rowind[i];
    Rows R(n_row);
    Cols C(n_col);
    if (!init_rows_cols(n_row, n_col, R, C, A, p)) return -2;
    int n_row2, n_col2, max_deg;
    init_scoring(n_row, n_col, R, C, A, p, 0.5, 0.5, &n_row2, &n_col2, &max_deg);
    find_ordering(n_row, n_col, Alen, R, C, A, p, n_col2, max_deg, 2 * nnz);
    order_children(n_col, C, p);
    perm.assign(p.begin(), p.begin() + n_col);
    return 0;
}

}  // namespace nuppn
