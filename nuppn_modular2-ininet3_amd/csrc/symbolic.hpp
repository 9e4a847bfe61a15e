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
    std::vector<int> dx_ptr, dx_ent;           // per reduced species: (l<<2 | role)
    std::vector<int> perm_c;                   // column order, SuperLU convention
};

// builds everything but the column order; returns 0 or a negative error
int build_symbolic(const HostNetwork &net, Symbolic &s);

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
