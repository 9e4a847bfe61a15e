/* Accessors compiled together with the reference's SuperLU 5.0 sources into _ref/libslu_ref.so.
 * Test infrastructure only.  They expose what c_fortran_dgssv_ / slu_solve compute internally so
 * the tests can compare integer results (column order, elimination tree, row pivots, factor
 * sizes) bit for bit:
 *   - ref_colperm: get_perm_c(COLAMD) followed by sp_preorder (column etree + postorder), the
 *     ordering dgssv uses (SuperLU_5.0/SRC/dgssv.c, get_perm_c.c:405, sp_preorder.c:140-150);
 *   - ref_factor_stats: perm_c, perm_r, nnz(L), nnz(U) held by a c_fortran_dgssv_ handle
 *     (the handle is a pointer to {L, U, perm_c, perm_r}, FORTRAN/c_fortran_dgssv.c:27-32,141-147).
 */
#include "slu_ddefs.h"

typedef struct {
    SuperMatrix *L;
    SuperMatrix *U;
    int *perm_c;
    int *perm_r;
} ref_factors_handle;

int ref_colperm(int n, int nnz, const int *colptr, const int *rowind,
                int *perm_colamd, int *perm_final, int *etree)
{
    SuperMatrix A, AC;
    superlu_options_t options;
    int i;
    int *cp = intMalloc(n + 1), *ri = intMalloc(nnz > 0 ? nnz : 1);
    double *val = doubleMalloc(nnz > 0 ? nnz : 1);
    for (i = 0; i <= n; ++i) cp[i] = colptr[i];
    for (i = 0; i < nnz; ++i) { ri[i] = rowind[i]; val[i] = 1.0; }
    dCreate_CompCol_Matrix(&A, n, n, nnz, val, ri, cp, SLU_NC, SLU_D, SLU_GE);
    set_default_options(&options);
    get_perm_c(options.ColPerm, &A, perm_colamd);
    for (i = 0; i < n; ++i) perm_final[i] = perm_colamd[i];
    sp_preorder(&options, &A, perm_final, etree, &AC);
    Destroy_CompCol_Permuted(&AC);
    Destroy_CompCol_Matrix(&A);
    return 0;
}

int ref_factor_stats(long long handle, int n, int *perm_c, int *perm_r, int *nnz_l, int *nnz_u,
                     int *nsuper)
{
    ref_factors_handle *f = (ref_factors_handle *) handle;
    SCformat *Ls = (SCformat *) f->L->Store;
    NCformat *Us = (NCformat *) f->U->Store;
    int i;
    for (i = 0; i < n; ++i) { perm_c[i] = f->perm_c[i]; perm_r[i] = f->perm_r[i]; }
    *nnz_l = Ls->nnz;
    *nnz_u = Us->nnz;
    *nsuper = Ls->nsuper + 1;
    return 0;
}

/* supernode partition of a factored handle: first column of each supernode (nsuper+1 entries) */
int ref_factor_supernodes(long long handle, int *xsup_out)
{
    ref_factors_handle *f = (ref_factors_handle *) handle;
    SCformat *Ls = (SCformat *) f->L->Store;
    int i;
    for (i = 0; i <= Ls->nsuper + 1; ++i) xsup_out[i] = Ls->sup_to_col[i];
    return Ls->nsuper + 1;
}
