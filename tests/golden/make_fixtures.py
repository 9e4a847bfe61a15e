"""Writes the reference's own known-answer data for the sparse LU as small fixtures.

Sources (read from the reference checkout; data only):
  * solver/source/SuperLU_5.0/EXAMPLE/g20.rua  - Harwell-Boeing matrix used by the Fortran example
  * solver/source/SuperLU_5.0/FORTRAN/f77exm.out - the recorded output of that example
    (f77_main.f: c_fortran_dgssv iopt=1/2/3 with b = ones): nnz(L), nnz(U), first 10 x
  * solver/source/slu_f90/matrixtest.f90:11-22 - the 20x20 CSR system (values, 0-based
    indices, right-hand side) fed to slu_solve
Run:  python tests/golden/make_fixtures.py
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/solver/source"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_hb(path):
    lines = open(path).read().splitlines()
    totcrd, ptrcrd, indcrd, valcrd = [int(t) for t in lines[1].split()[:4]]
    mxtype, nrow, ncol, nnz = lines[2].split()[:4]
    nrow, ncol, nnz = int(nrow), int(ncol), int(nnz)
    body = lines[4:]
    ptr_lines = body[:ptrcrd]
    ind_lines = body[ptrcrd:ptrcrd + indcrd]
    val_lines = body[ptrcrd + indcrd:ptrcrd + indcrd + valcrd]

    def ints(ls, width):
        out = []
        for l in ls:
            for i in range(0, len(l), width):
                t = l[i:i + width].strip()
                if t:
                    out.append(int(t))
        return out

    colptr = np.array(ints(ptr_lines, 5), dtype=np.int32) - 1
    rowind = np.array(ints(ind_lines, 5), dtype=np.int32) - 1
    vals = []
    for l in val_lines:
        for i in range(0, len(l), 15):
            t = l[i:i + 15].strip()
            if t:
                vals.append(float(t.replace("D", "E")))
    vals = np.array(vals, dtype=np.float64)
    assert colptr.size == ncol + 1 and rowind.size == nnz and vals.size == nnz
    return nrow, colptr, rowind, vals


def main():
    n, cp, ri, va = read_hb(os.path.join(REF, "SuperLU_5.0/EXAMPLE/g20.rua"))
    np.savez(os.path.join(HERE, "g20.npz"), n=n, colptr=cp, rowind=ri, values=va)

    out = open(os.path.join(REF, "SuperLU_5.0/FORTRAN/f77exm.out")).read()
    nl = int(re.search(r"factor L = (\d+)", out).group(1))
    nu = int(re.search(r"factor U = (\d+)", out).group(1))
    nlu = int(re.search(r"L\+U = (\d+)", out).group(1))
    tail = out.split("Solve succeeded")[1]
    x10 = [float(t) for t in tail.split()[:10]]
    json.dump({"nnz_l": nl, "nnz_u": nu, "nnz_lu": nlu, "x_first10": x10, "b": "ones"},
              open(os.path.join(HERE, "f77exm.json"), "w"), indent=1)

    src = open(os.path.join(REF, "slu_f90/matrixtest.f90")).read()

    def arr(name):
        m = re.search(name + r"\s*=\s*\[(.*?)\]", src, re.S)
        return [float(t) for t in m.group(1).replace("&", " ").replace(",", " ").split()]
    json.dump({"n": 20, "val": arr("val"), "colind0": [int(t) for t in arr("colind")],
               "rowptr0": [int(t) for t in arr("rowptr")], "rhs": arr("rhs")},
              open(os.path.join(HERE, "matrixtest.json"), "w"), indent=1)
    print("wrote g20.npz f77exm.json matrixtest.json")


if __name__ == "__main__":
    main()
