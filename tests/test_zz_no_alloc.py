"""GPU test of slu_solve_noalloc (the allocation-free slu_solve): through the C ABI, against
slu_solve on the same inputs - a 1,100-species network matrix, then a smaller one (matrixtest.f90's
20x20 system) in the next call, so the second call runs over the first call's leftovers in the
shared scratch array - bit for bit, with the same index shift and error codes; a system beyond
the fixed scratch is refused with NUPPN_ERR_UNSUPPORTED before anything is touched."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.hip


def _network_csr(pkg, oracle):
    w = pkg.network.config(1, nz=2)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    v = o.rates(w.t9[1, 0], w.rho[1, 0])
    n, rp, ci, va, mask = o.reduced_matrix(w.dt[1], w.x0[1], v)
    rhs = np.random.default_rng(7).standard_normal(n)
    return n, va, ci, rp, rhs


def _matrixtest():
    m = json.load(open(os.path.join(GOLD, "matrixtest.json")))
    return (m["n"], np.array(m["val"]), np.array(m["colind0"], np.int32),
            np.array(m["rowptr0"], np.int32), np.array(m["rhs"]))


def _both(hip, n, val, ci0, rp0, rhs):
    out = []
    for fn in (hip.slu_solve, hip.slu_solve_noalloc):
        ci1 = (ci0 + 1).astype(np.int32)
        rp1 = (rp0 + 1).astype(np.int32)
        b = rhs.copy()
        ierr = fn(val.copy(), ci1, rp1, b)
        out.append((ierr, b, ci1, rp1))
    return out


def test_noalloc_matches_slu_solve(pkg, hip, oracle):
    big = _network_csr(pkg, oracle)
    small = _matrixtest()
    for n, val, ci0, rp0, rhs in (big, small):
        print(f"n = {n}, nnz = {len(val)}", flush=True)
        (ie_a, b_a, ci_a, rp_a), (ie_b, b_b, ci_b, rp_b) = _both(hip, n, val, ci0, rp0, rhs)
        assert ie_a == ie_b == 0
        np.testing.assert_array_equal(b_b, b_a)
        # the index arrays come back shifted to 0-based, as slu_solve (and the reference) leaves
        np.testing.assert_array_equal(ci_b, ci0)
        np.testing.assert_array_equal(rp_b, rp0)
    # a singular system: the same SuperLU info (column index + 1)
    n, val, ci0, rp0, rhs = small
    (ie_a, *_), (ie_b, *_) = _both(hip, n, np.zeros_like(val), ci0, rp0, rhs)
    assert ie_a == ie_b > 0


def test_noalloc_refuses_beyond_scratch(hip):
    n = 3000   # dense-capacity factors of n = 3,000 need about 108 MB
    val = np.ones(n)
    ci1 = np.arange(1, n + 1, dtype=np.int32)
    rp1 = np.arange(1, n + 2, dtype=np.int32)
    b = np.ones(n)
    ierr = hip.slu_solve_noalloc(val, ci1, rp1, b)
    assert ierr == hip.ERR_UNSUPPORTED
    np.testing.assert_array_equal(ci1, np.arange(1, n + 1))
    np.testing.assert_array_equal(b, 1.0)
