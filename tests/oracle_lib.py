"""ctypes access to the CPU oracle (oracle/_build/libnuppn_oracle.so) and, through it, to the
reference SuperLU compiled from source (oracle/_ref/libslu_ref.so).  Test infrastructure: only the
tests, the smoke check and the CPU-baseline leg of the benchmark use this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libnuppn_oracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libslu_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

LS_GEPP = 0
LS_REF = 1
LS_EXACT = 2   # GEPP + quad-precision residual refinement: the FP64 system solved exactly


class Knobs(C.Structure):
    _fields_ = [("ittd", C.c_int), ("irdn", C.c_int), ("max_sub_steps", C.c_int),
                ("integration_method", C.c_int), ("interpolate_temp", C.c_int),
                ("interpolate_dens", C.c_int), ("dgrd", C.c_double), ("grdthreshold", C.c_double),
                ("cyminformal", C.c_double), ("min_rate_ov_dt", C.c_double),
                ("min_t9_for_slv", C.c_double), ("delta_state_max", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("nsubt", C.c_int), ("titers", C.c_int), ("iter", C.c_int), ("nvar1", C.c_int),
                ("nsucc", C.c_int), ("nfail", C.c_int), ("n_newton", C.c_longlong),
                ("n_factor", C.c_longlong), ("n_solve", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build()
    L = C.CDLL(ORACLE_SO)
    L.orc_default_knobs.argtypes = [C.POINTER(Knobs)]
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.POINTER(Knobs)]
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_set_knobs.argtypes = [C.c_void_p, C.POINTER(Knobs)]
    L.orc_set_linear_solver.argtypes = [C.c_void_p, C.c_int]
    L.orc_reset_state.argtypes = [C.c_void_p]
    L.orc_set_rates.argtypes = [C.c_void_p, _dp]
    L.orc_bd_state.argtypes = [C.c_void_p, _ip, _dp]
    L.orc_set_rate_cache.argtypes = [C.c_void_p, C.c_int]
    L.orc_rate_cache.restype = C.POINTER(C.c_double)
    L.orc_rate_cache.argtypes = [C.c_void_p]
    L.orc_bd_set_frst.argtypes = [C.c_void_p, C.c_int]
    L.orc_nnz_full.argtypes = [C.c_void_p]
    L.orc_full_pattern.argtypes = [C.c_void_p, _ip, _ip, _ip]
    L.orc_rates.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_dxdt.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.orc_jacobian.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, _ip, _ip, _bp]
    L.orc_reduce_jacobian.argtypes = [C.c_void_p, _dp, _ip, _ip, _bp, C.POINTER(C.c_int), _ip,
                                      _ip, _dp]
    L.orc_backward_euler.argtypes = [C.c_void_p, _dp, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(Stats)]
    L.orc_bader_deuflhard.argtypes = [C.c_void_p, _dp, C.c_double, C.POINTER(C.c_double),
                                      C.POINTER(Stats)]
    L.orc_integrate_zone.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.POINTER(C.c_double), C.c_double,
                                     C.POINTER(Stats)]
    L.orc_integrate_zone_v.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(C.c_double), C.c_double,
                                       C.POINTER(Stats)]
    L.orc_sortedsum.restype = C.c_double
    L.orc_sortedsum.argtypes = [C.c_int, _dp]
    L.orc_calculate_ye.restype = C.c_double
    L.orc_calculate_ye.argtypes = [C.c_void_p, _dp]
    L.orc_ypsplus.argtypes = [C.c_int, _dp]
    L.orc_bd_tables.argtypes = [_ip, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_lu_factor.argtypes = [C.c_int, _ip, _ip, _dp, C.c_void_p, C.POINTER(C.c_void_p)]
    L.orc_lu_solve.argtypes = [C.c_void_p, _dp]
    L.orc_lu_info.argtypes = [C.c_void_p, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_lu_free.argtypes = [C.c_void_p]
    _bp1 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    L.orc_zoneselect.argtypes = [C.c_double, _dp, C.c_int, _bp1]
    L.orc_ksum.restype = C.c_double
    L.orc_ksum.argtypes = [C.c_int, _dp]
    L.orc_mass_residual.restype = C.c_double
    L.orc_mass_residual.argtypes = [C.c_int, _dp]
    L.orc_renorm_1d.argtypes = [C.c_int, _dp]
    L.orc_mix_d_lagrangian.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.orc_matrixdiffus.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_double, _dp, _dp, _dp]
    L.orc_matrixdiffus_mesa.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_double, _dp, _dp,
                                        _dp]
    L.orc_tridag.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_do_mixing.argtypes = [C.c_int, C.c_int, _dp, _bp1, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                C.c_int, C.c_double]
    L.orc_mppnp_cycle.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                  C.c_double, C.c_double, C.c_int, C.c_double, _ip,
                                  C.POINTER(Stats)]
    L.orc_set_rate_physics.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _ip, _ip, C.c_int,
                                       C.c_int, _dp, _dp, _dp]
    L.orc_reverse_pairs.argtypes = [C.c_void_p, _ip]
    L.orc_rates_full.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _dp]
    L.orc_ref_open.argtypes = [C.c_char_p]
    L.orc_ref_slu_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp,
                                    C.POINTER(C.c_int)]
    L.orc_ref_dgssv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp, C.c_int,
                                C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    L.orc_ref_colperm.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip]
    L.orc_ref_factor_stats.argtypes = [C.c_longlong, C.c_int, _ip, _ip, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_ref_factor_supernodes.argtypes = [C.c_longlong, _ip]
    _lib = L
    return L


def ref_available() -> bool:
    L = lib()
    if os.path.exists(REF_SO):
        L.orc_ref_open(REF_SO.encode())
    return bool(L.orc_ref_available())


def default_knobs(**over) -> Knobs:
    k = Knobs()
    lib().orc_default_knobs(C.byref(k))
    for key, val in over.items():
        setattr(k, key, val)
    return k


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One oracle context over a network (network.Network)."""

    def __init__(self, net, knobs: Knobs | None = None, linear_solver: int = LS_GEPP):
        L = lib()
        self.net = net
        self.knobs = knobs if knobs is not None else default_knobs()
        self._k = i32(net.k.reshape(-1))
        self.h = L.orc_create(net.nsp, net.nre, self._k, i32(net.rtype), f64(net.coef.reshape(-1)),
                              f64(net.aw), f64(net.zw), 0, C.byref(self.knobs))
        if linear_solver == LS_REF and not ref_available():
            raise RuntimeError("reference SuperLU (oracle/_ref) is not built")
        if linear_solver == LS_EXACT:
            ref_available()   # its factors when built (faster), else the oracle GEPP's
        L.orc_set_linear_solver(self.h, linear_solver)
        self.nnz = L.orc_nnz_full(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_destroy(self.h)
            self.h = None

    def set_knobs(self, knobs: Knobs):
        self.knobs = knobs
        lib().orc_set_knobs(self.h, C.byref(knobs))

    def reset_state(self):
        lib().orc_reset_state(self.h)

    def full_pattern(self):
        n = self.net.nsp
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        mp = np.zeros(8 * self.net.nre, np.int32)
        lib().orc_full_pattern(self.h, rp, ci, mp)
        return rp, ci, mp.reshape(8, self.net.nre)

    def rates(self, t9, rho, ye=0.5):
        v = np.zeros(self.net.nre)
        lib().orc_rates(self.h, float(t9), float(rho), float(ye), v)
        return v

    def set_rate_physics(self, rp):
        """rp: network.RatePhysics (None switches the rate physics off)"""
        L = lib()
        if rp is None:
            z = np.zeros(1)
            L.orc_set_rate_physics(self.h, 0, z, z, np.zeros(1, np.int32), np.zeros(1, np.int32),
                                   0, 0, z, z, z)
            return
        L.orc_set_rate_physics(self.h, rp.pf_t9.size, f64(rp.pf_t9), f64(rp.pf_tab.reshape(-1)),
                               i32(rp.rev_ch), i32(rp.rev_part.reshape(-1)), int(rp.screen),
                               int(rp.detailed_balance), f64(rp.g_spin), f64(rp.mass_g),
                               f64(rp.q_mev))

    def reverse_pairs(self):
        out = np.zeros(self.net.nre, np.int32)
        lib().orc_reverse_pairs(self.h, out)
        return out

    def rates_full(self, t9, rho, x, ye=0.5):
        v = np.zeros(self.net.nre)
        lib().orc_rates_full(self.h, t9, rho, ye, f64(x), v)
        return v

    def dxdt(self, x, v):
        out = np.zeros(self.net.nsp)
        lib().orc_dxdt(self.h, f64(x), f64(v), out)
        return out

    def jacobian(self, dt, x, v):
        jac = np.zeros(self.nnz)
        row = np.zeros(self.nnz, np.int32)
        col = np.zeros(self.nnz, np.int32)
        mask = np.zeros(self.net.nsp, np.uint8)
        rc = lib().orc_jacobian(self.h, float(dt), f64(x), f64(v), jac, row, col, mask)
        return rc, jac, row, col, mask

    def reduced_matrix(self, dt, x, v):
        """(I - dt J) over the masked species as in newton_iteration (backward_euler.F90:133-137)."""
        rc, jac, row, col, mask = self.jacobian(dt, x, v)
        jac = jac * (-dt)
        n1 = C.c_int(0)
        rp = np.zeros(self.net.nsp + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        va = np.zeros(self.nnz)
        nn = lib().orc_reduce_jacobian(self.h, jac, row, col, mask, C.byref(n1), rp, ci, va)
        n = n1.value
        return n, rp[: n + 1].copy(), ci[:nn].copy(), va[:nn].copy(), mask

    def set_rates(self, v):
        lib().orc_set_rates(self.h, f64(v))

    def bd_state(self):
        iv = np.zeros(4, np.int32)
        dv = np.zeros(12)
        lib().orc_bd_state(self.h, iv, dv)
        return dict(frst=int(iv[0]), reduce=int(iv[1]), kopt=int(iv[2]), kmax=int(iv[3]),
                    errmax=dv[0], abuerr=dv[1], negerr=dv[2], sumerr=dv[3], err=dv[4:].copy())

    def set_rate_cache(self, use_cache=True):
        lib().orc_set_rate_cache(self.h, int(use_cache))

    def rate_cache(self):
        """the T9 table [5001, nre] (a copy), or None"""
        p = lib().orc_rate_cache(self.h)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(5001, self.net.nre)).copy()

    def bd_set_frst(self, frst):
        lib().orc_bd_set_frst(self.h, int(frst))

    def backward_euler(self, y, dt, v=None):
        if v is not None:
            self.set_rates(v)
        y = f64(y).copy()
        dtr = C.c_double(0.0)
        st = Stats()
        ierr = lib().orc_backward_euler(self.h, y, float(dt), C.byref(dtr), C.byref(st))
        return ierr, y, dtr.value, st

    def bader_deuflhard(self, y, dt, v=None):
        if v is not None:
            self.set_rates(v)
        y = f64(y).copy()
        dtr = C.c_double(0.0)
        st = Stats()
        ierr = lib().orc_bader_deuflhard(self.h, y, float(dt), C.byref(dtr), C.byref(st))
        return ierr, y, dtr.value, st

    def integrate_zone(self, y, t9_0, t9_1, rho_0, rho_1, dtwant, ye=0.5, v_mid=None):
        y = f64(y).copy()
        yeo = C.c_double(ye)
        st = Stats()
        if v_mid is None:
            ierr = lib().orc_integrate_zone(self.h, y, float(t9_0), float(t9_1), float(rho_0),
                                            float(rho_1), C.byref(yeo), float(dtwant),
                                            C.byref(st))
        else:
            ierr = lib().orc_integrate_zone_v(self.h, f64(v_mid), y, float(t9_0), float(t9_1),
                                              float(rho_0), float(rho_1), C.byref(yeo),
                                              float(dtwant), C.byref(st))
        return ierr, y, yeo.value, st


def mppnp_cycle(o, yps, t9, rho, xm, dq, rad, d_eul, dzeit, t9threshold=0.0, mesa=False,
                sig=1.0e10):
    """The oracle's mppnp cycle (orc_mppnp_cycle) on Oracle o: (rc, ierr, yps, stats list)."""
    y = f64(yps).copy()
    m = y.shape[0]
    ierr = np.zeros(m, np.int32)
    st = (Stats * m)()
    rc = lib().orc_mppnp_cycle(o.h, m, y, f64(t9), f64(rho), f64(xm), f64(dq), f64(rad),
                               f64(d_eul), float(dzeit), float(t9threshold), int(mesa),
                               float(sig), ierr, st)
    return rc, ierr, y, [st[i] for i in range(m)]


def do_mixing(yps, consider, xm, dq, rho, rad, d_eul, dzeit, mesa=False, sig=1.0e10):
    """orc_do_mixing on a copy of yps[m][nsp]: (rc, yps)."""
    y = f64(yps).copy()
    m, nsp = y.shape
    rc = lib().orc_do_mixing(m, nsp, y, np.ascontiguousarray(consider, np.uint8), f64(xm),
                             f64(dq), f64(rho), f64(rad), f64(d_eul), float(dzeit), int(mesa),
                             float(sig))
    return rc, y


def csr_to_csc(n, rowptr, colind, val):
    import scipy.sparse as sp
    m = sp.csr_matrix((val, colind, rowptr), shape=(n, n))
    m.sort_indices()
    c = m.tocsc()
    c.sort_indices()
    return i32(c.indptr), i32(c.indices), f64(c.data)


def gepp(n, colptr, rowind, val, perm_c=None):
    """Oracle GEPP on a 0-based CSC matrix; returns (info, perm_r, nnz_l, nnz_u, solve_fn)."""
    L = lib()
    h = C.c_void_p()
    pc = None if perm_c is None else i32(perm_c)
    info = L.orc_lu_factor(n, i32(colptr), i32(rowind), f64(val),
                           None if pc is None else pc.ctypes.data_as(C.c_void_p), C.byref(h))
    pr = np.zeros(n, np.int32)
    nl, nu = C.c_int(0), C.c_int(0)
    L.orc_lu_info(h, pr, C.byref(nl), C.byref(nu))

    def solve(b):
        x = f64(b).copy()
        L.orc_lu_solve(h, x)
        return x

    class _H:
        def __del__(self_inner):
            L.orc_lu_free(h)
    keep = _H()
    solve._keep = keep
    return info, pr, nl.value, nu.value, solve


def ref_colperm(n, colptr, rowind):
    """Reference get_perm_c(COLAMD) and the final (etree-postordered) perm_c plus etree."""
    if not ref_available():
        raise RuntimeError("reference SuperLU (oracle/_ref) is not built")
    p0 = np.zeros(n, np.int32)
    p1 = np.zeros(n, np.int32)
    et = np.zeros(n, np.int32)
    lib().orc_ref_colperm(n, int(colptr[n]), i32(colptr), i32(rowind), p0, p1, et)
    return p0, p1, et


def ref_factor(n, colptr, rowind, val, b=None):
    """c_fortran_dgssv_ iopt=1 (+2 if b), returning perm_c, perm_r, nnz(L), nnz(U), nsuper, x."""
    if not ref_available():
        raise RuntimeError("reference SuperLU (oracle/_ref) is not built")
    L = lib()
    cp1 = i32(np.asarray(colptr) + 1)
    ri1 = i32(np.asarray(rowind) + 1)
    vals = f64(val).copy()
    nnz = int(colptr[n])
    h = C.c_longlong(0)
    info = C.c_int(0)
    rhs = np.zeros(n) if b is None else f64(b).copy()
    L.orc_ref_dgssv(1, n, nnz, 1, vals, ri1, cp1, rhs, n, C.byref(h), C.byref(info))
    f_info = info.value
    pc = np.zeros(n, np.int32)
    pr = np.zeros(n, np.int32)
    nl, nu, ns = C.c_int(0), C.c_int(0), C.c_int(0)
    L.orc_ref_factor_stats(h.value, n, pc, pr, C.byref(nl), C.byref(nu), C.byref(ns))
    xsup = np.zeros(ns.value + 1, np.int32)
    L.orc_ref_factor_supernodes(h.value, xsup)
    x = None
    if b is not None and f_info == 0:
        L.orc_ref_dgssv(2, n, nnz, 1, vals, ri1, cp1, rhs, n, C.byref(h), C.byref(info))
        x = rhs
    L.orc_ref_dgssv(3, n, nnz, 1, vals, ri1, cp1, rhs, n, C.byref(h), C.byref(info))
    return dict(info=f_info, perm_c=pc, perm_r=pr, nnz_l=nl.value, nnz_u=nu.value,
                nsuper=ns.value, xsup=xsup, x=x)
