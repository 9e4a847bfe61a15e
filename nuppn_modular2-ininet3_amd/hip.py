"""Python host mirror of the C ABI in include/nuppn_hip.h (ctypes).

The names follow the reference's solver interface: ``Context.integrate_network`` is
integrate_network (solver/source/solver.F90:53) preceded by the cycle's evaluate_all_rates, and
``Context.integrate_batch`` replaces mppnp's per-zone worker loop (frames/mppnp/source/
mppnp.F90:713-753).  ``slu_solve`` and ``c_fortran_dgssv`` mirror the two SuperLU entry points
(solver/source/slu_f90/slu_f90.c:3, SuperLU_5.0/FORTRAN/c_fortran_dgssv.c:35).

The library is the in-tree build (lib/libnuppn_hip.so).  There is no fallback: if it is missing
or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnuppn_hip.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_PATTERN, ERR_SINGULAR, ERR_UNSUPPORTED, ERR_MIXING = -1, -2, -3, -4, -5, -6
IERR_NAMES = {0: "OK", 1: "CONVERGENCE", 2: "MAX_STEPS", 3: "BAD_STEP", 4: "NEGATIVE_X",
              5: "HUGE_X", 6: "SUM_X", 7: "NAN", 11: "DT_UNDRFLOW", -3: "PATTERN",
              -4: "SINGULAR"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Knobs(C.Structure):
    """nuppn_knobs = solver_knobs.F90's namelist &ppn_solver."""
    _fields_ = [("ittd", C.c_int), ("irdn", C.c_int), ("max_sub_steps", C.c_int),
                ("integration_method", C.c_int), ("interpolate_temp", C.c_int),
                ("interpolate_dens", C.c_int), ("dgrd", C.c_double), ("grdthreshold", C.c_double),
                ("cyminformal", C.c_double), ("min_rate_ov_dt", C.c_double),
                ("min_t9_for_slv", C.c_double), ("delta_state_max", C.c_double)]


class ZoneStats(C.Structure):
    _fields_ = [("nsubt", C.c_int), ("titers", C.c_int), ("iter", C.c_int), ("nvar1", C.c_int),
                ("nsucc", C.c_int), ("nfail", C.c_int), ("n_newton", C.c_longlong),
                ("n_factor", C.c_longlong), ("n_solve", C.c_longlong), ("lu_nnz", C.c_longlong),
                ("n_refine", C.c_longlong), ("n_factor_fp64", C.c_longlong)]


class BatchOpts(C.Structure):
    """nuppn_batch_opts: factor precision (32 / 64) and the FP64 refinement tolerances."""
    _fields_ = [("factor_precision", C.c_int), ("max_refine", C.c_int),
                ("refine_tol", C.c_double), ("accept_tol", C.c_double),
                ("team_zones", C.c_int), ("team_waves", C.c_int),
                ("queue_passes", C.c_int), ("use_graph", C.c_int)]


class MppnpOpts(C.Structure):
    """nuppn_mppnp_opts: zoneselect threshold, mixing limit and matrix form, integration path."""
    _fields_ = [("t9threshold", C.c_double), ("sig_term_limit", C.c_double),
                ("code_source", C.c_int), ("path", C.c_int)]


class RatePhysicsC(C.Structure):
    """nuppn_rate_physics"""
    _fields_ = [("pf_n", C.c_int), ("pf_t9", C.c_void_p), ("pf_tab", C.c_void_p),
                ("rev_ch", C.c_void_p), ("rev_part", C.c_void_p), ("detailed_balance", C.c_int),
                ("g_spin", C.c_void_p), ("mass_g", C.c_void_p), ("q_mev", C.c_void_p),
                ("screen", C.c_int)]


class TuneResult(C.Structure):
    """nuppn_tune_result"""
    _fields_ = [("block_size", C.c_int), ("team_zones", C.c_int), ("team_waves", C.c_int),
                ("from_cache", C.c_int), ("ncand", C.c_int), ("consistent", C.c_int),
                ("cand_nb", C.c_int * 8), ("cand_tz", C.c_int * 8), ("cand_tw", C.c_int * 8),
                ("ms", C.c_double * 8)]


class Network(C.Structure):
    _fields_ = [("nsp", C.c_int), ("nre", C.c_int), ("k", C.POINTER(C.c_int)),
                ("rtype", C.POINTER(C.c_int)), ("coef", C.POINTER(C.c_double)),
                ("aw", C.POINTER(C.c_double)), ("zw", C.POINTER(C.c_double)),
                ("igamma", C.c_int)]


STATS_DTYPE = np.dtype([("nsubt", np.int32), ("titers", np.int32), ("iter", np.int32),
                        ("nvar1", np.int32), ("nsucc", np.int32), ("nfail", np.int32),
                        ("n_newton", np.int64), ("n_factor", np.int64), ("n_solve", np.int64),
                        ("lu_nnz", np.int64), ("n_refine", np.int64), ("n_factor_fp64", np.int64)])

_lib = None


def load():
    """Load the in-tree HIP library (raises when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    # torch ships its own HIP runtime under the same soname; importing it first makes this
    # library bind to that one, so device pointers from torch tensors and this library's kernels
    # share one runtime in the process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.nuppn_default_knobs.argtypes = [C.POINTER(Knobs)]
    L.nuppn_column_order.argtypes = [C.c_int, _ip, _ip, _ip, C.c_void_p]
    L.nuppn_ctx_create.restype = C.c_void_p
    L.nuppn_ctx_create.argtypes = [C.POINTER(Network), C.POINTER(Knobs), C.POINTER(C.c_int)]
    L.nuppn_ctx_destroy.argtypes = [C.c_void_p]
    L.nuppn_set_column_order.argtypes = [C.c_void_p, _ip]
    L.nuppn_get_column_order.argtypes = [C.c_void_p, _ip]
    L.nuppn_reduced_size.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.nuppn_reduced_pattern.argtypes = [C.c_void_p, _ip, _ip, _ip]
    L.nuppn_rates_eval.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, _dp]
    L.nuppn_dxdt.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.nuppn_jacobian_reduced.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
    L.nuppn_lu_solve_reduced.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
    L.nuppn_integrate_zone.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(C.c_double), C.c_double,
                                       C.POINTER(C.c_int), C.POINTER(ZoneStats)]
    L.nuppn_integrate_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _ip, C.c_void_p]
    L.nuppn_integrate_batch_dev.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 10
    L.nuppn_last_timing.argtypes = [C.c_void_p, _dp, C.c_int]
    L.nuppn_default_mppnp_opts.argtypes = [C.POINTER(MppnpOpts)]
    L.nuppn_set_rate_physics.argtypes = [C.c_void_p, C.c_void_p]
    L.nuppn_autotune.argtypes = ([C.c_void_p, C.c_int] + [_dp] * 6 +
                                 [C.c_char_p, C.POINTER(TuneResult)])
    L.nuppn_reverse_pairs.argtypes = [C.c_void_p, _ip]
    L.nuppn_evaluate_all_rates.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
    L.nuppn_mppnp_cycle.argtypes = ([C.c_void_p, C.c_int] + [_dp] * 7 +
                                    [C.c_double, C.POINTER(MppnpOpts), C.POINTER(BatchOpts), _ip,
                                     C.c_void_p])
    L.nuppn_mppnp_cycle_dev.argtypes = ([C.c_void_p, C.c_int] + [C.c_void_p] * 7 +
                                        [C.c_double, C.POINTER(MppnpOpts), C.POINTER(BatchOpts),
                                         C.c_void_p, C.c_void_p, C.c_void_p])
    L.nuppn_default_batch_opts.argtypes = [C.POINTER(BatchOpts)]
    L.nuppn_batch_block_info.argtypes = [C.c_void_p, _ip, C.c_int]
    L.nuppn_set_block_form.argtypes = [C.c_void_p, C.c_int]
    L.nuppn_integrate_batch_mixed.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                              _dp, _ip, C.c_void_p, C.POINTER(BatchOpts)]
    L.nuppn_integrate_batch_mixed_dev.argtypes = ([C.c_void_p, C.c_int] + [C.c_void_p] * 9 +
                                                  [C.POINTER(BatchOpts), C.c_void_p])
    L.slu_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp,
                            C.POINTER(C.c_int)]
    L.slu_solve_noalloc.argtypes = L.slu_solve.argtypes
    L.nuppn_set_rate_cache.argtypes = [C.c_void_p, C.c_int]
    L.nuppn_rate_cache_table.argtypes = [C.c_void_p, _dp]
    L.c_fortran_dgssv_.argtypes = [C.POINTER(C.c_int)] * 4 + [_dp, _ip, _ip, _dp,
                                                               C.POINTER(C.c_int),
                                                               C.POINTER(C.c_longlong),
                                                               C.POINTER(C.c_int)]
    _lib = L
    return L


def default_knobs(**over) -> Knobs:
    k = Knobs()
    load().nuppn_default_knobs(C.byref(k))
    for key, val in over.items():
        setattr(k, key, val)
    return k


def batch_opts(**over) -> BatchOpts:
    o = BatchOpts()
    load().nuppn_default_batch_opts(C.byref(o))
    for key, val in over.items():
        setattr(o, key, val)
    return o


def mppnp_opts(**over) -> MppnpOpts:
    o = MppnpOpts()
    load().nuppn_default_mppnp_opts(C.byref(o))
    for k, v in over.items():
        setattr(o, k, v)
    return o


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def column_order(n, colptr, rowind, with_etree=False):
    """SuperLU's column order (COLAMD + column-etree post-order) computed on the host."""
    pc = np.zeros(n, np.int32)
    et = np.zeros(n, np.int32)
    rc = load().nuppn_column_order(n, _i32(colptr), _i32(rowind), pc, et.ctypes.data_as(C.c_void_p))
    if rc != OK:
        raise RuntimeError(f"nuppn_column_order failed ({rc})")
    return (pc, et) if with_etree else pc


class Context:
    """A network on the device (nuppn_ctx)."""

    def __init__(self, net, knobs: Knobs | None = None):
        L = load()
        self.net = net
        self._k = _i32(net.k.reshape(-1))
        self._rtype = _i32(net.rtype)
        self._coef = _f64(net.coef.reshape(-1))
        self._aw = _f64(net.aw)
        self._zw = _f64(net.zw)
        nw = Network(net.nsp, net.nre, self._k.ctypes.data_as(C.POINTER(C.c_int)),
                     self._rtype.ctypes.data_as(C.POINTER(C.c_int)),
                     self._coef.ctypes.data_as(C.POINTER(C.c_double)),
                     self._aw.ctypes.data_as(C.POINTER(C.c_double)),
                     self._zw.ctypes.data_as(C.POINTER(C.c_double)), 0)
        self.knobs = knobs if knobs is not None else default_knobs()
        st = C.c_int(0)
        self.h = L.nuppn_ctx_create(C.byref(nw), C.byref(self.knobs), C.byref(st))
        if not self.h:
            raise RuntimeError(f"nuppn_ctx_create failed ({st.value})")
        n, nnz = C.c_int(0), C.c_int(0)
        L.nuppn_reduced_size(self.h, C.byref(n), C.byref(nnz))
        self.n, self.nnzA = n.value, nnz.value

    def close(self):
        if getattr(self, "h", None):
            load().nuppn_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- symbolic
    def reduced_pattern(self):
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(self.nnzA, np.int32)
        r2s = np.zeros(self.n, np.int32)
        load().nuppn_reduced_pattern(self.h, cp, ri, r2s)
        return cp, ri, r2s

    def column_order(self):
        pc = np.zeros(self.n, np.int32)
        load().nuppn_get_column_order(self.h, pc)
        return pc

    def set_column_order(self, perm_c):
        rc = load().nuppn_set_column_order(self.h, _i32(perm_c))
        if rc != OK:
            raise RuntimeError(f"nuppn_set_column_order failed ({rc})")

    # -- single-state pieces (test hooks)
    def rates(self, t9, rho):
        t9 = np.atleast_1d(_f64(t9))
        rho = np.atleast_1d(_f64(rho))
        v = np.zeros(t9.size * self.net.nre)
        rc = load().nuppn_rates_eval(self.h, t9.size, t9, rho, None, v)
        if rc != OK:
            raise RuntimeError(f"nuppn_rates_eval failed ({rc})")
        return v.reshape(t9.size, self.net.nre)

    def autotune(self, yps, t9_0, t9_1, rho_0, rho_1, dtwant, cache_path=None):
        """nuppn_autotune on a sample batch: the chosen shapes and the candidates' times"""
        r = TuneResult()
        y = _f64(yps)
        rc = load().nuppn_autotune(self.h, y.shape[0], y, _f64(t9_0), _f64(t9_1), _f64(rho_0),
                                   _f64(rho_1), _f64(dtwant),
                                   None if cache_path is None else str(cache_path).encode(),
                                   C.byref(r))
        if rc != OK:
            raise RuntimeError(f"nuppn_autotune failed ({rc})")
        n = r.ncand
        return {"block_size": r.block_size, "team_zones": r.team_zones,
                "team_waves": r.team_waves, "from_cache": bool(r.from_cache),
                "consistent": bool(r.consistent),
                "candidates": [(r.cand_nb[i], r.cand_tz[i], r.cand_tw[i], r.ms[i])
                               for i in range(n)]}

    def set_rate_physics(self, rp):
        """rp: network.RatePhysics, or None for the bare REACLIB sum"""
        if rp is None:
            rc = load().nuppn_set_rate_physics(self.h, None)
        else:
            keep = [_f64(rp.pf_t9), _f64(rp.pf_tab.reshape(-1)), _i32(rp.rev_ch),
                    _i32(rp.rev_part.reshape(-1)), _f64(rp.g_spin), _f64(rp.mass_g),
                    _f64(rp.q_mev)]
            c = RatePhysicsC(rp.pf_t9.size, *(a.ctypes.data for a in keep[:4]),
                             int(rp.detailed_balance), *(a.ctypes.data for a in keep[4:]),
                             int(rp.screen))
            rc = load().nuppn_set_rate_physics(self.h, C.byref(c))
        if rc != OK:
            raise RuntimeError(f"nuppn_set_rate_physics failed ({rc})")

    def reverse_pairs(self):
        out = np.zeros(self.net.nre, np.int32)
        rc = load().nuppn_reverse_pairs(self.h, out)
        if rc != OK:
            raise RuntimeError(f"nuppn_reverse_pairs failed ({rc})")
        return out

    def evaluate_all_rates(self, t9, rho, x):
        t9 = np.atleast_1d(_f64(t9))
        rho = np.atleast_1d(_f64(rho))
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        v = np.zeros(t9.size * self.net.nre)
        rc = load().nuppn_evaluate_all_rates(self.h, t9.size, t9, rho, x, v)
        if rc != OK:
            raise RuntimeError(f"nuppn_evaluate_all_rates failed ({rc})")
        return v.reshape(t9.size, self.net.nre)

    def dxdt(self, x, v):
        out = np.zeros(self.net.nsp)
        rc = load().nuppn_dxdt(self.h, _f64(x), _f64(v), out)
        if rc != OK:
            raise RuntimeError(f"nuppn_dxdt failed ({rc})")
        return out

    def set_rate_cache(self, use_cache=True):
        """nuppn_set_rate_cache: the REACLIB T9 cache on (built on first use) or off"""
        rc = load().nuppn_set_rate_cache(self.h, int(use_cache))
        if rc != OK:
            raise RuntimeError(f"nuppn_set_rate_cache failed ({rc})")

    def rate_cache_table(self):
        out = np.zeros(5001 * self.net.nre)
        rc = load().nuppn_rate_cache_table(self.h, out)
        if rc != OK:
            raise RuntimeError(f"nuppn_rate_cache_table failed ({rc})")
        return out.reshape(5001, self.net.nre)

    def jacobian_reduced(self, dt, x, v):
        out = np.zeros(self.nnzA)
        rc = load().nuppn_jacobian_reduced(self.h, float(dt), _f64(x), _f64(v), out)
        return rc, out

    def lu_solve(self, vals, b=None):
        b2 = None if b is None else _f64(np.atleast_2d(b)).copy()
        pr = np.zeros(self.n, np.int32)
        nlu = np.zeros(2, np.int32)
        rc = load().nuppn_lu_solve_reduced(
            self.h, _f64(vals), 0 if b2 is None else b2.shape[0],
            None if b2 is None else b2.ctypes.data_as(C.c_void_p), pr.ctypes.data_as(C.c_void_p),
            nlu.ctypes.data_as(C.c_void_p))
        return rc, (None if b2 is None else b2.reshape(np.shape(b))), pr, nlu

    # -- integration
    def integrate_network(self, yps, t9_0, t9_1, rho_0, rho_1, dtwant, ye=0.5):
        """evaluate_all_rates(mid-point) + integrate_network for one zone."""
        y = _f64(yps).copy()
        yeo = C.c_double(ye)
        ierr = C.c_int(0)
        st = ZoneStats()
        rc = load().nuppn_integrate_zone(self.h, y, float(t9_0), float(t9_1), float(rho_0),
                                         float(rho_1), C.byref(yeo), float(dtwant), C.byref(ierr),
                                         C.byref(st))
        if rc != OK:
            raise RuntimeError(f"nuppn_integrate_zone failed ({rc})")
        return ierr.value, y, yeo.value, st

    def integrate_batch(self, yps, t9_0, t9_1, rho_0, rho_1, dtwant, ye=None):
        y = _f64(yps).copy()
        nz = y.shape[0]
        ye = np.full(nz, 0.5) if ye is None else _f64(ye).copy()
        ierr = np.zeros(nz, np.int32)
        stats = np.zeros(nz, STATS_DTYPE)
        rc = load().nuppn_integrate_batch(self.h, nz, y, _f64(t9_0), _f64(t9_1), _f64(rho_0),
                                          _f64(rho_1), ye, _f64(dtwant), ierr,
                                          stats.ctypes.data_as(C.c_void_p))
        if rc != OK:
            raise RuntimeError(f"nuppn_integrate_batch failed ({rc})")
        return ierr, y, ye, stats

    def integrate_batch_dev(self, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats,
                            stream=0):
        """Device-pointer variant (torch tensors' data_ptr() or raw addresses)."""
        rc = load().nuppn_integrate_batch_dev(self.h, nz, yps, t9_0, t9_1, rho_0, rho_1, ye,
                                              dtwant, ierr, stats, stream)
        if rc != OK:
            raise RuntimeError(f"nuppn_integrate_batch_dev failed ({rc})")

    def set_block_form(self, nb):
        """rebuilds the batched path as block size 32 / 64 or the band form (-1)
        (nuppn_set_block_form); returns the status code."""
        return load().nuppn_set_block_form(self.h, int(nb))

    def block_info(self):
        """dict of the batched path's block structure (nuppn_batch_block_info)."""
        info = np.zeros(7, np.int32)
        rc = load().nuppn_batch_block_info(self.h, info, 7)
        keys = ("NB", "nblk", "nbd", "ni", "bandwidth", "slots", "lds_bytes")
        d = dict(zip(keys, info.tolist()))
        d["available"] = rc == OK
        return d

    def integrate_batch_mixed(self, yps, t9_0, t9_1, rho_0, rho_1, dtwant, ye=None, opts=None):
        """The batched path (block factorisation + FP64 refinement) on host arrays."""
        y = _f64(yps).copy()
        nz = y.shape[0]
        ye = np.full(nz, 0.5) if ye is None else _f64(ye).copy()
        ierr = np.zeros(nz, np.int32)
        stats = np.zeros(nz, STATS_DTYPE)
        o = opts if opts is not None else batch_opts()
        rc = load().nuppn_integrate_batch_mixed(self.h, nz, y, _f64(t9_0), _f64(t9_1),
                                                _f64(rho_0), _f64(rho_1), ye, _f64(dtwant), ierr,
                                                stats.ctypes.data_as(C.c_void_p), C.byref(o))
        if rc != OK:
            raise RuntimeError(f"nuppn_integrate_batch_mixed failed ({rc})")
        return ierr, y, ye, stats

    def integrate_batch_mixed_dev(self, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr,
                                  stats, opts=None, stream=0):
        """Device-pointer variant of the batched path."""
        o = opts if opts is not None else batch_opts()
        rc = load().nuppn_integrate_batch_mixed_dev(self.h, nz, yps, t9_0, t9_1, rho_0, rho_1, ye,
                                                    dtwant, ierr, stats, C.byref(o), stream)
        if rc != OK:
            raise RuntimeError(f"nuppn_integrate_batch_mixed_dev failed ({rc})")

    def mppnp_cycle(self, yps, t9, rho, xm, dq, rad, d_eul, dzeit, opts=None, bopts=None):
        """One mppnp cycle (zoneselect, batch integration, renormalisation, mixing) on host
        arrays, yps[m][nsp] zone-major.  Returns (rc, ierr, yps, stats); rc is NUPPN_OK or
        NUPPN_ERR_MIXING (other codes raise)."""
        y = _f64(yps).copy()
        m = y.shape[0]
        ierr = np.zeros(m, np.int32)
        stats = np.zeros(m, STATS_DTYPE)
        mo = opts if opts is not None else mppnp_opts()
        bo = bopts if bopts is not None else batch_opts()
        rc = load().nuppn_mppnp_cycle(self.h, m, y, _f64(t9), _f64(rho), _f64(xm), _f64(dq),
                                      _f64(rad), _f64(d_eul), float(dzeit), C.byref(mo),
                                      C.byref(bo), ierr, stats.ctypes.data_as(C.c_void_p))
        if rc not in (OK, ERR_MIXING):
            raise RuntimeError(f"nuppn_mppnp_cycle failed ({rc})")
        return rc, ierr, y, stats

    def mppnp_cycle_dev(self, m, yps, t9, rho, xm, dq, rad, d_eul, dzeit, ierr, stats=None,
                        opts=None, bopts=None, stream=0):
        """Device-pointer variant of mppnp_cycle; returns the code (OK or ERR_MIXING)."""
        mo = opts if opts is not None else mppnp_opts()
        bo = bopts if bopts is not None else batch_opts()
        rc = load().nuppn_mppnp_cycle_dev(self.h, m, yps, t9, rho, xm, dq, rad, d_eul,
                                          float(dzeit), C.byref(mo), C.byref(bo), ierr, stats,
                                          stream)
        if rc not in (OK, ERR_MIXING):
            raise RuntimeError(f"nuppn_mppnp_cycle_dev failed ({rc})")
        return rc

    def last_timing(self):
        """dict: rates_ms, total_ms, newton_ms, control_ms, newton_launches, passes,
        graph_launches"""
        out = np.zeros(7)
        rc = load().nuppn_last_timing(self.h, out, 7)
        if rc != OK:
            raise RuntimeError(f"nuppn_last_timing failed ({rc})")
        keys = ("rates_ms", "total_ms", "newton_ms", "control_ms", "newton_launches", "passes",
                "graph_launches")
        return dict(zip(keys, out.tolist()))


def slu_solve(val, colind, rowptr, rhs, isrow=1, shiftindex=1):
    """slu_solve(n, nnz, isrow, shiftindex, val, colind, rowptr, rhs, ierr): rhs is overwritten;
    with shiftindex the index arrays are shifted to 0-based in place, as the reference does."""
    n = rhs.shape[0]
    ierr = C.c_int(0)
    load().slu_solve(n, val.shape[0], isrow, shiftindex, val, colind, rowptr, rhs, C.byref(ierr))
    return ierr.value


def slu_solve_noalloc(val, colind, rowptr, rhs, isrow=1, shiftindex=1):
    """slu_solve_noalloc: slu_solve with every device buffer in the library's fixed scratch
    array (no device allocation); the same arguments and results."""
    n = rhs.shape[0]
    ierr = C.c_int(0)
    load().slu_solve_noalloc(n, val.shape[0], isrow, shiftindex, val, colind, rowptr, rhs,
                             C.byref(ierr))
    return ierr.value


def c_fortran_dgssv(iopt, n, nnz, nrhs, values, rowind, colptr, b, ldb, factors):
    """c_fortran_dgssv_(iopt, ...) with Fortran (1-based CSC) conventions; returns (info, handle)."""
    io, nn, nz, nr, lb = (C.c_int(iopt), C.c_int(n), C.c_int(nnz), C.c_int(nrhs), C.c_int(ldb))
    f = C.c_longlong(factors)
    info = C.c_int(0)
    load().c_fortran_dgssv_(C.byref(io), C.byref(nn), C.byref(nz), C.byref(nr), values, rowind,
                            colptr, b, C.byref(lb), C.byref(f), C.byref(info))
    return info.value, f.value
