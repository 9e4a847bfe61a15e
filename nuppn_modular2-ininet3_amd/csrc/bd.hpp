// Bader-Deuflhard step on the device: bader_deuflhard_integration (solver/source/
// bader_deuflhard.F90:139-272), semi_implicit_extrapolation and bader_deuflhard_step
// (276-316, 410-488), extrapolate (371-407) and compute_error (325-366), for one zone on one
// wavefront.  The linear algebra comes from a solver object:
//   int jacobian(y, dt)  J at y once per step (calculate_jacobian); 0 or E_BAD_STEP
//   int level(h)         I - h J assembled and factored; 0, E_BAD_STEP (a zero where the
//                        reference stops) or a negative library error
//   void dxdt(y, out)    reduced dX/dt at y (calculate_dxdt + reduce_rhs_bd)
//   int solve(rhs, x)    x = (I - h J)^-1 rhs; 0 or a negative library error
// The plain solver below factors with the reference's GEPP (pivot order of SuperLU); the
// batched path's block solver is in batched.hpp.
#pragma once
#include "zone.hpp"

namespace nuppn {

// per-wavefront arrays of one BD step: yk, yext, yscal [nsp]; delta, rhs, x [n];
// qcol, yerr [8 * nsp] (column c of the reference's (nsp, nlevels) tableau at c * nsp);
// sortbuf [pow2 >= nsp]
struct BDWork {
    double *yk, *yext, *yscal, *delta, *rhs, *x, *qcol, *yerr, *sortbuf;
};

// extrapolate (bader_deuflhard.F90:371-407): Neville's scheme towards x -> 0 over the
// abscissae xcol; each lane runs the whole recurrence for its species
__device__ inline void bd_extrapolate(int nsp, int k, double xest, const double *xcol,
                                      const double *yest, double *yext, double *qcol,
                                      double *yerr) {
    double *yk_err = yerr + (size_t) (k - 1) * nsp;
    for (int j = lane_id(); j < nsp; j += WAVE) {
        const double ye = yest[j];
        yk_err[j] = ye;
        if (k == 1) {
            qcol[j] = ye;
            yext[j] = ye;
            continue;
        }
        double yz = ye, d = ye;
        for (int k1 = 1; k1 <= k - 1; ++k1) {
            const double delta = 1.0 / (xcol[k - k1 - 1] - xest);
            const double f1 = xest * delta;
            const double f2 = xcol[k - k1 - 1] * delta;
            const double q = qcol[(size_t) (k1 - 1) * nsp + j];
            qcol[(size_t) (k1 - 1) * nsp + j] = yk_err[j];
            const double dl = d - q;
            yk_err[j] = f1 * dl;
            d = f2 * dl;
            yz = yz + yk_err[j];
        }
        qcol[(size_t) (k - 1) * nsp + j] = yk_err[j];
        yext[j] = yz;
    }
    wave_sync();
}

// compute_error (bader_deuflhard.F90:325-366)
__device__ inline void bd_compute_error(const DevNet &N, const DevKnobs &K, int k,
                                        const double *yk, const double *yscal, const double *yerr,
                                        double *sortbuf, BDState &bd) {
    const int nsp = N.nsp;
    const double *e_k = yerr + (size_t) (k - 1) * nsp;
    double ab = -__DBL_MAX__, ng = -__DBL_MAX__;
    for (int j = lane_id(); j < nsp; j += WAVE) {
        const double e = fabs(e_k[j] / yscal[j]);
        if (yk[j] > 1.0e-27) ab = fmax(ab, e);
        if (yk[j] < K.grdthreshold) ng = fmax(ng, -yk[j]);
    }
    ab = wave_max(ab);
    ng = wave_max(ng);
    bd.abuerr = ab / K.dgrd;
    bd.negerr = ng / K.grdthreshold;
    if (bd.abuerr < 1.0e-3) {
        bd.errmax = bd.abuerr;
    } else {
        bd.sumerr = fabs(1.0 - sorted_sum(yk, nsp, sortbuf)) / SUM_TOL;
        bd.errmax = fmax(bd.abuerr, bd.sumerr);
    }
    bd.errmax = fmax(bd.errmax, bd.negerr);
}

// any y > 2 (bader_deuflhard_step, bader_deuflhard.F90:468-479) or any NaN
__device__ inline bool wave_any_above(const double *y, int n, double lim) {
    bool f = false;
    for (int i = lane_id(); i < n; i += WAVE) f = f || (y[i] > lim);
    return ballot(f);
}
__device__ inline bool wave_any_nan(const double *y, int n) {
    bool f = false;
    for (int i = lane_id(); i < n; i += WAVE) f = f || (y[i] != y[i]);
    return ballot(f);
}

// One step of bader_deuflhard_integration for zone z from yps = Z.yps[z] over s.dttry.  Sets
// s.step_ierr and s.dtrcmd (or s.lib_err), updates bd, and leaves the new abundances in yps
// unless the step ended early (the caller restores yps from yps_old on any error).
template <class Solver>
__device__ void bd_step(const KArgs &a, int z, ZoneState &s, BDState &bd, const BDWork &W,
                        Solver &sv) {
    const DevNet &N = a.N;
    const DevKnobs &K = a.K;
    const BDTables &T = a.T;
    const int nsp = N.nsp, n = N.n;
    const double safe1 = 0.25, safe2 = 0.7, redmin = 0.7, redmax = 1.0e-5, scalmx = 0.25;
    double *yps = zrow(a.Z.yps, a.Z.nsp_stride, z);
    const double dt = s.dttry;
    double scal = 1.0, red = -1.0, dtrcmd = 0.0, xcol[8], err[8];
    int km = -1, ierr = 0, k = 1;
    bool copy_out = true;
    s.st.iter = 0;
    if (sv.jacobian(yps, dt)) {
        s.step_ierr = E_BAD_STEP;
        s.dtrcmd = dt * 0.2;
        return;
    }
    for (k = 1; k <= T.kmax; ++k) {
        copy_vec(W.yk, yps, nsp);
        // semi_implicit_extrapolation (276-316): h and the quirky abscissa (h / nseq)^2
        const int ns = T.nseq[k - 1];
        const double h = dt / (double) ns;
        const double xest = (h / ns) * (h / ns);
        const int e = sv.level(h);
        if (e > 0) {   // reduce_jacobian's stop (semi_implicit_extrapolation, 287-289)
            s.step_ierr = E_BAD_STEP;
            s.dtrcmd = dt * 0.2;
            return;
        }
        if (e) {
            s.lib_err = e;
            return;
        }
        bool huge = false;
        for (int stage = 0; stage <= ns && !huge; ++stage) {
            // FIRST: rhs = h f(y); MID: h f(y) - delta, delta += 2x; LAST: h f(y) - delta,
            // delta = x; y += delta after each (bader_deuflhard_step, 410-488)
            sv.dxdt(W.yk, W.rhs);
            for (int r = lane_id(); r < n; r += WAVE)
                W.rhs[r] = (stage == 0) ? h * W.rhs[r] : h * W.rhs[r] - W.delta[r];
            wave_sync();
            const int e2 = sv.solve(W.rhs, W.x);
            if (e2) {
                s.lib_err = e2;
                return;
            }
            const bool mid = stage > 0 && stage < ns;
            for (int r = lane_id(); r < n; r += WAVE) {
                const double dl = mid ? W.delta[r] + 2.0 * W.x[r] : W.x[r];
                W.delta[r] = dl;
                const int sp = N.red2sp[r];
                W.yk[sp] = W.yk[sp] + dl;
            }
            wave_sync();
            huge = wave_any_above(W.yk, nsp, 2.0);
        }
        if (huge) {
            ierr = E_HUGE_ABUNDS;
            dtrcmd = dt * 0.2;
            copy_out = false;
            break;
        }
        s.st.iter += ns;
        s.st.nvar1 = sv.nvar1;
        xcol[k - 1] = xest;
        bd_extrapolate(nsp, k, xest, xcol, W.yk, W.yext, W.qcol, W.yerr);
        copy_vec(W.yk, W.yext, nsp);
        if (wave_any_nan(W.yk, nsp)) {
            ierr = E_GOT_NAN;
            dtrcmd = dt * 0.2;
            copy_out = false;
            break;
        }
        for (int j = lane_id(); j < nsp; j += WAVE) W.yscal[j] = fmax(fabs(W.yk[j]), K.grdthreshold);
        wave_sync();
        if (k != 1) {
            bd_compute_error(N, K, k, W.yk, W.yscal, W.yerr, W.sortbuf, bd);
            km = k - 1;
            err[km - 1] = pow(bd.errmax / safe1, 1.0 / (2.0 * km + 1.0));
        }
        // convergence in the window kopt-1 .. kopt+1, or step reduction (204-246)
        if (k != 1 && (k >= bd.kopt - 1 || bd.frst)) {
            if (bd.errmax <= 1.0) {
                bd.reduce = 0;
                break;
            }
            if (k == T.kmax || k == bd.kopt + 1) {
                bd.reduce = 1;
                red = safe2 / err[km - 1];
            } else if (k == bd.kopt) {
                if (T.alf[bd.kopt - 2][bd.kopt - 1] < err[km - 1]) {
                    bd.reduce = 1;
                    red = safe2 / err[km - 1];
                }
            } else if (bd.kopt == T.kmax) {
                if (T.alf[km - 1][T.kmax - 2] < err[km - 1]) {
                    bd.reduce = 1;
                    red = T.alf[km - 1][T.kmax - 2] * safe2 / err[km - 1];
                }
            } else if (T.alf[km - 1][bd.kopt - 1] < err[km - 1]) {
                bd.reduce = 1;
                red = T.alf[km - 1][bd.kopt - 2] / err[km - 1];
            }
        }
    }
    if (copy_out) {
        if (bd.reduce) {   // 247-252
            red = fmax(redmax, fmin(red, redmin));
            dtrcmd = dt * red;
            if (bd.negerr > bd.sumerr && bd.negerr > bd.abuerr) ierr = E_NEG_ABUNDS;
            else if (bd.sumerr > bd.abuerr) ierr = E_SUM_X;
            else ierr = E_SLV_CNVRG;
        } else {           // order and step size for the next step (253-268)
            double workmin = 1.0e35;
            for (int kk = 1; kk <= km; ++kk) {
                const double fact = fmax(err[kk - 1], scalmx);
                const double work = fact * T.a[kk];
                if (work < workmin) {
                    scal = fact;
                    workmin = work;
                    bd.kopt = kk + 1;
                }
            }
            dtrcmd = 2.0 * dt / scal;
            if (bd.kopt >= k && bd.kopt != T.kmax) {
                const double fact = fmax(scal / T.alf[bd.kopt - 2][bd.kopt - 1], scalmx);
                if (T.a[bd.kopt] * fact <= workmin) {
                    dtrcmd = dtrcmd * scal / fact;
                    bd.kopt = bd.kopt + 1;
                }
            }
        }
        copy_vec(yps, W.yk, nsp);
    }
    s.step_ierr = ierr;
    s.dtrcmd = dtrcmd;
}

// The plain solver: the reference's matrix (jac_cpy * (-h) + I, reduce_jacobian's drops
// reported as a pattern error) factored by the wavefront GEPP with SuperLU's column order and
// pivot policy, solved by its triangular solves.
struct PlainBDSolver {
    const KArgs &a;
    const Slot &S;
    LUWork w;
    ZoneState &s;
    const double *v;
    int nvar1;

    // calculate_jacobian (jac_rhs.F90:282-455) once per step: the mask and the entries of the
    // reduced matrix (every entry of every species on the static network), the 1e-99 seed, and
    // the stop on an exact zero off the diagonal among the entries kept (444-451)
    __device__ int jacobian(const double *y, double dt) {
        const DevNet &N = a.N;
        const DynWork DW{S.first, S.rflag, S.eflag};
        molar(N, y, S.xn);
        wave_sync();
        reaction_terms(N, S.xn, v, S.vx, S.st1, S.st2);
        wave_sync();
        if (a.K.dynamic) {
            nvar1 = dyn_network<Team<1>>(N, a.K, dt, S.xn, v, S.st1, S.st2, N.dyn_q, N.nnzA, DW,
                                         nullptr);
        } else {
            for (int r = lane_id(); r < N.n; r += WAVE) S.first[N.red2sp[r]] = 0;
            for (int q = lane_id(); q < N.nnzA; q += WAVE) S.eflag[q] = 1;
            nvar1 = N.n;
            wave_sync();
        }
        int bad = 0;
        for (int q = lane_id(); q < N.nnzA; q += WAVE) {
            const double jv = jac_entry(N, q, S.st1, S.st2);
            S.jac_cpy[q] = jv;
            if (!N.A_isdiag[q] && S.eflag[q] && dyn_marked(DW, N.red2sp[N.A_rowind[q]]) &&
                jv == 0.0)
                bad = 1;
        }
        wave_sync();
        return ballot(bad) ? E_BAD_STEP : 0;
    }

    // I - h J with reduce_jacobian's drops (jac_rhs.F90:138-161); E_BAD_STEP where the reference
    // stops.  A dropped entry leaves the reduced matrix for the rest of the step, as the
    // reference's row(ice) = -1 stays set until the next calculate_jacobian; species outside
    // the mask keep identity rows.
    __device__ int level(double h) {
        const DevNet &N = a.N;
        const DynWork DW{S.first, S.rflag, S.eflag};
        int stop = 0;
        for (int q = lane_id(); q < N.nnzA; q += WAVE) {
            const bool diag = N.A_isdiag[q];
            double val;
            if (!dyn_marked(DW, N.red2sp[N.A_rowind[q]])) {
                val = diag ? 1.0 : 0.0;
            } else if (diag) {
                val = reduced_value(S.jac_cpy[q], h, true, stop);
            } else if (!S.eflag[q]) {
                val = 0.0;
            } else {
                val = S.jac_cpy[q] * (-h);
                if (fabs(val) == 0.0 || val == -1.0) {
                    S.eflag[q] = 0;
                    val = 0.0;
                }
            }
            S.Aval[q] = val;
        }
        wave_sync();
        if (ballot(stop)) return E_BAD_STEP;
        const int info = gepp_factor(N.n, N.A_colptr, N.A_rowind, S.Aval, N.iperm_c, w, S.lu);
        s.st.n_factor++;
        if (info) return -4;
        s.st.lu_nnz += (long long) w.Lp[N.n] + (long long) S.lu.Up[N.n];
        return 0;
    }

    __device__ void dxdt(const double *y, double *out) {
        const DevNet &N = a.N;
        molar(N, y, S.xn);
        wave_sync();
        reaction_terms(N, S.xn, v, S.vx, S.st1, S.st2);
        wave_sync();
        for (int r = lane_id(); r < N.n; r += WAVE) out[r] = dxdt_species(N, r, S.vx);
        wave_sync();
    }

    // reduce_rhs_bd packs the masked species; the others get x = 0 (bader_deuflhard_step, 420)
    __device__ int solve(const double *rhs, double *x) {
        const DevNet &N = a.N;
        const DynWork DW{S.first, S.rflag, S.eflag};
        for (int r = lane_id(); r < N.n; r += WAVE)
            x[r] = dyn_marked(DW, N.red2sp[r]) ? rhs[r] : 0.0;
        wave_sync();
        gepp_solve(N.n, N.perm_c, w, S.lu, x);
        s.st.n_solve++;
        return 0;
    }
};

// the plain path's worker for a zone in PH_NEWTON under integration_method 1
__device__ inline void zone_bd(const KArgs &a, const Slot &S, LUWork w, int z) {
    const ZoneData &Z = a.Z;
    ZoneState s = Z.state[z];
    if (s.phase != PH_NEWTON) return;
    BDState bd = Z.bd[z];
    PlainBDSolver sv{a, S, w, s, zrow(Z.v, Z.nre_stride, z), a.N.n};
    BDWork W{S.y_cur, S.yext, S.yscal, S.delta, S.rhs, S.y_new, S.qcol, S.yerr, S.sortbuf};
    bd_step(a, z, s, bd, W, sv);
    s.phase = PH_AFTER_NEWTON;
    if (lane_id() == 0) {
        Z.state[z] = s;
        Z.bd[z] = bd;
    }
    wave_sync();
}

}  // namespace nuppn
