// Per-zone driver shared by the kernels: the per-slot scratch layout, the LDS carve-up and
// integrate_network for one zone on one wavefront (solver/source/solver.F90:53-238).
#pragma once
#include "device_common.hpp"
#include "integrate.hpp"
#include "dynamic.hpp"

namespace nuppn {

// ------------------------------------------------------------------------------------------
// per-slot scratch layout

enum SlotArr {
    S_YOLD, S_YCUR, S_YNEW, S_YPS, S_YPSOLD, S_XN, S_DXDT, S_VX, S_ST1, S_ST2, S_VBUF, S_AVAL,
    S_RHS, S_SORT, S_JACCPY, S_DELTA, S_QCOL, S_YERR, S_YSCAL, S_YEXT, S_LI, S_LX, S_UI, S_UX,
    S_UP, S_FIRST, S_RFLAG, S_EFLAG, S_COUNT
};

struct SlotLayout {
    size_t off[S_COUNT];
    size_t bytes;
};

inline SlotLayout slot_layout(int nsp, int nre, int n, int nnzA) {
    const size_t capL = (size_t) n * (n - 1) / 2 + 1, capU = (size_t) n * (n + 1) / 2 + 1;
    size_t pow2 = 1;
    while (pow2 < (size_t) nsp) pow2 <<= 1;
    const size_t len[S_COUNT] = {
        8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * nsp, 8ull * nsp,
        8ull * nre, 8ull * nre, 8ull * nre, 8ull * nre, 8ull * nnzA, 8ull * n, 8ull * pow2,
        8ull * nnzA, 8ull * nsp, 64ull * nsp, 64ull * nsp, 8ull * nsp, 8ull * nsp,
        4ull * capL, 8ull * capL, 4ull * capU, 8ull * capU, 4ull * (n + 1), 4ull * nsp,
        (size_t) nre, (size_t) nnzA};
    SlotLayout L;
    size_t o = 0;
    for (int i = 0; i < S_COUNT; ++i) {
        L.off[i] = o;
        o += (len[i] + 255) & ~(size_t) 255;
    }
    L.bytes = o;
    return L;
}

// Per-zone state of the sub-step controller and the backward-Euler Newton loop, kept in
// global memory between launches (solver.F90:65-107 locals, backward_euler.F90:31-53 locals).
enum ZonePhase { PH_SUBSTEP = 0, PH_NEWTON = 1, PH_AFTER_NEWTON = 2, PH_DONE = 3 };

// pad0: factor precision of the batched path (0 FP32, 1 FP64); step_ierr: the ierr of the last
// Bader-Deuflhard step (its dtrcmd goes to dtrcmd); ye_cur: integrate_network's ye, which
// do_one_time_step recomputes from every step's raw result (solver.F90:272) and a failed step
// restores (solver.F90:190,197)
struct ZoneState {
    int phase, i, iter, ierr, lib_err, solved, pad0, step_ierr;
    double dtdone, dttry, dtrcmd, my_t9_last, my_rho_last, grd, ye_entry, ye_cur;
    ZoneStats st;
};

// per-zone arrays (zone-major, fixed strides) and the I/O of one batch
struct ZoneData {
    double *yps, *yps_old, *y_old, *y_cur, *y_new, *v, *work;
    size_t nsp_stride, nre_stride, work_stride;
    const double *t9_0, *t9_1, *rho_0, *rho_1, *dtwant;
    double *ye;
    int *ierr;
    ZoneStats *stats;   // may be null
    ZoneState *state;
    BDState *bd;        // [nz] Bader-Deuflhard state carried between steps
    BDState bd0;        // the state every zone starts from (the context's)
    int *flag;          // [0] set when some zone needs a Newton iteration or a BD step,
                        // [1] the number of such zones
    int *active;        // [nz] those zones, in no particular order (the workers' work list)
    int nact;           // entries of active the worker launch takes, or, when nact_dev is set,
    const int *nact_dev;   // read on the device from there (the control pass's count)
    const int *ctl;     // the zones the control pass takes: the previous pass's work list, or
    int nctl;           // every zone (ctl null) in the first pass; when nctl_dev is set, the
    const int *nctl_dev;   // count is read on the device from there (the previous pass's)
    int nz;
    int init_prec;      // factor precision a zone starts with on the batched path (0 FP32, 1 FP64)
};

// the length of a work list: the host's value, or the count a previous kernel left on the device
// (passes are queued ahead of the host knowing their counts)
__device__ __forceinline__ int list_len(const int *dev, int n) { return dev ? *dev : n; }

struct KArgs {
    DevNet N;
    DevKnobs K;
    BDTables T;
    ZoneData Z;
    char *slot_base;
    size_t slot_bytes;
    size_t off[S_COUNT];
};

__device__ inline Slot make_slot(char *base, const size_t *off) {
    Slot S;
    S.y_old = (double *) (base + off[S_YOLD]);
    S.y_cur = (double *) (base + off[S_YCUR]);
    S.y_new = (double *) (base + off[S_YNEW]);
    S.yps = (double *) (base + off[S_YPS]);
    S.yps_old = (double *) (base + off[S_YPSOLD]);
    S.xn = (double *) (base + off[S_XN]);
    S.dxdt = (double *) (base + off[S_DXDT]);
    S.vx = (double *) (base + off[S_VX]);
    S.st1 = (double *) (base + off[S_ST1]);
    S.st2 = (double *) (base + off[S_ST2]);
    S.vbuf = (double *) (base + off[S_VBUF]);
    S.Aval = (double *) (base + off[S_AVAL]);
    S.rhs = (double *) (base + off[S_RHS]);
    S.sortbuf = (double *) (base + off[S_SORT]);
    S.jac_cpy = (double *) (base + off[S_JACCPY]);
    S.delta = (double *) (base + off[S_DELTA]);
    S.qcol = (double *) (base + off[S_QCOL]);
    S.yerr = (double *) (base + off[S_YERR]);
    S.yscal = (double *) (base + off[S_YSCAL]);
    S.yext = (double *) (base + off[S_YEXT]);
    S.lu.Li = (int *) (base + off[S_LI]);
    S.lu.Lx = (double *) (base + off[S_LX]);
    S.lu.Ui = (int *) (base + off[S_UI]);
    S.lu.Ux = (double *) (base + off[S_UX]);
    S.lu.Up = (int *) (base + off[S_UP]);
    S.first = (int *) (base + off[S_FIRST]);
    S.rflag = (unsigned char *) (base + off[S_RFLAG]);
    S.eflag = (unsigned char *) (base + off[S_EFLAG]);
    return S;
}

// LDS: x[n] doubles, perm_r / pivrow / touched [n] ints, Lp [n+1] ints, flag [n] bytes
inline size_t lds_bytes(int n) {
    size_t b = 8ull * n + 4ull * n * 4 + 4ull * (n + 1);
    b = (b + 15) & ~(size_t) 15;
    b += (size_t) n;
    return (b + 255) & ~(size_t) 255;
}

__device__ inline LUWork make_work(char *smem, int n) {
    LUWork w;
    NUPPN_LDS char *p = (NUPPN_LDS char *) smem;
    w.x = (NUPPN_LDS double *) p;
    p += 8ull * n;
    w.perm_r = (NUPPN_LDS int *) p;
    p += 4ull * n;
    w.pivrow = (NUPPN_LDS int *) p;
    p += 4ull * n;
    w.touched = (NUPPN_LDS int *) p;
    p += 4ull * n;
    w.ulist = (NUPPN_LDS int *) p;
    p += 4ull * n;
    w.Lp = (NUPPN_LDS int *) p;
    p += 4ull * (n + 1);
    p = (NUPPN_LDS char *) (((uintptr_t) p + 15) & ~(uintptr_t) 15);
    w.flag = (NUPPN_LDS unsigned char *) p;
    return w;
}

// linear interpolant over [0, dtwant] (solver.F90:84-91, interp.F90:29-50)
__device__ inline double interp2(double x1, double y0, double y1, double xw) {
    double x = fmax(xw, 0.0);
    x = fmin(x, x1);
    return y0 + (y1 - y0) * (x - 0.0) / (x1 - 0.0);
}


// -------------------------------------------------------------------------------------------
// Sub-step controller (solver.F90:53-238) and the backward-Euler decisions
// (backward_euler.F90:29-109) as a per-zone state machine.  One wavefront per zone.

__device__ inline double *zrow(double *base, size_t stride, int z) { return base + (size_t) z * stride; }

// set-up of a zone at the start of a batch: yps_old = yps, entry Ye, phase SUBSTEP
// sb: LDS of pow2(nsp) doubles for the sorted sums
__device__ inline void zone_init(const KArgs &a, int z, NUPPN_LDS double *sb) {
    const DevNet &N = a.N;
    const ZoneData &Z = a.Z;
    double *yps = zrow(Z.yps, Z.nsp_stride, z);
    copy_vec(zrow(Z.yps_old, Z.nsp_stride, z), yps, N.nsp);
    double *work = zrow(Z.work, Z.work_stride, z);
    const double ye = calculate_ye(N, yps, work, sb);
    if (lane_id() == 0) {
        ZoneState s = {};
        s.phase = PH_SUBSTEP;
        s.pad0 = Z.init_prec;   // factor precision of the batched path
        s.i = 0;
        const double t9_0 = Z.t9_0[z], t9_1 = Z.t9_1[z], rho_0 = Z.rho_0[z], rho_1 = Z.rho_1[z];
        s.dttry = Z.dtwant[z];
        s.my_t9_last = 0.5 * (t9_0 + t9_1);
        s.my_rho_last = 0.5 * (rho_0 + rho_1);
        s.ye_entry = ye;
        s.ye_cur = ye;
        Z.state[z] = s;
        Z.bd[z] = Z.bd0;
    }
    wave_sync();
}

// final bookkeeping of a zone (solver.F90:229-237): clip, Ye, outputs.  Under backward Euler
// the Ye of the last step's result equals the Ye of the final abundances (ypsplus leaves every
// value >= 1e-99, so the clips change nothing); under Bader-Deuflhard the renormalisation and the
// clip come after the last Ye (solver.F90:234-237), which the zone state carries.
__device__ inline void zone_finish(const KArgs &a, int z, ZoneState &s, int ierr, bool loop_done,
                                   NUPPN_LDS double *sb) {
    const DevNet &N = a.N;
    const ZoneData &Z = a.Z;
    double *yps = zrow(Z.yps, Z.nsp_stride, z);
    if (loop_done) {
        // DEUFL renormalises by the plain sum (solver.F90:229-235), then the final clip
        const double sm = a.K.integration_method == 1 ? seq_sum_staged(yps, N.nsp, sb) : 1.0;
        for (int j = lane_id(); j < N.nsp; j += WAVE) {
            const double y = a.K.integration_method == 1 ? yps[j] / sm : yps[j];
            yps[j] = fmax(y, 1.0e-99);
        }
        wave_sync();
    }
    // a zone that solved under backward Euler gets the Ye of its final abundances from
    // zone_final_ye after the last pass (its sorted sum would hold up the whole pass here)
    const double ye = a.K.integration_method == 1 ? s.ye_cur : s.ye_entry;
    s.phase = PH_DONE;
    s.ierr = ierr;
    if (!a.K.dynamic) s.st.nvar1 = N.n;
    if (lane_id() == 0) {
        Z.ye[z] = ye;
        Z.ierr[z] = s.lib_err ? s.lib_err : ierr;
        if (Z.stats) Z.stats[z] = s.st;
    }
}

// Ye of zone z once every pass has ended: the final abundances' Ye for a zone that solved under
// backward Euler (the Ye of the last step's result, solver.F90:272, equals it: ypsplus leaves
// every value >= 1e-99, so the final clip changes nothing); zone_finish wrote the entry Ye,
// which stands for every other zone
__device__ inline void zone_final_ye(const KArgs &a, int z, NUPPN_LDS double *sb) {
    const ZoneState s = a.Z.state[z];
    if (a.K.integration_method == 1 || s.phase != PH_DONE || !s.solved) return;
    const double ye = calculate_ye(a.N, zrow(a.Z.yps, a.Z.nsp_stride, z),
                                   zrow(a.Z.work, a.Z.work_stride, z), sb);
    if (lane_id() == 0) a.Z.ye[z] = ye;
}

// one pass of the controller for zone z: consume the result of a Newton iteration if there is
// one, then advance sub-steps until the zone needs another Newton iteration or is done.
__device__ inline void zone_control(const KArgs &a, int z, NUPPN_LDS double *sb) {
    const DevNet &N = a.N;
    const DevKnobs &K = a.K;
    const ZoneData &Z = a.Z;
    const int nsp = N.nsp;
    ZoneState s = Z.state[z];
    if (s.phase == PH_DONE || s.phase == PH_NEWTON) return;
    double *yps = zrow(Z.yps, Z.nsp_stride, z);
    double *yps_old = zrow(Z.yps_old, Z.nsp_stride, z);
    double *y_old = zrow(Z.y_old, Z.nsp_stride, z);
    double *y_cur = zrow(Z.y_cur, Z.nsp_stride, z);
    double *y_new = zrow(Z.y_new, Z.nsp_stride, z);
    double *v = zrow(Z.v, Z.nre_stride, z);
    const double t9_0 = Z.t9_0[z], t9_1 = Z.t9_1[z], rho_0 = Z.rho_0[z], rho_1 = Z.rho_1[z];
    const double dtwant = Z.dtwant[z];
    const double rmin = 0.2, rmax = 2.0, dgrd = K.dgrd;
    const double sum_check = K.grdthreshold * (nsp - 1) * dgrd;
    int ierr = 0;
    bool step_ended = false;   // a backward-Euler step just returned (ierr set)

    if (s.phase == PH_AFTER_NEWTON && K.integration_method == 1) {
        // a Bader-Deuflhard step returned (do_one_time_step, solver.F90:164-179): Ye of the
        // step's result before any clip (solver.F90:272); kept on success and on the returns
        // that follow at once (BAD_STEP, SMALL_DT), replaced by ye_old on a failed step
        const double ye_step = calculate_ye(N, yps, zrow(Z.work, Z.work_stride, z), sb);
        if (s.lib_err || s.step_ierr == E_BAD_STEP) {
            s.ye_cur = ye_step;
            zone_finish(a, z, s, E_BAD_STEP, false, sb);
            if (lane_id() == 0) Z.state[z] = s;
            return;
        }
        ierr = s.step_ierr;
        if (ierr == 0 || s.dttry < SMALL_C) s.ye_cur = ye_step;
        step_ended = true;
    } else if (s.phase == PH_AFTER_NEWTON) {
        if (s.lib_err || s.step_ierr == E_BAD_STEP) {
            // the reference stops here (a zero in the matrix, a singular factor)
            zone_finish(a, z, s, 3, false, sb);
            if (lane_id() == 0) Z.state[z] = s;
            return;
        }
        s.iter += 1;
        s.st.titers++;
        s.st.iter++;
        s.st.n_newton++;
        // y_new is read from HBM once, into the LDS sort buffer: the negative-abundance check
        // on the way, then the mass sum and the copy below from LDS
        double mloc = -__builtin_huge_val();
        for (int i = lane_id(); i < nsp; i += WAVE) {
            const double y = y_new[i];
            sb[i] = y;
            mloc = fmax(mloc, -y);
        }
        const double mneg = wave_max(mloc);
        wave_sync();
        const double sm = seq_sum(sb, nsp);
        if (s.grd <= dgrd && mneg < sum_check && fabs(1.0 - sm) <= SUM_TOL) {
            // yps = y_new, then ypsplus
            for (int i = lane_id(); i < nsp; i += WAVE) yps[i] = ypsplus_value(sb[i]);
            wave_sync();
            s.dtrcmd = 2.0 * fmax(fmin(sqrt(dgrd / fmax(s.grd, SMALL_C)), rmax), rmin) * s.dttry;
            ierr = 0;
            step_ended = true;
        } else {
            for (int i = lane_id(); i < nsp; i += WAVE) y_cur[i] = sb[i];
            wave_sync();
            if (s.iter > K.ittd) {
                double err, tol;
                if (mneg >= sum_check) { err = mneg; tol = sum_check; ierr = 4; }
                else if (fabs(1.0 - sm) > SUM_TOL) { err = fabs(1.0 - sm); tol = SUM_TOL; ierr = 6; }
                else { err = s.grd; tol = dgrd; ierr = 1; }
                s.dtrcmd = fmax(fmin(sqrt(tol / fmax(err, SMALL_C)), rmax), rmin) * s.dttry;
                step_ended = true;
            } else {
                s.phase = PH_NEWTON;
                if (lane_id() == 0) { Z.state[z] = s; Z.flag[0] = 1; Z.active[atomicAdd(&Z.flag[1], 1)] = z; }
                return;
            }
        }
    }

    for (;;) {
        if (step_ended) {
            step_ended = false;
            // back in integrate_network after do_one_time_step (solver.F90:165-225)
            if (s.dttry < SMALL_C) { zone_finish(a, z, s, 11, false, sb); break; }
            if (ierr == 0) {
                s.dtdone = s.dtdone + s.dttry;
                // yps = max(yps, 1e-99), then yps_old = yps (one read of yps)
                batched_for<4>(
                    lane_id(), nsp, WAVE, [&](int j) { return fmax(yps[j], 1.0e-99); },
                    [&](int j, double y) {
                        yps[j] = y;
                        yps_old[j] = y;
                    });
                wave_sync();
                s.st.nsucc++;
                if (s.dtdone >= dtwant) { s.st.nsubt = s.i; zone_finish(a, z, s, 0, true, sb); break; }
            } else {
                copy_vec(yps, yps_old, nsp);
                s.st.nfail++;
            }
            s.dttry = fmin(s.dtrcmd, dtwant - s.dtdone);
            // MAX_SUBSTEPS and SMALL_DT return before nsubt is assigned (solver.F90:174-178,
            // 218-225): the per-call count stays 0
            if (s.i == K.max_sub_steps) { zone_finish(a, z, s, 2, false, sb); break; }
        }
        // next sub-step (solver.F90:108-164)
        s.i += 1;
        if (s.i > 1 && lane_id() == 0) Z.bd[z].frst = 0;   // solver.F90:109
        if (s.i > K.max_sub_steps) {   // loop ran out through the halving branch
            s.st.nsubt = s.i;
            zone_finish(a, z, s, ierr, true, sb);
            break;
        }
        double my_t9, my_rho, my_t9_0 = 0.0, my_t9_1 = 0.0, my_rho_0 = 0.0, my_rho_1 = 0.0;
        if (K.interpolate_temp) {
            const double t_0 = s.dtdone, t_1 = fmin(s.dtdone + s.dttry, dtwant);
            my_t9_0 = interp2(dtwant, t9_0, t9_1, t_0);
            my_t9_1 = interp2(dtwant, t9_0, t9_1, t_1);
            my_t9 = 0.5 * (my_t9_0 + my_t9_1);
        } else {
            my_t9 = 0.5 * (t9_0 + t9_1);
        }
        if (K.interpolate_dens) {
            const double t_0 = s.dtdone, t_1 = fmin(s.dtdone + s.dttry, dtwant);
            my_rho_0 = interp2(dtwant, rho_0, rho_1, t_0);
            my_rho_1 = interp2(dtwant, rho_0, rho_1, t_1);
            my_rho = 0.5 * (my_rho_0 + my_rho_1);
        } else {
            my_rho = 0.5 * (rho_0 + rho_1);
        }
        double diff = fabs(my_t9_0 - my_t9_1) / my_t9_0;
        diff = fmax(diff, fabs(my_rho_0 - my_rho_1) / my_rho_0);
        if (diff > K.delta_state_max && t9_0 > K.min_t9_for_slv && t9_1 > K.min_t9_for_slv) {
            s.dttry = 0.5 * s.dttry;
            continue;
        }
        if (my_t9 != s.my_t9_last || my_rho != s.my_rho_last) {
            zone_rates(N, my_t9, my_rho, v, yps, sb);
            s.my_rho_last = my_rho;
            s.my_t9_last = my_t9;
        }
        ierr = 0;
        if (my_t9 >= K.min_t9_for_slv && K.integration_method == 1) {
            s.solved = 1;
            s.phase = PH_NEWTON;   // the worker runs bader_deuflhard_integration
            if (lane_id() == 0) { Z.state[z] = s; Z.flag[0] = 1; Z.active[atomicAdd(&Z.flag[1], 1)] = z; }
            return;
        }
        if (my_t9 >= K.min_t9_for_slv) {
            // backward_euler_newton set-up (backward_euler.F90:49-53)
            batched_for<4>(
                lane_id(), nsp, WAVE, [&](int j) { return yps[j]; },
                [&](int j, double y) {
                    y_old[j] = y;
                    y_cur[j] = y;
                    y_new[j] = y;
                });
            wave_sync();
            s.iter = 0;
            s.grd = 0.0;
            s.st.titers = 0;
            s.st.iter = 0;
            s.solved = 1;
            if (0 > K.ittd) {
                const double mneg = wave_maxneg(y_new, nsp);
                const double sm = seq_sum_staged(y_new, nsp, sb);
                double err, tol;
                if (mneg >= sum_check) { err = mneg; tol = sum_check; ierr = 4; }
                else if (fabs(1.0 - sm) > SUM_TOL) { err = fabs(1.0 - sm); tol = SUM_TOL; ierr = 6; }
                else { err = s.grd; tol = dgrd; ierr = 1; }
                s.dtrcmd = fmax(fmin(sqrt(tol / fmax(err, SMALL_C)), rmax), rmin) * s.dttry;
                step_ended = true;
                continue;
            }
            s.phase = PH_NEWTON;
            if (lane_id() == 0) { Z.state[z] = s; Z.flag[0] = 1; Z.active[atomicAdd(&Z.flag[1], 1)] = z; }
            return;
        }
        // below min_t9_for_slv: no nucleosynthesis, the step counts as done (solver.F90:180-183)
        step_ended = true;
        ierr = 0;
    }
    if (lane_id() == 0) Z.state[z] = s;
}

// One Newton iteration (newton_iteration, backward_euler.F90:120-176) for zone z, using the
// per-wave scratch S and LDS work w.  Writes y_new and grd (or lib_err) and sets AFTER_NEWTON.
__device__ inline void zone_newton(const KArgs &a, const Slot &S, LUWork w, int z) {
    const DevNet &N = a.N;
    const DevKnobs &K = a.K;
    const ZoneData &Z = a.Z;
    ZoneState s = Z.state[z];
    if (s.phase != PH_NEWTON) return;
    const double *y_old = zrow(Z.y_old, Z.nsp_stride, z);
    const double *y_cur = zrow(Z.y_cur, Z.nsp_stride, z);
    double *y_new = zrow(Z.y_new, Z.nsp_stride, z);
    const double *v = zrow(Z.v, Z.nre_stride, z);
    const double dt = s.dttry;
    molar(N, y_cur, S.xn);
    wave_sync();
    reaction_terms(N, S.xn, v, S.vx, S.st1, S.st2);
    wave_sync();
    const DynWork DW{S.first, S.rflag, S.eflag};
    int stop;
    if (K.dynamic) {
        s.st.nvar1 = dyn_network<Team<1>>(N, K, dt, S.xn, v, S.st1, S.st2, N.dyn_q, N.nnzA, DW,
                                          nullptr);
        stop = assemble_matrix_dyn(N, dt, S.st1, S.st2, DW, S.Aval);
    } else {
        stop = assemble_matrix(N, dt, S.st1, S.st2, S.Aval);
    }
    int e = stop ? 1 : 0;
    for (int r = lane_id(); r < N.n; r += WAVE) {
        const int sp = N.red2sp[r];
        const double d = dxdt_species(N, r, S.vx);
        // reduce_rhs packs the masked species only: the others solve to an exact zero
        S.rhs[r] = (!K.dynamic || dyn_marked(DW, sp)) ? y_old[sp] - y_cur[sp] + dt * d : 0.0;
    }
    wave_sync();
    if (!e) {
        const int info = gepp_factor(N.n, N.A_colptr, N.A_rowind, S.Aval, N.iperm_c, w, S.lu);
        s.st.n_factor++;
        if (info) e = -4;
        else s.st.lu_nnz += (long long) w.Lp[N.n] + (long long) S.lu.Up[N.n];
    }
    if (!e) {
        gepp_solve(N.n, N.perm_c, w, S.lu, S.rhs);
        s.st.n_solve++;
        for (int r = lane_id(); r < N.n; r += WAVE) {
            const int sp = N.red2sp[r];
            y_new[sp] = y_cur[sp] + S.rhs[r];
        }
        wave_sync();
        double g = 0.0;
        for (int i = lane_id(); i < N.nsp; i += WAVE) {
            const double yc = y_cur[i];
            const double d = y_new[i] - yc;
            const double ga = (yc > K.grdthreshold) ? fabs(d / yc) : 0.0;
            g = fmax(g, ga);
        }
        s.grd = wave_max(g);
    }
    // a zero where the reference stops (calculate_jacobian / reduce_jacobian) ends the zone
    // with BAD_STEP; a singular factorisation is the library error NUPPN_ERR_SINGULAR
    s.lib_err = stop ? 0 : e;
    s.step_ierr = stop ? E_BAD_STEP : 0;
    s.phase = PH_AFTER_NEWTON;
    if (lane_id() == 0) Z.state[z] = s;
    wave_sync();
}

}  // namespace nuppn
