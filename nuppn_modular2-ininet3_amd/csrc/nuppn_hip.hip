// MI355X (gfx950) network-integration path: kernels, host set-up and the C ABI of
// include/nuppn_hip.h.
//
// Kernels
//   rates_kernel      REACLIB rates for (reactions x zones): a coalesced structure-of-arrays
//                     stream, one lane per (reaction, zone), zones along gridDim.y.
//   zone_init/zone_control/zone_newton kernels: integrate_network for a batch of zones as a
//                     per-zone state machine (one wavefront per zone): the control kernel runs the
//                     sub-step and Newton decisions, the Newton kernel one Newton iteration
//                     (Jacobian, rhs, GEPP, solve) for every zone that needs one; the host loops
//                     until no zone needs another iteration.
//   lu_kernel         factor + solve of given reduced-pattern matrices (test entry and the
//                     SuperLU drop-ins).
//   dxdt_kernel / jac_kernel  single-state assembly (test entries).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <climits>
#include <mutex>
#include <vector>

#include "../../include/nuppn_hip.h"
#include "device_common.hpp"
#include "integrate.hpp"
#include "zone.hpp"
#include "batched.hpp"
#include "mppnp.hpp"
#include "symbolic.hpp"

using namespace nuppn;

#define HIP_OK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "nuppn_hip: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                  \
            return NUPPN_ERR_HIP;                                                         \
        }                                                                                 \
    } while (0)

// ------------------------------------------------------------------------------------------
// kernels

// reactions per thread of the rate kernels: the zone's T9 powers (a pow, a log, two divisions)
// or cache position are computed once per thread and serve RATE_PER_THREAD rates
constexpr int RATE_PER_THREAD = 8;

__global__ void __launch_bounds__(256) rates_kernel(DevNet N, int nz,
                                                    const double *__restrict__ t9_0,
                                                    const double *__restrict__ t9_1,
                                                    const double *__restrict__ rho_0,
                                                    const double *__restrict__ rho_1,
                                                    double *__restrict__ v) {
    const int z = blockIdx.y;
    if (z >= nz) return;
    const int nre = N.nre;
    // evaluate_all_rates at the cycle mid-point (ppn.F90:287-298, mppnp.F90:719)
    const double t9 = 0.5 * (t9_0[z] + t9_1[z]);
    const double rho = 0.5 * (rho_0[z] + rho_1[z]);
    double *vz = v + (size_t) z * nre;
    const int l0 = blockIdx.x * blockDim.x * RATE_PER_THREAD + threadIdx.x;
    if (N.rcache) {
        const CachePos c = cache_pos(N, t9);
#pragma unroll
        for (int r = 0; r < RATE_PER_THREAD; ++r) {
            const int l = l0 + r * blockDim.x;
            if (l < nre) vz[l] = cached_rate(N, l, N.rtype[l], c, rho);
        }
    } else if (N.P.pf_n > 0) {
        const T9Pow p = t9_powers(t9);
        const PfPos q = pf_pos(N.P, t9);
        for (int r = 0; r < RATE_PER_THREAD; ++r) {
            const int l = l0 + r * blockDim.x;
            if (l >= nre) break;
            const bool rev = N.P.rev_ch[l] > 0;
            vz[l] = reaclib_rate(N.coef, nre, l, N.rtype[l], p, rho,
                                 rev ? reverse_gfac(N.P, nre, l, q) : 1.0, rev);
        }
    } else {
        const T9Pow p = t9_powers(t9);
#pragma unroll
        for (int r = 0; r < RATE_PER_THREAD; ++r) {
            const int l = l0 + r * blockDim.x;
            if (l < nre) vz[l] = reaclib_rate(N.coef, nre, l, N.rtype[l], p, rho);
        }
    }
}

// evaluate_all_rates per zone with the rate physics on (partition functions, detailed balance,
// screening at the zone's abundances yps): one wavefront per zone, LDS >= nsp doubles
__global__ void __launch_bounds__(64) rates_phys_kernel(DevNet N, int nz,
                                                        const double *__restrict__ t9_0,
                                                        const double *__restrict__ t9_1,
                                                        const double *__restrict__ rho_0,
                                                        const double *__restrict__ rho_1,
                                                        const double *__restrict__ yps,
                                                        size_t nsp_stride, double *v) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS double *buf = (NUPPN_LDS double *) smem;
    for (int z = blockIdx.x; z < nz; z += gridDim.x) {
        const double t9 = 0.5 * (t9_0[z] + t9_1[z]);
        const double rho = 0.5 * (rho_0[z] + rho_1[z]);
        zone_rates(N, t9, rho, v + (size_t) z * N.nre, yps + (size_t) z * nsp_stride, buf);
    }
}

// reaclib_preprocessor (physics/source/reaclib.F90:609-646): the rates at rho = 1 on the T9
// grid, each floored at 1e-99 (the formula's own floors)
__global__ void __launch_bounds__(256) rate_cache_kernel(DevNet N, const double *t9grid,
                                                         double *cache) {
    const int i = blockIdx.y;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNT9 || l >= N.nre) return;
    const T9Pow p = t9_powers(t9grid[i]);
    // the preprocessor interpolates the partition functions at every grid point
    // (reaclib.F90:627): the reverse rates are stored with their factors
    const bool rev = N.P.pf_n > 0 && N.P.rev_ch[l] > 0;
    const double gf = rev ? reverse_gfac(N.P, N.nre, l, pf_pos(N.P, t9grid[i])) : 1.0;
    cache[(size_t) i * N.nre + l] = reaclib_rate(N.coef, N.nre, l, N.rtype[l], p, 1.0, gf, rev);
}

static dim3 rates_grid(int nre, int nz) {
    return dim3((nre + 256 * RATE_PER_THREAD - 1) / (256 * RATE_PER_THREAD), nz);
}



// dynamic LDS of both: pow2(nsp) doubles, the sort buffer of the sorted sums
__global__ void __launch_bounds__(64) zone_init_kernel(KArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    zone_init(a, blockIdx.x, (NUPPN_LDS double *) smem);
}

// the final Ye of every zone of a batch (after the last pass)
__global__ void __launch_bounds__(64) zone_ye_kernel(KArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    zone_final_ye(a, blockIdx.x, (NUPPN_LDS double *) smem);
}

// one controller pass over the zones of the previous pass's work list (every zone at first: the
// zones not on it are done)
__global__ void __launch_bounds__(64) zone_control_kernel(KArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int nctl = list_len(a.Z.nctl_dev, a.Z.nctl);
    for (int i = blockIdx.x; i < nctl; i += gridDim.x)
        zone_control(a, a.Z.ctl ? a.Z.ctl[i] : i, (NUPPN_LDS double *) smem);
}

// one Newton iteration (or Bader-Deuflhard step) for every zone in PH_NEWTON; each wavefront
// owns one scratch slot and strides over the zones
__global__ void __launch_bounds__(64) zone_newton_kernel(KArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const LUWork w = make_work(smem, a.N.n);
    const Slot S = make_slot(a.slot_base + (size_t) blockIdx.x * a.slot_bytes, a.off);
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x) zone_newton(a, S, w, a.Z.active[i]);
}

// one Bader-Deuflhard step for every zone in PH_NEWTON (integration_method 1)
__global__ void __launch_bounds__(64) zone_bd_kernel(KArgs a) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const LUWork w = make_work(smem, a.N.n);
    const Slot S = make_slot(a.slot_base + (size_t) blockIdx.x * a.slot_bytes, a.off);
    const int nact = list_len(a.Z.nact_dev, a.Z.nact);
    for (int i = blockIdx.x; i < nact; i += gridDim.x) zone_bd(a, S, w, a.Z.active[i]);
}

__global__ void flag_reset_kernel(int *flag) {
    flag[0] = 0;
    flag[1] = 0;
}

namespace nuppn {
template __global__ void zone_newton_block_kernel<32, 1>(KArgs a, BKArgs b);
template __global__ void zone_newton_block_kernel<32, 2>(KArgs a, BKArgs b);
template __global__ void zone_newton_block_kernel<64, 1>(KArgs a, BKArgs b);
template __global__ void zone_newton_block_kernel<64, 2>(KArgs a, BKArgs b);
template __global__ void zone_newton_team_kernel<32, 2>(KArgs a, BKArgs b);
template __global__ void zone_newton_team_kernel<32, 4>(KArgs a, BKArgs b);
template __global__ void zone_newton_team_kernel<64, 2>(KArgs a, BKArgs b);
template __global__ void zone_newton_team_kernel<64, 4>(KArgs a, BKArgs b);
template __global__ void zone_bd_block_kernel<32>(KArgs a, BKArgs b);
template __global__ void zone_bd_block_kernel<64>(KArgs a, BKArgs b);
}  // namespace nuppn

// single-state assembly entries (test hooks): mode 0 dX/dt, mode 1 reduced matrix values
__global__ void __launch_bounds__(64) assemble_kernel(KArgs a, int mode, const double *x,
                                                     const double *v, double dt, double *out,
                                                     int *status) {
    const DevNet &N = a.N;
    const Slot S = make_slot(a.slot_base, a.off);
    molar(N, x, S.xn);
    wave_sync();
    reaction_terms(N, S.xn, v, S.vx, S.st1, S.st2);
    wave_sync();
    if (mode == 0) {
        for (int i = lane_id(); i < N.nsp; i += WAVE) out[i] = 0.0;
        wave_sync();
        for (int r = lane_id(); r < N.n; r += WAVE) out[N.red2sp[r]] = dxdt_species(N, r, S.vx);
    } else {
        const int e = assemble_matrix(N, dt, S.st1, S.st2, out);
        if (lane_id() == 0) status[0] = e;
    }
}

// factor + solve given matrices of a fixed pattern (test hook and SuperLU drop-ins).
// Acp/Ari/perm_c/iperm_c: pattern; vals: [nrhs_sets? no: one matrix]; b: [nrhs*n] in/out
__global__ void __launch_bounds__(64) lu_kernel(int n, const int *Acp, const int *Ari,
                                               const double *vals, const int *perm_c,
                                               const int *iperm_c, int nrhs, double *b,
                                               LUStore st, int *perm_r_out, int *Lp_out,
                                               int *info_out, int factor) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    LUWork w = make_work(smem, n);
    int info = 0;
    if (factor) {
        info = gepp_factor(n, Acp, Ari, vals, iperm_c, w, st);
        for (int i = lane_id(); i < n; i += WAVE) perm_r_out[i] = w.perm_r[i];
        for (int i = lane_id(); i <= n; i += WAVE) Lp_out[i] = w.Lp[i];
        wave_sync();
    } else {
        for (int i = lane_id(); i < n; i += WAVE) w.perm_r[i] = perm_r_out[i];
        for (int i = lane_id(); i <= n; i += WAVE) w.Lp[i] = Lp_out[i];
        wave_sync();
    }
    if (info == 0)
        for (int r = 0; r < nrhs; ++r) gepp_solve(n, perm_c, w, st, b + (size_t) r * n);
    if (lane_id() == 0) info_out[0] = info;
}

// ------------------------------------------------------------------------------------------
// host side

namespace {

template <class T>
int dev_upload(T **dptr, const T *src, size_t count) {
    if (count == 0) count = 1;
    HIP_OK(hipMalloc((void **) dptr, sizeof(T) * count));
    if (src) HIP_OK(hipMemcpy(*dptr, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return 0;
}

template <class T>
int dev_upload(T **dptr, const std::vector<T> &v) {
    return dev_upload(dptr, v.data(), v.size());
}

void dev_free(void *p) {
    if (p) (void) hipFree(p);
}

// passes queued ahead of the host: each has its own flag pair on the device and event set
constexpr int kPassRing = 16;

// a chunk of kPassRing passes built as a graph (and its executable), replayed while the work list
// stays within the bound its grids were sized for; freed at the end of a batch call
struct PassGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int bound = -1;
    // the kernel arguments of the graph's passes, kept for the graph's lifetime
    KArgs ka[kPassRing];
    int *flag[kPassRing];
    void reset() {
        if (exec) (void) hipGraphExecDestroy(exec);
        if (graph) (void) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        bound = -1;
    }
    ~PassGraph() { reset(); }
};

}  // namespace

struct nuppn_ctx {
    HostNetwork net;
    Symbolic sym;
    nuppn_knobs kn;
    std::vector<int> iperm_c;
    // device copies of the network and symbolic data
    int *d_k = nullptr, *d_rtype = nullptr, *d_red2sp = nullptr, *d_dx_ptr = nullptr;
    int *d_dx_ent = nullptr, *d_jq_ptr = nullptr, *d_jq_ent = nullptr, *d_Acp = nullptr;
    int *d_Ari = nullptr, *d_perm_c = nullptr, *d_iperm_c = nullptr, *d_dyn_q = nullptr;
    unsigned char *d_isdiag = nullptr;
    double *d_coef = nullptr, *d_aw = nullptr, *d_zw = nullptr;
    // persistent-kernel slots
    SlotLayout layout{};
    size_t lds = 0;
    int max_slots = 0, nslots = 0;
    char *d_slots = nullptr;
    int *d_flag = nullptr;
    int *h_flag = nullptr;   // pinned
    // per-zone state for batches
    int nz_state_cap = 0;
    double *d_yps_old = nullptr, *d_y_old = nullptr, *d_y_cur = nullptr, *d_y_new = nullptr;
    double *d_work = nullptr;
    ZoneState *d_state = nullptr;
    BDState *d_bd = nullptr;      // [nz] per-zone Bader-Deuflhard state of the running call
    int *d_active = nullptr;      // [2][nz] the work lists of consecutive passes
    size_t sort_lds = 0;          // LDS of the control and init kernels (sorted-sum buffer)
    BDTables bdt{};               // bader_deuflhard_init's tables
    BDState bd{};                 // the reference's SAVE'd BD state, carried between calls
    size_t work_stride = 0;
    // staging for host-array calls
    int nz_cap = 0, nz_vcap = 0;
    double *d_yps = nullptr, *d_t9_0 = nullptr, *d_t9_1 = nullptr, *d_rho_0 = nullptr;
    double *d_rho_1 = nullptr, *d_dt = nullptr, *d_ye = nullptr, *d_v = nullptr;
    int *d_ierr = nullptr;
    ZoneStats *d_stats = nullptr;
    int *d_status = nullptr;
    // batched path (block-tridiagonal pivot-free factorisation)
    BlockSym bsym;
    int bnb = 0;   // block size, 0 when the network's interior is not block tridiagonal
    int *d_b_vpos = nullptr, *d_b_diagk = nullptr, *d_b_ent_ptr = nullptr, *d_b_ent_k = nullptr;
    int *d_b_ent_loc = nullptr, *d_b_ent_col = nullptr, *d_b_csr_col = nullptr, *d_b_ent_dst = nullptr;
    int *d_b_csr_row = nullptr, *d_b_dyn_k = nullptr;
    // the band form's entering lists (band_lu.hpp)
    int *d_bd_ent_ptr = nullptr, *d_bd_ent_k = nullptr, *d_bd_ent_dst = nullptr;
    int *d_bd_ent_col = nullptr;
    double *d_bd_ent_ys = nullptr;
    unsigned char *d_b_isdiag = nullptr;
    double *d_b_mass = nullptr, *d_b_ent_ys = nullptr;
    int *d_seg_i[12] = {};        // jac, dxdt, resid: e_idx, lane_len, lane_s0, round_base
    int *d_b_resid_pos = nullptr; // A entry -> physical position in the residual stream
    double *d_seg_c[3] = {};      // dxdt, resid: e_coef; [0]: J's per-entry ratio, high parts
    double *d_jac_lo = nullptr;   // J's per-entry ratio, low parts
    double *d_rcache = nullptr, *d_rc_t9 = nullptr;   // REACLIB T9 cache, grid
    // rate physics (nuppn_set_rate_physics)
    int pf_n = 0, ndb = 0, screen = 0;
    double *d_pf_t9 = nullptr, *d_pf_tab = nullptr, *d_g_spin = nullptr, *d_mass_g = nullptr;
    double *d_q_mev = nullptr;
    int *d_rev_ch = nullptr, *d_rev_part = nullptr, *d_db_fwd = nullptr, *d_db_rev = nullptr;
    double rc_dt9 = 0.0;
    int use_cache = 0;
    BLayout blayout{};
    size_t blds32 = 0, blds64 = 0;   // LDS per wavefront with FP32 / FP64 factors
    size_t blds_team = 0;            // LDS of a multi-wavefront kernel's workgroup
    int team_ok2 = 0;                // the two-per-CU four-wavefront variant fits
    int team_ok[5] = {};             // [2], [4]: that multi-wavefront kernel fits the device
    int ncu = 0;                     // compute units of the device
    int bmax_slots = 0, bslots = 0;
    int bgrid[3] = {0, 0, 0};   // resident wavefronts of the worker held to 1 / 2 per SIMD
    char *d_bslots = nullptr;
    hipEvent_t ev[4] = {};
    // per pass of a chunk (pass % kPassRing): control start / end, worker end
    hipEvent_t evc[kPassRing][2] = {}, evn[kPassRing][1] = {};
    // the same for captured pass graphs (events a graph records are used by graphs only)
    hipEvent_t gevc[kPassRing][2] = {}, gevn[kPassRing][1] = {};
    double newton_ms = 0.0, control_ms = 0.0;
    long long newton_launches = 0, graph_launches = 0;
    hipGraphNode_t glast = nullptr;     // the last node of a pass graph being built
    // the autotuner's choice, the defaults of the batched path once it has run
    int tuned = 0;
    nuppn_batch_opts tuned_opts{};
    // the mppnp cycle (zone selection, batch gather / scatter, mixing): device buffers for m zones
    int mp_cap = 0;
    int *mp_zidx = nullptr, *mp_ierrb = nullptr, *mp_small = nullptr;   // small: count, bad
    double *mp_yb = nullptr, *mp_t9b = nullptr, *mp_rhob = nullptr, *mp_dtb = nullptr;
    double *mp_yeb = nullptr, *mp_dl = nullptr, *mp_a = nullptr, *mp_b = nullptr, *mp_c = nullptr;
    double *mp_bet = nullptr, *mp_gam = nullptr, *mp_res = nullptr, *mp_max = nullptr;
    ZoneStats *mp_stb = nullptr;
    long long last_passes = 0;
    std::mutex mu;
    ~nuppn_ctx();
};

// frees every device and pinned allocation (also on the error returns of nuppn_ctx_create,
// whose unique_ptr deletes a half-built context)
nuppn_ctx::~nuppn_ctx() {
    (void) hipDeviceSynchronize();
    for (void *p : {(void *) d_k, (void *) d_rtype, (void *) d_red2sp, (void *) d_dx_ptr,
                    (void *) d_dx_ent, (void *) d_jq_ptr, (void *) d_jq_ent, (void *) d_Acp,
                    (void *) d_Ari, (void *) d_perm_c, (void *) d_iperm_c, (void *) d_isdiag,
                    (void *) d_coef, (void *) d_aw, (void *) d_zw, (void *) d_slots,
                    (void *) d_flag, (void *) d_yps_old, (void *) d_y_old, (void *) d_y_cur,
                    (void *) d_y_new, (void *) d_work, (void *) d_state, (void *) d_bd,
                    (void *) d_active, (void *) d_yps, (void *) d_t9_0, (void *) d_t9_1,
                    (void *) d_rho_0, (void *) d_rho_1, (void *) d_dt, (void *) d_ye, (void *) d_v,
                    (void *) d_ierr, (void *) d_stats, (void *) d_status, (void *) d_b_vpos,
                    (void *) d_b_diagk, (void *) d_b_ent_ptr, (void *) d_b_ent_k,
                    (void *) d_b_ent_loc, (void *) d_b_ent_col, (void *) d_b_csr_col,
                    (void *) d_b_isdiag, (void *) d_b_mass, (void *) d_b_ent_ys, (void *) d_bslots,
                    (void *) d_b_ent_dst, (void *) d_jac_lo, (void *) d_rcache, (void *) d_rc_t9,
                    (void *) d_b_resid_pos, (void *) d_dyn_q, (void *) d_b_csr_row,
                    (void *) d_b_dyn_k, (void *) d_bd_ent_ptr, (void *) d_bd_ent_k,
                    (void *) d_bd_ent_dst, (void *) d_bd_ent_col, (void *) d_bd_ent_ys})
        if (p) (void) hipFree(p);
    for (void *p : {(void *) d_pf_t9, (void *) d_pf_tab, (void *) d_g_spin, (void *) d_mass_g,
                    (void *) d_q_mev, (void *) d_rev_ch, (void *) d_rev_part, (void *) d_db_fwd,
                    (void *) d_db_rev})
        if (p) (void) hipFree(p);
    for (void *p : {(void *) mp_zidx, (void *) mp_ierrb, (void *) mp_small, (void *) mp_yb,
                    (void *) mp_t9b, (void *) mp_rhob, (void *) mp_dtb, (void *) mp_yeb,
                    (void *) mp_dl, (void *) mp_a, (void *) mp_b, (void *) mp_c, (void *) mp_bet,
                    (void *) mp_gam, (void *) mp_res, (void *) mp_max, (void *) mp_stb})
        if (p) (void) hipFree(p);
    for (void *p : d_seg_i)
        if (p) (void) hipFree(p);
    for (void *p : d_seg_c)
        if (p) (void) hipFree(p);
    for (auto &e : ev)
        if (e) (void) hipEventDestroy(e);
    for (auto &r : evn)
        for (auto &e : r)
            if (e) (void) hipEventDestroy(e);
    for (auto &r : evc)
        for (auto &e : r)
            if (e) (void) hipEventDestroy(e);
    for (auto &r : gevn)
        for (auto &e : r)
            if (e) (void) hipEventDestroy(e);
    for (auto &r : gevc)
        for (auto &e : r)
            if (e) (void) hipEventDestroy(e);
    if (h_flag) (void) hipHostFree(h_flag);
}

static DevNet dev_net(const nuppn_ctx *c) {
    DevNet N;
    N.nsp = c->net.nsp;
    N.nre = c->net.nre;
    N.n = c->sym.n;
    N.nnzA = c->sym.nnzA;
    N.igamma = c->net.igamma;
    N.k = c->d_k;
    N.rtype = c->d_rtype;
    N.coef = c->d_coef;
    N.aw = c->d_aw;
    N.zw = c->d_zw;
    N.red2sp = c->d_red2sp;
    N.dx_ptr = c->d_dx_ptr;
    N.dx_ent = c->d_dx_ent;
    N.jq_ptr = c->d_jq_ptr;
    N.jq_ent = c->d_jq_ent;
    N.A_colptr = c->d_Acp;
    N.A_rowind = c->d_Ari;
    N.A_isdiag = c->d_isdiag;
    N.perm_c = c->d_perm_c;
    N.iperm_c = c->d_iperm_c;
    N.dyn_q = c->d_dyn_q;
    N.rcache = c->use_cache ? c->d_rcache : nullptr;
    N.rc_t9 = c->d_rc_t9;
    N.rc_dt9 = c->rc_dt9;
    DevPhys &P = N.P;
    P.pf_n = c->pf_n;
    P.pf_t9 = c->d_pf_t9;
    P.pf_tab = c->d_pf_tab;
    P.rev_ch = c->d_rev_ch;
    P.rev_part = c->d_rev_part;
    P.ndb = c->ndb;
    P.db_fwd = c->d_db_fwd;
    P.db_rev = c->d_db_rev;
    P.g_spin = c->d_g_spin;
    P.mass_g = c->d_mass_g;
    P.q_mev = c->d_q_mev;
    P.screen = c->screen;
    return N;
}

static bool rate_physics_on(const nuppn_ctx *c) { return c->pf_n > 0 || c->screen; }

// rates of a batch at its cycle mid-points into c->d_v: the REACLIB kernel over (reaction,
// zone) or, with the rate physics on, one wavefront per zone at the zones' abundances
static void launch_rates(nuppn_ctx *c, int nz, const double *t9_0, const double *t9_1,
                         const double *rho_0, const double *rho_1, const double *yps,
                         hipStream_t s) {
    if (rate_physics_on(c))
        hipLaunchKernelGGL(rates_phys_kernel, dim3(std::min(nz, 8192)), dim3(64),
                           (size_t) c->net.nsp * 8, s, dev_net(c), nz, t9_0, t9_1, rho_0, rho_1,
                           yps, (size_t) c->net.nsp, c->d_v);
    else
        hipLaunchKernelGGL(rates_kernel, rates_grid(c->net.nre, nz), dim3(256), 0, s, dev_net(c),
                           nz, t9_0, t9_1, rho_0, rho_1, c->d_v);
}

static DevKnobs dev_knobs(const nuppn_knobs &k) {
    DevKnobs K;
    K.ittd = k.ittd;
    K.max_sub_steps = k.max_sub_steps;
    K.integration_method = k.integration_method;
    K.interpolate_temp = k.interpolate_temp;
    K.interpolate_dens = k.interpolate_dens;
    K.dgrd = k.dgrd;
    K.grdthreshold = k.grdthreshold;
    K.min_t9_for_slv = k.min_t9_for_slv;
    K.delta_state_max = k.delta_state_max;
    K.dynamic = k.cyminformal >= 0.0;
    K.cyminformal = k.cyminformal;
    K.min_rate_ov_dt = k.min_rate_ov_dt;
    return K;
}

static DevBlock dev_block(const nuppn_ctx *c) {
    DevBlock B;
    const BlockSym &b = c->bsym;
    B.NB = b.NB;
    B.nbd = b.nbd;
    B.ni = b.ni;
    B.nblk = b.nblk;
    B.nip = b.nip;
    B.n = c->sym.n;
    B.nnzA = c->sym.nnzA;
    B.vpos = c->d_b_vpos;
    B.mass = c->d_b_mass;
    B.diag_a = c->d_b_diagk;
    B.ent_ptr = c->d_b_ent_ptr;
    B.ent_a = c->d_b_ent_k;
    B.ent_loc = c->d_b_ent_loc;
    B.ent_dst = c->d_b_ent_dst;
    B.ent_col = c->d_b_ent_col;
    B.ent_ys = c->d_b_ent_ys;
    B.csr_col = c->d_b_csr_col;
    B.csr_row = c->d_b_csr_row;
    B.dyn_k = c->d_b_dyn_k;
    B.csr_isdiag = c->d_b_isdiag;
    B.resid_pos = c->d_b_resid_pos;
    const SegStream *sc[3] = {&b.jac, &b.dxdt, &b.resid};
    DevSeg *ds[3] = {&B.jac, &B.dxdt, &B.resid};
    for (int i = 0; i < 3; ++i) {
        ds[i]->nseg = sc[i]->nseg;
        ds[i]->e_idx = c->d_seg_i[4 * i];
        ds[i]->lane_len = c->d_seg_i[4 * i + 1];
        ds[i]->lane_s0 = c->d_seg_i[4 * i + 2];
        ds[i]->round_base = c->d_seg_i[4 * i + 3];
        ds[i]->e_coef = c->d_seg_c[i];
    }
    B.jac.e_coef = nullptr;   // the J stream's counts are packed in its operand words
    B.jac_rhi = c->d_seg_c[0];
    B.jac_rlo = c->d_jac_lo;
    return B;
}

// the band form's structure (ni = 0 for the block form)
static DevBand dev_band(const nuppn_ctx *c) {
    DevBand D{};
    const BlockSym &b = c->bsym;
    if (c->bnb != kBandForm) return D;
    D.ni = b.ni;
    D.nbd = b.nbd;
    D.bw = b.bandwidth;
    D.W = b.band_w;
    D.WT = b.band_w + b.nbd;
    D.ent_ptr = c->d_bd_ent_ptr;
    D.ent_k = c->d_bd_ent_k;
    D.ent_dst = c->d_bd_ent_dst;
    D.ent_col = c->d_bd_ent_col;
    D.ent_ys = c->d_bd_ent_ys;
    return D;
}

// The batched path's Bader-Deuflhard step is switched off: its kernel took a memory aperture
// violation on the GPU after the segmented sums moved to the streamed form, with no cause found
// (the same sums run in the backward-Euler kernel, and a CPU emulation of the BD kernel under
// AddressSanitizer is clean).  The plain path's Bader-Deuflhard step is unaffected.
static bool batched_method_enabled(const nuppn_ctx *c) { return c->kn.integration_method == 0; }

// the batched path's worker: BD step (method 1) or Newton iteration, the latter held to one or
// two resident wavefronts per SIMD (waves)
static const void *block_kernel_ptr(int nb, int method, int waves) {
    if (method == 1)
        return nb == 64 ? (const void *) zone_bd_block_kernel<64>
                        : (const void *) zone_bd_block_kernel<32>;
    if (nb == 64)
        return waves == 2 ? (const void *) zone_newton_block_kernel<64, 2>
                          : (const void *) zone_newton_block_kernel<64, 1>;
    return waves == 2 ? (const void *) zone_newton_block_kernel<32, 2>
                      : (const void *) zone_newton_block_kernel<32, 1>;
}

static const void *team_kernel_ptr(int nb, int nw) {
    if (nb == 64)
        return nw == 4 ? (const void *) zone_newton_team_kernel<64, 4>
                       : (const void *) zone_newton_team_kernel<64, 2>;
    return nw == 4 ? (const void *) zone_newton_team_kernel<32, 4>
                   : (const void *) zone_newton_team_kernel<32, 2>;
}

// The runtime gives every thread a private (scratch) segment of hipLimitStackSize bytes (1,024
// by default).  A kernel whose register spills and private arrays need more would run past it,
// so the limit is raised to the largest private size of the kernels a context launches.
static int ensure_stack(const void *kfn) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, kfn) != hipSuccess) return -1;
    size_t cur = 0;
    if (hipDeviceGetLimit(&cur, hipLimitStackSize) != hipSuccess) return -1;
    if (attr.localSizeBytes > cur &&
        hipDeviceSetLimit(hipLimitStackSize, attr.localSizeBytes) != hipSuccess)
        return -1;
    return 0;
}

// resident wavefronts per CU of a kernel with `lds` bytes of dynamic LDS
static int kernel_occupancy(const void *kfn, size_t lds) {
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) !=
            hipSuccess)
        return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 64, lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return per_cu;
}

// the batched path's device arrays (before a rebuild with another block size)
static void free_batched(nuppn_ctx *c) {
    for (void **p : {(void **) &c->d_b_vpos, (void **) &c->d_b_diagk, (void **) &c->d_b_ent_ptr,
                     (void **) &c->d_b_ent_k, (void **) &c->d_b_ent_loc, (void **) &c->d_b_ent_col,
                     (void **) &c->d_b_csr_col, (void **) &c->d_b_isdiag, (void **) &c->d_b_mass,
                     (void **) &c->d_b_ent_ys, (void **) &c->d_b_ent_dst, (void **) &c->d_jac_lo,
                     (void **) &c->d_b_resid_pos, (void **) &c->d_bslots,
                     (void **) &c->d_b_csr_row, (void **) &c->d_b_dyn_k,
                     (void **) &c->d_bd_ent_ptr, (void **) &c->d_bd_ent_k,
                     (void **) &c->d_bd_ent_dst, (void **) &c->d_bd_ent_col,
                     (void **) &c->d_bd_ent_ys}) {
        dev_free(*p);
        *p = nullptr;
    }
    for (int *&p : c->d_seg_i) {
        dev_free(p);
        p = nullptr;
    }
    for (double *&p : c->d_seg_c) {
        dev_free(p);
        p = nullptr;
    }
    c->bslots = 0;
    c->bnb = 0;
}

// block structure + device upload + occupancy of the batched Newton kernels (once per context,
// again when the autotuner tries another block size): force_nb 0 takes the smallest block size
// whose blocks hold the interior band
static int setup_batched(nuppn_ctx *c, int force_nb = 0) {
    const int nb = force_nb ? force_nb : choose_block_size(c->net, c->sym);
    if (nb == 0) return 0;   // the batched path stays unavailable; the plain path is unaffected
    const bool band = nb == kBandForm;
    c->bsym = BlockSym();
    if (build_block_symbolic(c->net, c->sym, band ? 0 : nb, c->bsym) != 0 || c->bsym.nbd > 8)
        return 0;
    const BlockSym &b = c->bsym;
    if (band && (dev_upload(&c->d_bd_ent_ptr, b.bent_ptr) || dev_upload(&c->d_bd_ent_k, b.bent_k) ||
                 dev_upload(&c->d_bd_ent_dst, b.bent_dst) ||
                 dev_upload(&c->d_bd_ent_col, b.bent_col) ||
                 dev_upload(&c->d_bd_ent_ys, b.bent_ys)))
        return NUPPN_ERR_HIP;
    if (dev_upload(&c->d_b_vpos, b.vpos) || dev_upload(&c->d_b_diagk, b.diagk) ||
        dev_upload(&c->d_b_ent_ptr, b.ent_ptr) || dev_upload(&c->d_b_ent_k, b.ent_k) ||
        dev_upload(&c->d_b_ent_loc, b.ent_loc) || dev_upload(&c->d_b_ent_col, b.ent_col) ||
        dev_upload(&c->d_b_csr_col, b.csr_col) || dev_upload(&c->d_b_isdiag, b.csr_isdiag) ||
        dev_upload(&c->d_b_mass, b.mass) || dev_upload(&c->d_b_ent_ys, b.ent_ys) ||
        dev_upload(&c->d_b_ent_dst, b.ent_dst) || dev_upload(&c->d_b_csr_row, b.csr_row) ||
        dev_upload(&c->d_b_dyn_k, b.dyn_k))
        return NUPPN_ERR_HIP;
    const SegStream *sc[3] = {&b.jac, &b.dxdt, &b.resid};
    for (int i = 0; i < 3; ++i) {
        if (dev_upload(&c->d_seg_i[4 * i], sc[i]->e_idx) ||
            dev_upload(&c->d_seg_i[4 * i + 1], sc[i]->lane_len) ||
            dev_upload(&c->d_seg_i[4 * i + 2], sc[i]->lane_s0) ||
            dev_upload(&c->d_seg_i[4 * i + 3], sc[i]->round_base) ||
            dev_upload(&c->d_seg_c[i], i == 0 ? b.jac_rhi : sc[i]->e_coef))
            return NUPPN_ERR_HIP;
    }
    if (dev_upload(&c->d_jac_lo, b.jac_rlo) || dev_upload(&c->d_b_resid_pos, b.resid_pos))
        return NUPPN_ERR_HIP;
    if (band) {
        // one workgroup of four wavefronts per zone, the factorisation's window in its registers
        static_assert(kBandWindowMax == kBandMaxWT, "band window limits differ");
        static_assert(kBandListGroup == kBandEnterChunk && kBandListCap == kBandEnterCap,
                      "band entering-list groups differ");
        c->blayout = block_layout(c->net.nsp, c->net.nre, c->sym.n, c->sym.nnzA, 1, 0, b.ni,
                                  b.nbd, b.resid.round_base[kSegRounds],
                                  (size_t) b.ni * b.bandwidth, (size_t) b.ni * b.band_w);
        c->blds64 = band_lds_bytes(c->sym.n, b.band_w, b.nbd, 8);
        c->blds32 = band_lds_bytes(c->sym.n, b.band_w, b.nbd, 4);
        if (c->blds64 > 160 * 1024) return 0;
        const void *k = (const void *) zone_newton_band_kernel;
        if (ensure_stack(k)) return NUPPN_ERR_HIP;
        if (c->blds64 > 64 * 1024 &&
            hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int) c->blds64) != hipSuccess)
            return 0;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, c->blds64) !=
                hipSuccess || per_cu <= 0)
            return 0;
        int dev = 0, ncu = 0;
        (void) hipGetDevice(&dev);
        (void) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        size_t free_b = 0, total_b = 0;
        (void) hipMemGetInfo(&free_b, &total_b);
        const int by_mem = (int) std::max<size_t>(1, (free_b / 4) / c->blayout.bytes);
        c->bmax_slots = std::max(1, std::min(ncu * per_cu, by_mem));
        c->bgrid[1] = c->bgrid[2] = c->bmax_slots;
        c->ncu = std::max(1, ncu);
        c->team_ok[2] = c->team_ok[4] = 0;
        c->bnb = kBandForm;
        return 0;
    }
    c->blayout = block_layout(c->net.nsp, c->net.nre, c->sym.n, c->sym.nnzA, nb, b.nblk, b.nip,
                              b.nbd, b.resid.round_base[kSegRounds]);
    auto rnd = [](size_t x) { return (x + 255) & ~(size_t) 255; };
    c->blds32 = rnd(nb == 64 ? block_lds_bytes<float, 64>(b.nip, b.nbd)
                             : block_lds_bytes<float, 32>(b.nip, b.nbd));
    c->blds64 = rnd(nb == 64 ? block_lds_bytes<double, 64>(b.nip, b.nbd)
                             : block_lds_bytes<double, 32>(b.nip, b.nbd));
    if (c->blds64 > 160 * 1024) return 0;
    // one kernel runs both factor precisions: it takes the FP64 LDS size
    const int method = c->kn.integration_method;
    if (ensure_stack(block_kernel_ptr(nb, method, 1)) || ensure_stack(block_kernel_ptr(nb, method, 2)))
        return NUPPN_ERR_HIP;
    const int occ1 = kernel_occupancy(block_kernel_ptr(nb, method, 1), c->blds64);
    const int occ2 = kernel_occupancy(block_kernel_ptr(nb, method, 2), c->blds64);
    if (occ1 <= 0 || occ2 <= 0) return 0;
    const int occ = std::max(occ1, occ2);
    int dev = 0, ncu = 0;
    (void) hipGetDevice(&dev);
    (void) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    size_t free_b = 0, total_b = 0;
    (void) hipMemGetInfo(&free_b, &total_b);
    const int by_mem = (int) std::max<size_t>(1, (free_b / 4) / c->blayout.bytes);
    c->bmax_slots = std::max(1, std::min(ncu * occ, by_mem));
    c->bgrid[1] = std::max(1, std::min(ncu * occ1, c->bmax_slots));
    c->bgrid[2] = std::max(1, std::min(ncu * occ2, c->bmax_slots));
    c->ncu = std::max(1, ncu);
    // the multi-wavefront kernels for passes with few zones: one workgroup of 2 or 4 wavefronts
    // per zone, with a second block-operand area in LDS for the bottom-up sweep
    c->blds_team = rnd(nb == 64 ? block_lds_bytes<double, 64>(b.nip, b.nbd, 2)
                                : block_lds_bytes<double, 32>(b.nip, b.nbd, 2));
    for (int nw = 2; nw <= 4; nw += 2) {
        c->team_ok[nw] = 0;
        if (method != 0 || c->blds_team > 160 * 1024) continue;
        const void *tk = team_kernel_ptr(nb, nw);
        int per_cu = 0;
        if (ensure_stack(tk)) return NUPPN_ERR_HIP;
        if ((c->blds_team <= 64 * 1024 ||
             hipFuncSetAttribute(tk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) c->blds_team) == hipSuccess) &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tk, 64 * nw, c->blds_team) ==
                hipSuccess &&
            per_cu >= 1)
            c->team_ok[nw] = 1;
    }
    c->team_ok2 = 0;
    if (nb == 32 && method == 0 && c->team_ok[4]) {
        const void *tk2 = (const void *) zone_newton_team_kernel<32, 4, 2>;
        int per_cu = 0;
        if (ensure_stack(tk2)) return NUPPN_ERR_HIP;
        if ((c->blds_team <= 64 * 1024 ||
             hipFuncSetAttribute(tk2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) c->blds_team) == hipSuccess) &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tk2, 256, c->blds_team) ==
                hipSuccess &&
            per_cu >= 2)
            c->team_ok2 = 1;
    }
    c->bnb = nb;
    return 0;
}

static int ensure_bslots(nuppn_ctx *c, int want) {
    want = std::max(1, std::min(want, c->bmax_slots));
    if (want <= c->bslots) return 0;
    dev_free(c->d_bslots);
    c->d_bslots = nullptr;
    c->bslots = 0;
    HIP_OK(hipMalloc((void **) &c->d_bslots, c->blayout.bytes * (size_t) want));
    // zero once: the factorisation's dense input areas are only ever written at the pattern's
    // positions (block_factor), so their other positions must start at zero
    HIP_OK(hipMemset(c->d_bslots, 0, c->blayout.bytes * (size_t) want));
    c->bslots = want;
    return 0;
}

static int ensure_slots(nuppn_ctx *c, int want) {
    want = std::max(1, std::min(want, c->max_slots));
    if (want <= c->nslots) return 0;
    dev_free(c->d_slots);
    c->d_slots = nullptr;
    c->nslots = 0;
    HIP_OK(hipMalloc((void **) &c->d_slots, c->layout.bytes * (size_t) want));
    c->nslots = want;
    return 0;
}

static int ensure_zone_buffers(nuppn_ctx *c, int nz) {
    if (nz <= c->nz_cap) return 0;
    for (void *p : {(void *) c->d_yps, (void *) c->d_t9_0, (void *) c->d_t9_1, (void *) c->d_rho_0,
                    (void *) c->d_rho_1, (void *) c->d_dt, (void *) c->d_ye, (void *) c->d_ierr,
                    (void *) c->d_stats})
        dev_free(p);
    const size_t nsp = c->net.nsp;
    HIP_OK(hipMalloc((void **) &c->d_yps, sizeof(double) * nsp * nz));
    HIP_OK(hipMalloc((void **) &c->d_t9_0, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_t9_1, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_rho_0, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_rho_1, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_dt, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_ye, sizeof(double) * nz));
    HIP_OK(hipMalloc((void **) &c->d_ierr, sizeof(int) * nz));
    HIP_OK(hipMalloc((void **) &c->d_stats, sizeof(ZoneStats) * nz));
    c->nz_cap = nz;
    return 0;
}

static int ensure_rate_buffer(nuppn_ctx *c, int nz) {
    if (nz <= c->nz_vcap) return 0;
    dev_free(c->d_v);
    HIP_OK(hipMalloc((void **) &c->d_v, sizeof(double) * (size_t) c->net.nre * nz));
    c->nz_vcap = nz;
    return 0;
}

static int ensure_zone_state(nuppn_ctx *c, int nz) {
    if (nz <= c->nz_state_cap) return 0;
    for (void *p : {(void *) c->d_yps_old, (void *) c->d_y_old, (void *) c->d_y_cur,
                    (void *) c->d_y_new, (void *) c->d_work, (void *) c->d_state, (void *) c->d_bd,
                    (void *) c->d_active})
        dev_free(p);
    const size_t nsp = c->net.nsp;
    size_t pow2 = 1;
    while (pow2 < nsp) pow2 <<= 1;
    c->work_stride = nsp + pow2;
    HIP_OK(hipMalloc((void **) &c->d_yps_old, sizeof(double) * nsp * nz));
    HIP_OK(hipMalloc((void **) &c->d_y_old, sizeof(double) * nsp * nz));
    HIP_OK(hipMalloc((void **) &c->d_y_cur, sizeof(double) * nsp * nz));
    HIP_OK(hipMalloc((void **) &c->d_y_new, sizeof(double) * nsp * nz));
    HIP_OK(hipMalloc((void **) &c->d_work, sizeof(double) * c->work_stride * nz));
    HIP_OK(hipMalloc((void **) &c->d_state, sizeof(ZoneState) * nz));
    HIP_OK(hipMalloc((void **) &c->d_bd, sizeof(BDState) * nz));
    HIP_OK(hipMalloc((void **) &c->d_active, 2 * sizeof(int) * nz));
    c->nz_state_cap = nz;
    return 0;
}

static int upload_order(nuppn_ctx *c) {
    const int n = c->sym.n;
    c->iperm_c.assign(n, 0);
    for (int i = 0; i < n; ++i) c->iperm_c[c->sym.perm_c[i]] = i;
    HIP_OK(hipMemcpy(c->d_perm_c, c->sym.perm_c.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_iperm_c, c->iperm_c.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    return 0;
}

// bader_deuflhard_init (bader_deuflhard.F90:66-124) with the reference's settings (47-59):
// use_weights, cf = 1, cj = 0, cs = 10, clr = cj, safe1 = 1/4.  nseq(1:3) = 2, 6, 10, then
// nint(nseq(i-1) / (5/7)); work a(1) = cj + clr + (nseq(1)+1)(cf+cs), a(i) = a(i-1) +
// nseq(i)(cf+cs) + clr + cs for i < 9 (a(9) stays 0); correction factors alf(k, iq) =
// (dgrd safe1)^((a(k+1)-a(iq+1)) / ((a(iq+1)-a(1)+1)(2k+1))); kopt is the first row with
// a(kopt+1) > a(kopt) alf(kopt-1, kopt) (kmax = kopt; kmax stays 0 if none).
static void bd_tables(double dgrd, BDTables &T, BDState &bd) {
    const double cf = 1.0, cj = 0.0, cs = 10.0, clr = 0.0, safe1 = 0.25;
    memset(&T, 0, sizeof(T));
    T.nseq[0] = 2;
    T.nseq[1] = 6;
    T.nseq[2] = 10;
    for (int i = 3; i < 9; ++i) T.nseq[i] = (int) std::lround((double) T.nseq[i - 1] / (5.0 / 7.0));
    T.a[0] = cj + clr + (T.nseq[0] + 1) * (cf + cs);
    for (int i = 1; i < 8; ++i) T.a[i] = T.a[i - 1] + T.nseq[i] * (cf + cs) + clr + cs;
    for (int iq = 2; iq <= 8; ++iq)
        for (int k = 1; k <= iq - 1; ++k)
            T.alf[k - 1][iq - 1] =
                std::pow(dgrd * safe1, (T.a[k] - T.a[iq]) / ((T.a[iq] - T.a[0] + 1.0) * (2.0 * k + 1)));
    int ko;
    T.kmax = 0;
    for (ko = 2; ko <= 7; ++ko)
        if (T.a[ko] > T.a[ko - 1] * T.alf[ko - 2][ko - 1]) {
            T.kmax = ko;
            break;
        }
    T.kopt0 = ko;
    memset(&bd, 0, sizeof(bd));
    bd.frst = 1;
    bd.reduce = 0;
    bd.kopt = ko;
}

static KArgs kargs(nuppn_ctx *c) {
    KArgs a;
    memset(&a, 0, sizeof(a));
    a.N = dev_net(c);
    a.K = dev_knobs(c->kn);
    a.T = c->bdt;
    a.slot_base = c->d_slots;
    a.slot_bytes = c->layout.bytes;
    for (int i = 0; i < S_COUNT; ++i) a.off[i] = c->layout.off[i];
    return a;
}

extern "C" {

void nuppn_default_knobs(nuppn_knobs *k) {
    k->ittd = 6;
    k->irdn = 2;
    k->max_sub_steps = 4000;
    k->integration_method = 0;
    k->interpolate_temp = 1;
    k->interpolate_dens = 1;
    k->dgrd = 1.0e-3;
    k->grdthreshold = 1.0e-15;
    k->cyminformal = -1.0;
    k->min_rate_ov_dt = -1.0e99;
    k->min_t9_for_slv = 0.01;
    k->delta_state_max = 1.0e99;
}

int nuppn_column_order(int n, const int *colptr, const int *rowind, int *perm_c, int *etree) {
    if (n <= 0 || !colptr || !rowind || !perm_c) return NUPPN_ERR_ARG;
    std::vector<int> pc, et;
    const int rc = superlu_column_order(n, colptr, rowind, pc, &et);
    if (rc) return NUPPN_ERR_ARG;
    std::copy(pc.begin(), pc.end(), perm_c);
    if (etree) std::copy(et.begin(), et.end(), etree);
    return NUPPN_OK;
}

nuppn_ctx *nuppn_ctx_create(const nuppn_network *net, const nuppn_knobs *knobs, int *status) {
    int dummy;
    if (!status) status = &dummy;
    *status = NUPPN_ERR_ARG;
    if (!net || !net->k || !net->rtype || !net->coef || !net->aw || !net->zw) return nullptr;
    nuppn_knobs kn;
    if (knobs) kn = *knobs;
    else nuppn_default_knobs(&kn);
    if (kn.irdn != 2) {   // calculate_jacobian stops on any other choice (jac_rhs.F90:297-304)
        *status = NUPPN_ERR_UNSUPPORTED;
        return nullptr;
    }
    if (kn.integration_method != 0 && kn.integration_method != 1) {
        *status = NUPPN_ERR_UNSUPPORTED;
        return nullptr;
    }
    if (kn.cyminformal < 1.0e-99) kn.min_rate_ov_dt = -1.0e99;   // solver_knobs.F90:66-69
    auto c = std::make_unique<nuppn_ctx>();
    c->kn = kn;
    bd_tables(kn.dgrd, c->bdt, c->bd);
    HostNetwork &h = c->net;
    h.nsp = net->nsp;
    h.nre = net->nre;
    h.igamma = net->igamma;
    h.k.assign(net->k, net->k + 8 * (size_t) net->nre);
    h.rtype.assign(net->rtype, net->rtype + net->nre);
    h.coef.assign(net->coef, net->coef + 7 * (size_t) net->nre);
    h.aw.assign(net->aw, net->aw + net->nsp);
    h.zw.assign(net->zw, net->zw + net->nsp);
    if (build_symbolic(h, c->sym) != 0) return nullptr;
    Symbolic &s = c->sym;
    if (superlu_column_order(s.n, s.A_colptr.data(), s.A_rowind.data(), s.perm_c) != 0) return nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        *status = NUPPN_ERR_HIP;
        return nullptr;
    }
    *status = NUPPN_ERR_HIP;
    if (dev_upload(&c->d_k, h.k) || dev_upload(&c->d_rtype, h.rtype) ||
        dev_upload(&c->d_coef, h.coef) || dev_upload(&c->d_aw, h.aw) ||
        dev_upload(&c->d_zw, h.zw) || dev_upload(&c->d_red2sp, s.red2sp) ||
        dev_upload(&c->d_dx_ptr, s.dx_ptr) || dev_upload(&c->d_dx_ent, s.dx_ent) ||
        dev_upload(&c->d_jq_ptr, s.jq_ptr) || dev_upload(&c->d_jq_ent, s.jq_ent) ||
        dev_upload(&c->d_Acp, s.A_colptr) || dev_upload(&c->d_Ari, s.A_rowind) ||
        dev_upload(&c->d_isdiag, s.A_isdiag) || dev_upload(&c->d_dyn_q, s.dyn_q) ||
        dev_upload(&c->d_perm_c, (const int *) nullptr, (size_t) s.n) ||
        dev_upload(&c->d_iperm_c, (const int *) nullptr, (size_t) s.n) ||
        dev_upload(&c->d_flag, (const int *) nullptr, 2 * kPassRing) ||
        dev_upload(&c->d_status, (const int *) nullptr, 4))
        return nullptr;
    if (upload_order(c.get())) return nullptr;
    c->layout = slot_layout(h.nsp, h.nre, s.n, s.nnzA);
    {
        size_t pow2 = 1;
        while (pow2 < (size_t) h.nsp) pow2 <<= 1;
        c->sort_lds = 8 * pow2;
        if (c->sort_lds > 160 * 1024) {
            *status = NUPPN_ERR_UNSUPPORTED;
            return nullptr;
        }
        if (c->sort_lds > 64 * 1024 &&
            (hipFuncSetAttribute((const void *) zone_control_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) c->sort_lds) != hipSuccess ||
             hipFuncSetAttribute((const void *) zone_init_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) c->sort_lds) != hipSuccess ||
             hipFuncSetAttribute((const void *) zone_ye_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int) c->sort_lds) != hipSuccess))
            return nullptr;
    }
    c->lds = lds_bytes(s.n);
    if (c->lds > 160 * 1024) {
        *status = NUPPN_ERR_UNSUPPORTED;
        return nullptr;
    }
    const void *plain_kernel = kn.integration_method == 1 ? (const void *) zone_bd_kernel
                                                          : (const void *) zone_newton_kernel;
    if (ensure_stack(plain_kernel) || ensure_stack((const void *) zone_control_kernel) ||
        ensure_stack((const void *) zone_init_kernel) ||
        ensure_stack((const void *) zone_ye_kernel) || ensure_stack((const void *) lu_kernel) ||
        ensure_stack((const void *) assemble_kernel))
        return nullptr;
    if (c->lds > 64 * 1024 &&
        hipFuncSetAttribute(plain_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int) c->lds) != hipSuccess)
        return nullptr;
    if (c->lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) lu_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) c->lds);
    int dev = 0, ncu = 0, per_cu = 0;
    (void) hipGetDevice(&dev);
    (void) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, plain_kernel, 64,
                                                     c->lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    size_t free_b = 0, total_b = 0;
    (void) hipMemGetInfo(&free_b, &total_b);
    const size_t budget = free_b / 2;
    int by_mem = (int) std::max<size_t>(1, budget / c->layout.bytes);
    c->max_slots = std::max(1, std::min(ncu * per_cu, by_mem));
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
    for (auto &r : c->evn)
        for (auto &e : r)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
    for (auto &r : c->evc)
        for (auto &e : r)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
    for (auto &r : c->gevn)
        for (auto &e : r)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
    for (auto &r : c->gevc)
        for (auto &e : r)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
    if (hipHostMalloc((void **) &c->h_flag, 2 * kPassRing * sizeof(int), hipHostMallocDefault) !=
        hipSuccess)
        return nullptr;
    if (ensure_slots(c.get(), 1)) return nullptr;
    if (setup_batched(c.get())) return nullptr;
    *status = NUPPN_OK;
    return c.release();
}

void nuppn_ctx_destroy(nuppn_ctx *c) { delete c; }

int nuppn_set_column_order(nuppn_ctx *c, const int *perm_c) {
    if (!c || !perm_c) return NUPPN_ERR_ARG;
    const int n = c->sym.n;
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i) {
        if (perm_c[i] < 0 || perm_c[i] >= n || seen[perm_c[i]]) return NUPPN_ERR_ARG;
        seen[perm_c[i]] = 1;
    }
    std::lock_guard<std::mutex> g(c->mu);
    c->sym.perm_c.assign(perm_c, perm_c + n);
    return upload_order(c);
}

int nuppn_get_column_order(nuppn_ctx *c, int *perm_c) {
    if (!c || !perm_c) return NUPPN_ERR_ARG;
    std::copy(c->sym.perm_c.begin(), c->sym.perm_c.end(), perm_c);
    return NUPPN_OK;
}

int nuppn_set_rate_cache(nuppn_ctx *c, int use_cache) {
    if (!c) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    if (use_cache && !c->d_rcache) {
        const int nre = c->net.nre;
        c->rc_dt9 = (10.0 - 0.01) / (kNT9 - 1);
        std::vector<double> grid(kNT9);
        for (int i = 0; i < kNT9; ++i) grid[i] = 0.01 + i * c->rc_dt9;
        if (dev_upload(&c->d_rc_t9, grid) ||
            dev_upload(&c->d_rcache, (const double *) nullptr, (size_t) kNT9 * nre))
            return NUPPN_ERR_HIP;
        hipLaunchKernelGGL(rate_cache_kernel, dim3((nre + 255) / 256, kNT9), dim3(256), 0, 0,
                           dev_net(c), c->d_rc_t9, c->d_rcache);
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
    }
    c->use_cache = use_cache ? 1 : 0;
    return NUPPN_OK;
}

// reverse_init (reverse.F90:25-105): each forward reaction (not a REACLIB reverse rate, not
// weak, at most one heavy product of a kind) paired with the first reaction of mirrored
// topology, kept when that one is a REACLIB reverse rate (ground states only here)
static void reverse_pairs(const HostNetwork &h, const int *rev_ch, std::vector<int> &fwd,
                          std::vector<int> &rev) {
    const int nre = h.nre;
    const int *k = h.k.data();
    fwd.clear();
    rev.clear();
    for (int i = 0; i < nre; ++i) {
        if (h.rtype[i] == 2 || k[7 * nre + i] > 1 || rev_ch[i] > 0) continue;
        for (int j = 0; j < nre; ++j) {
            const int ri[8] = {k[i], k[nre + i], k[2 * nre + i], k[3 * nre + i], k[4 * nre + i],
                               k[5 * nre + i], k[6 * nre + i], k[7 * nre + i]};
            const int rj[8] = {k[6 * nre + j], k[7 * nre + j], k[4 * nre + j], k[5 * nre + j],
                               k[2 * nre + j], k[3 * nre + j], k[j], k[nre + j]};
            if (std::equal(ri, ri + 8, rj)) {
                if (rev_ch[j] > 0) {
                    fwd.push_back(i);
                    rev.push_back(j);
                }
                break;
            }
        }
    }
}

int nuppn_set_rate_physics(nuppn_ctx *c, const nuppn_rate_physics *rp) {
    if (!c) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    const int nsp = c->net.nsp, nre = c->net.nre;
    if (rp && rp->pf_n > 0) {
        if (rp->pf_n < 2 || !rp->pf_t9 || !rp->pf_tab || !rp->rev_ch || !rp->rev_part ||
            (rp->detailed_balance && (!rp->g_spin || !rp->mass_g || !rp->q_mev)))
            return NUPPN_ERR_ARG;
        for (int i = 1; i < rp->pf_n; ++i)
            if (!(rp->pf_t9[i] > rp->pf_t9[i - 1])) return NUPPN_ERR_ARG;
        for (size_t e = 0; e < 6 * (size_t) nre; ++e)
            if (rp->rev_part[e] < -1 || rp->rev_part[e] >= nsp) return NUPPN_ERR_ARG;
    }
    for (void *p : {(void *) c->d_pf_t9, (void *) c->d_pf_tab, (void *) c->d_g_spin,
                    (void *) c->d_mass_g, (void *) c->d_q_mev, (void *) c->d_rev_ch,
                    (void *) c->d_rev_part, (void *) c->d_db_fwd, (void *) c->d_db_rev})
        dev_free(p);
    c->d_pf_t9 = c->d_pf_tab = c->d_g_spin = c->d_mass_g = c->d_q_mev = nullptr;
    c->d_rev_ch = c->d_rev_part = c->d_db_fwd = c->d_db_rev = nullptr;
    c->pf_n = c->ndb = c->screen = 0;
    if (rp) {
        c->screen = rp->screen ? 1 : 0;
        if (rp->pf_n > 0) {
            if (dev_upload(&c->d_pf_t9, rp->pf_t9, rp->pf_n) ||
                dev_upload(&c->d_pf_tab, rp->pf_tab, (size_t) nsp * rp->pf_n) ||
                dev_upload(&c->d_rev_ch, rp->rev_ch, nre) ||
                dev_upload(&c->d_rev_part, rp->rev_part, 6 * (size_t) nre))
                return NUPPN_ERR_HIP;
            c->pf_n = rp->pf_n;
            if (rp->detailed_balance) {
                std::vector<int> fwd, rev;
                reverse_pairs(c->net, rp->rev_ch, fwd, rev);
                if (!fwd.empty() &&
                    (dev_upload(&c->d_db_fwd, fwd) || dev_upload(&c->d_db_rev, rev) ||
                     dev_upload(&c->d_g_spin, rp->g_spin, nsp) ||
                     dev_upload(&c->d_mass_g, rp->mass_g, nsp) ||
                     dev_upload(&c->d_q_mev, rp->q_mev, nsp)))
                    return NUPPN_ERR_HIP;
                c->ndb = (int) fwd.size();
            }
        }
    }
    // a T9 cache built before holds the rates without the partition-function factors
    if (c->d_rcache) {
        hipLaunchKernelGGL(rate_cache_kernel, dim3((nre + 255) / 256, kNT9), dim3(256), 0, 0,
                           dev_net(c), c->d_rc_t9, c->d_rcache);
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
    }
    return NUPPN_OK;
}

int nuppn_reverse_pairs(nuppn_ctx *c, int *irev) {
    if (!c || !irev) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    const int nre = c->net.nre;
    for (int i = 0; i < nre; ++i) irev[i] = -1;
    if (c->ndb == 0) return NUPPN_OK;
    std::vector<int> fwd(c->ndb), rev(c->ndb);
    HIP_OK(hipMemcpy(fwd.data(), c->d_db_fwd, sizeof(int) * c->ndb, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(rev.data(), c->d_db_rev, sizeof(int) * c->ndb, hipMemcpyDeviceToHost));
    for (int d = 0; d < c->ndb; ++d) irev[fwd[d]] = rev[d];
    return NUPPN_OK;
}

int nuppn_evaluate_all_rates(nuppn_ctx *c, int nz, const double *t9, const double *rho,
                             const double *x, double *v) {
    if (!c || nz <= 0 || !t9 || !rho || !x || !v) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    if (ensure_zone_buffers(c, nz) || ensure_rate_buffer(c, nz)) return NUPPN_ERR_HIP;
    const size_t nsp = c->net.nsp, nre = c->net.nre;
    HIP_OK(hipMemcpy(c->d_t9_0, t9, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_0, rho, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_yps, x, sizeof(double) * nsp * nz, hipMemcpyHostToDevice));
    launch_rates(c, nz, c->d_t9_0, c->d_t9_0, c->d_rho_0, c->d_rho_0, c->d_yps, nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(v, c->d_v, sizeof(double) * nre * (size_t) nz, hipMemcpyDeviceToHost));
    return NUPPN_OK;
}

int nuppn_rate_cache_table(nuppn_ctx *c, double *table) {
    if (!c || !table) return NUPPN_ERR_ARG;
    if (!c->d_rcache) return NUPPN_ERR_UNSUPPORTED;
    HIP_OK(hipMemcpy(table, c->d_rcache, sizeof(double) * kNT9 * (size_t) c->net.nre,
                     hipMemcpyDeviceToHost));
    return NUPPN_OK;
}

int nuppn_reduced_size(nuppn_ctx *c, int *n, int *nnz) {
    if (!c || !n || !nnz) return NUPPN_ERR_ARG;
    *n = c->sym.n;
    *nnz = c->sym.nnzA;
    return NUPPN_OK;
}

int nuppn_reduced_pattern(nuppn_ctx *c, int *colptr, int *rowind, int *red2sp) {
    if (!c) return NUPPN_ERR_ARG;
    if (colptr) std::copy(c->sym.A_colptr.begin(), c->sym.A_colptr.end(), colptr);
    if (rowind) std::copy(c->sym.A_rowind.begin(), c->sym.A_rowind.end(), rowind);
    if (red2sp) std::copy(c->sym.red2sp.begin(), c->sym.red2sp.end(), red2sp);
    return NUPPN_OK;
}

int nuppn_rates_eval(nuppn_ctx *c, int nz, const double *t9, const double *rho, const double *ye,
                     double *v) {
    (void) ye;   // rho*ye scaling applies to '  ec' rates, which the network format excludes
    if (!c || nz <= 0 || !t9 || !rho || !v) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    if (ensure_zone_buffers(c, nz) || ensure_rate_buffer(c, nz)) return NUPPN_ERR_HIP;
    HIP_OK(hipMemcpy(c->d_t9_0, t9, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_0, rho, sizeof(double) * nz, hipMemcpyHostToDevice));
    const int nre = c->net.nre;
    hipLaunchKernelGGL(rates_kernel, rates_grid(nre, nz), dim3(256), 0, 0, dev_net(c), nz,
                       c->d_t9_0, c->d_t9_0, c->d_rho_0, c->d_rho_0, c->d_v);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(v, c->d_v, sizeof(double) * nre * (size_t) nz, hipMemcpyDeviceToHost));
    return NUPPN_OK;
}

static int assemble_call(nuppn_ctx *c, int mode, const double *x, const double *v, double dt,
                         double *out, size_t out_count) {
    std::lock_guard<std::mutex> g(c->mu);
    if (ensure_slots(c, 1)) return NUPPN_ERR_HIP;
    double *d_x = nullptr, *d_v = nullptr, *d_out = nullptr;
    if (dev_upload(&d_x, x, c->net.nsp) || dev_upload(&d_v, v, c->net.nre) ||
        dev_upload(&d_out, (const double *) nullptr, out_count))
        return NUPPN_ERR_HIP;
    KArgs a = kargs(c);
    hipLaunchKernelGGL(assemble_kernel, dim3(1), dim3(64), 0, 0, a, mode, d_x, d_v, dt, d_out,
                       c->d_status);
    int st = 0;
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(out, d_out, sizeof(double) * out_count, hipMemcpyDeviceToHost));
    if (mode == 1) HIP_OK(hipMemcpy(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost));
    dev_free(d_x);
    dev_free(d_v);
    dev_free(d_out);
    return st;
}

int nuppn_dxdt(nuppn_ctx *c, const double *x, const double *v, double *dxdt) {
    if (!c || !x || !v || !dxdt) return NUPPN_ERR_ARG;
    return assemble_call(c, 0, x, v, 0.0, dxdt, c->net.nsp);
}

int nuppn_jacobian_reduced(nuppn_ctx *c, double dt, const double *x, const double *v,
                           double *csc_val) {
    if (!c || !x || !v || !csc_val) return NUPPN_ERR_ARG;
    return assemble_call(c, 1, x, v, dt, csc_val, c->sym.nnzA);
}

int nuppn_lu_solve_reduced(nuppn_ctx *c, const double *csc_val, int nrhs, double *b, int *perm_r,
                           int *nnz_lu) {
    if (!c || !csc_val || nrhs < 0 || (nrhs > 0 && !b)) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    if (ensure_slots(c, 1)) return NUPPN_ERR_HIP;
    const int n = c->sym.n;
    double *d_val = nullptr, *d_b = nullptr;
    int *d_pr = nullptr, *d_lp = nullptr;
    if (dev_upload(&d_val, csc_val, c->sym.nnzA) ||
        dev_upload(&d_b, nrhs ? b : (const double *) nullptr, (size_t) n * std::max(nrhs, 1)) ||
        dev_upload(&d_pr, (const int *) nullptr, n) ||
        dev_upload(&d_lp, (const int *) nullptr, n + 1))
        return NUPPN_ERR_HIP;
    const Slot dummy{};
    (void) dummy;
    LUStore st;
    char *base = c->d_slots;
    st.Li = (int *) (base + c->layout.off[S_LI]);
    st.Lx = (double *) (base + c->layout.off[S_LX]);
    st.Ui = (int *) (base + c->layout.off[S_UI]);
    st.Ux = (double *) (base + c->layout.off[S_UX]);
    st.Up = (int *) (base + c->layout.off[S_UP]);
    hipLaunchKernelGGL(lu_kernel, dim3(1), dim3(64), c->lds, 0, n, c->d_Acp, c->d_Ari, d_val,
                       c->d_perm_c, c->d_iperm_c, nrhs, d_b, st, d_pr, d_lp, c->d_status, 1);
    HIP_OK(hipGetLastError());
    int info = 0;
    HIP_OK(hipMemcpy(&info, c->d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (nrhs) HIP_OK(hipMemcpy(b, d_b, sizeof(double) * n * (size_t) nrhs, hipMemcpyDeviceToHost));
    if (perm_r) HIP_OK(hipMemcpy(perm_r, d_pr, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (nnz_lu) {
        int lpn = 0, upn = 0;
        HIP_OK(hipMemcpy(&lpn, d_lp + n, sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&upn, st.Up + n, sizeof(int), hipMemcpyDeviceToHost));
        nnz_lu[0] = lpn;
        nnz_lu[1] = upn;
    }
    dev_free(d_val);
    dev_free(d_b);
    dev_free(d_pr);
    dev_free(d_lp);
    return info ? NUPPN_ERR_SINGULAR : NUPPN_OK;
}

}  // extern "C"

// integrate_network for a batch of zones: rates at the cycle mid-point, then the per-zone state
// machine (control kernel: sub-step and Newton decisions; Newton kernel: one Newton iteration per
// zone that needs one) until every zone is done.  `bo` null: the plain path (device GEPP with the
// reference's pivot order); otherwise the block-factorisation Newton kernels with those options.
static int run_batch(nuppn_ctx *c, int nz, double *yps, const double *t9_0, const double *t9_1,
                     const double *rho_0, const double *rho_1, double *ye, const double *dtwant,
                     int *ierr, nuppn_zone_stats *stats, const nuppn_batch_opts *bo,
                     hipStream_t s) {
    const bool block = bo != nullptr;
    if (block) {
        if (ensure_bslots(c, nz)) return NUPPN_ERR_HIP;
    } else if (ensure_slots(c, nz)) {
        return NUPPN_ERR_HIP;
    }
    if (ensure_rate_buffer(c, nz) || ensure_zone_state(c, nz)) return NUPPN_ERR_HIP;
    const int nre = c->net.nre;
    HIP_OK(hipEventRecord(c->ev[0], s));
    launch_rates(c, nz, t9_0, t9_1, rho_0, rho_1, yps, s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(c->ev[1], s));
    KArgs a = kargs(c);
    ZoneData &Z = a.Z;
    Z.yps = yps;
    Z.yps_old = c->d_yps_old;
    Z.y_old = c->d_y_old;
    Z.y_cur = c->d_y_cur;
    Z.y_new = c->d_y_new;
    Z.v = c->d_v;
    Z.work = c->d_work;
    Z.nsp_stride = (size_t) c->net.nsp;
    Z.nre_stride = (size_t) nre;
    Z.work_stride = c->work_stride;
    Z.t9_0 = t9_0;
    Z.t9_1 = t9_1;
    Z.rho_0 = rho_0;
    Z.rho_1 = rho_1;
    Z.dtwant = dtwant;
    Z.ye = ye;
    Z.ierr = ierr;
    Z.stats = (ZoneStats *) stats;
    Z.state = c->d_state;
    Z.bd = c->d_bd;
    Z.active = c->d_active;
    Z.ctl = nullptr;
    Z.nctl = nz;
    Z.bd0 = c->bd;
    Z.flag = c->d_flag;
    Z.nz = nz;
    Z.init_prec = (block && bo->factor_precision == 64) ? 1 : 0;
    BKArgs b;
    memset(&b, 0, sizeof(b));
    if (block) {
        b.B = dev_block(c);
        b.D = dev_band(c);
        b.O.fp64_factors = bo->factor_precision == 64;
        b.O.max_refine = bo->max_refine;
        b.O.refine_tol = bo->refine_tol;
        b.O.accept_tol = bo->accept_tol;
        b.slot_base = c->d_bslots;
        b.slot_bytes = c->blayout.bytes;
        for (int i = 0; i < B_COUNT; ++i) b.off[i] = c->blayout.off[i];
        for (int i = 0; i < F_COUNT; ++i) b.foff[i] = c->blayout.foff[i];
    }
    const int grid_n = std::min(nz, block ? c->bslots : c->nslots);
    HIP_OK(hipEventRecord(c->ev[2], s));
    hipLaunchKernelGGL(zone_init_kernel, dim3(nz), dim3(64), c->sort_lds, s, a);
    HIP_OK(hipGetLastError());
    // Each pass is a flag reset, a control launch and a worker launch; it advances every zone by
    // at least one Newton iteration or finishes it, so the work-list counts never grow.  The
    // counts stay on the device (pass p uses flag pair p % kPassRing): the control kernel reads
    // the previous pass's count and the worker its own, and the host queues up to kPassRing
    // passes ahead, sizing their grids by the last count it has read (an upper bound; surplus
    // workgroups find no list entry).  It reads the counts back after each chunk of passes.
    const long long max_pass = (long long) c->kn.max_sub_steps * (c->kn.ittd + 2) + 16;
    const bool bd = c->kn.integration_method == 1;
    const int queue = block ? std::max(1, std::min(bo->queue_passes, kPassRing)) : kPassRing;
    const int team_max = block ? (bo->team_zones < 0 ? INT_MAX : bo->team_zones) : 0;
    c->newton_ms = 0.0;
    c->control_ms = 0.0;
    c->newton_launches = 0;
    c->graph_launches = 0;
    PassGraph graph;
    bool graph_off = false;   // set when a capture fails
    long long pass = 0, done_pass = -1;
    int bound = nz;   // upper bound of the zones the next control pass takes
    // one pass: its kernels and timing events (on a capturing stream they become graph nodes)
    // (st: the stream launched on; capturing: st records into a graph, where launch errors
    // surface at the end of the capture)
    const char *failed_at = "";
    // one pass: its kernels and timing events, launched on stream s, or (g set) appended to
    // graph g as a chain of kernel and event-record nodes (built node by node: HIP runtimes that
    // cannot capture event records into a graph still take event-record nodes)
    auto enqueue_pass = [&](long long p, int bnd, hipGraph_t g) -> hipError_t {
        hipEvent_t(&evc)[kPassRing][2] = g ? c->gevc : c->evc;
        hipEvent_t(&evn)[kPassRing][1] = g ? c->gevn : c->evn;
        hipGraphNode_t last = g ? c->glast : nullptr;   // chain after the chunk's previous pass
        auto kernel = [&](const void *f, dim3 grid, dim3 blk, void **args, size_t lds) {
            if (!g) return hipLaunchKernel(f, grid, blk, args, lds, s);
            hipKernelNodeParams kp;
            memset(&kp, 0, sizeof(kp));
            kp.func = const_cast<void *>(f);
            kp.gridDim = grid;
            kp.blockDim = blk;
            kp.sharedMemBytes = (unsigned) lds;
            kp.kernelParams = args;
            hipGraphNode_t n = nullptr;
            const hipError_t e = hipGraphAddKernelNode(&n, g, last ? &last : nullptr,
                                                       last ? 1 : 0, &kp);
            if (e == hipSuccess) last = n;
            return e;
        };
        auto event = [&](hipEvent_t ev) {
            if (!g) return hipEventRecord(ev, s);
            hipGraphNode_t n = nullptr;
            const hipError_t e = hipGraphAddEventRecordNode(&n, g, last ? &last : nullptr,
                                                            last ? 1 : 0, ev);
            if (e == hipSuccess) last = n;
            return e;
        };
        const int slot = (int) (p % kPassRing);
        KArgs ka_local;
        int *flag_local = nullptr;
        KArgs &ka = g ? graph.ka[slot] : ka_local;
        int *&flag = g ? graph.flag[slot] : flag_local;
        flag = c->d_flag + 2 * slot;
        ka = a;
        ka.Z.flag = flag;
        ka.Z.active = c->d_active + (size_t) (p & 1) * nz;
        if (p == 0) {
            ka.Z.ctl = nullptr;
            ka.Z.nctl = nz;
            ka.Z.nctl_dev = nullptr;
        } else {
            ka.Z.ctl = c->d_active + (size_t) ((p - 1) & 1) * nz;
            ka.Z.nctl = 0;
            ka.Z.nctl_dev = c->d_flag + 2 * (int) ((p - 1) % kPassRing) + 1;
        }
        ka.Z.nact = 0;
        ka.Z.nact_dev = flag + 1;
        failed_at = "flag reset";
        void *fargs[] = {&flag};
        hipError_t e = kernel((const void *) flag_reset_kernel, dim3(1), dim3(1), fargs, 0);
        if (e != hipSuccess) return e;
        failed_at = "control start event";
        if ((e = event(evc[slot][0])) != hipSuccess) return e;
        failed_at = "control launch";
        void *cargs[] = {&ka};
        e = kernel((const void *) zone_control_kernel, dim3(std::max(1, bnd)), dim3(64), cargs,
                   c->sort_lds);
        if (e != hipSuccess) return e;
        failed_at = "control end event";
        if ((e = event(evc[slot][1])) != hipSuccess) return e;
        failed_at = "worker launch";
        const int nl = std::max(1, std::min(bnd, grid_n));
        if (!block) {
            void *args[] = {&ka};
            e = kernel(bd ? (const void *) zone_bd_kernel : (const void *) zone_newton_kernel,
                       dim3(nl), dim3(64), args, c->lds);
        } else if (c->bnb == kBandForm) {
            void *args[] = {&ka, &b};
            e = kernel((const void *) zone_newton_band_kernel,
                       dim3(std::max(1, std::min(nl, c->bgrid[1]))), dim3(256), args, c->blds64);
        } else {
            void *args[] = {&ka, &b};
            // several wavefronts per zone shorten each one's iteration; with four a zone takes
            // a whole CU, and the zones of a longer list follow one another as workgroups finish
            // (measured faster than one wavefront per zone even with 2,048 zones in the list)
            int nw = bo->team_waves ? bo->team_waves : 4;
            if (!c->team_ok[nw]) nw = c->team_ok[2] ? 2 : 0;
            if (!bd && nw && bnd <= team_max) {
                // a list longer than one zone per CU: two four-wavefront workgroups per CU
                const bool two = nw == 4 && c->bnb == 32 && bnd > c->ncu && c->team_ok2;
                e = kernel(two ? (const void *) zone_newton_team_kernel<32, 4, 2>
                               : team_kernel_ptr(c->bnb, nw),
                           dim3(nl), dim3(64 * nw), args, c->blds_team);
            } else {
                // the register-lean two-per-SIMD variant only when the single-wavefront one
                // cannot give every zone of the list its own wavefront
                const int waves = (!bd && bnd > c->bgrid[1]) ? 2 : 1;
                const int grid = std::max(1, std::min(nl, c->bgrid[waves]));
                e = kernel(block_kernel_ptr(c->bnb, bd ? 1 : 0, waves), dim3(grid), dim3(64),
                           args, c->blds64);
            }
        }
        if (e != hipSuccess) return e;
        failed_at = "worker end event";
        if ((e = event(evn[slot][0])) != hipSuccess) return e;
        if (g) c->glast = last;
        return hipSuccess;
    };
    while (done_pass < 0) {
        if (pass >= max_pass) return NUPPN_ERR_HIP;
        // chunk length: one pass at a time while the lists are long (long passes; their grids
        // and kernel variants follow the counts exactly), more as they shorten
        int ch = 1;
        if (pass > 0 && (!block || bound <= team_max))
            ch = (!block || bound <= c->ncu) ? queue : std::min(queue, 4);
        // graph chunks start at a ring boundary: a shorter chunk first reaches one
        if (block && bo->use_graph && ch == kPassRing && pass % kPassRing != 0)
            ch = kPassRing - (int) (pass % kPassRing);
        // a chunk that starts at a ring boundary with a short list can be replayed from a
        // prebuilt graph: pass p's arguments depend only on p % kPassRing, except the grids
        const bool use_graph = block && bo->use_graph && !graph_off && ch == kPassRing &&
                               pass % kPassRing == 0;
        if (use_graph) {
            // grids sized for the next power of two at or above the count (at most the CU
            // count, which a graph chunk's count never exceeds), so that a built chunk
            // serves until the count halves
            int gb = 1;
            while (gb < bound) gb <<= 1;
            gb = std::max(bound, std::min(gb, c->ncu));
            if (!graph.exec || graph.bound != gb) {
                graph.reset();
                hipGraph_t g = nullptr;
                HIP_OK(hipGraphCreate(&g, 0));
                graph.graph = g;
                c->glast = nullptr;
                hipError_t e = hipSuccess;
                for (int q = 0; q < ch && e == hipSuccess; ++q) e = enqueue_pass(pass + q, gb, g);
                hipError_t e3 = hipSuccess;
                if (e == hipSuccess) e3 = hipGraphInstantiate(&graph.exec, g, nullptr, nullptr, 0);
                if (e != hipSuccess || e3 != hipSuccess) {
                    // the rest of the call runs on plain launches
                    fprintf(stderr, "nuppn_hip: pass graph not built (%s: %s; "
                            "instantiation: %s); plain launches follow\n", failed_at,
                            hipGetErrorString(e), hipGetErrorString(e3));
                    (void) hipGetLastError();
                    graph.reset();
                    graph_off = true;
                }
                graph.bound = gb;
            }
        }
        if (use_graph && !graph_off) {
            HIP_OK(hipGraphLaunch(graph.exec, s));
            c->graph_launches++;
        } else {
            for (int q = 0; q < ch; ++q) HIP_OK(enqueue_pass(pass + q, bound, nullptr));
            HIP_OK(hipGetLastError());
        }
        HIP_OK(hipMemcpyAsync(c->h_flag, c->d_flag, 2 * kPassRing * sizeof(int),
                              hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        const bool graphed = use_graph && !graph_off;
        hipEvent_t(&evc)[kPassRing][2] = graphed ? c->gevc : c->evc;
        hipEvent_t(&evn)[kPassRing][1] = graphed ? c->gevn : c->evn;
        for (int q = 0; q < ch; ++q) {
            const int slot = (int) ((pass + q) % kPassRing);
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, evc[slot][0], evc[slot][1]));
            c->control_ms += ms;
            if (c->h_flag[2 * slot] == 0) {   // no zone needed a Newton iteration: all done
                done_pass = pass + q;
                break;
            }
            HIP_OK(hipEventElapsedTime(&ms, evc[slot][1], evn[slot][0]));
            c->newton_ms += ms;
            c->newton_launches++;
            bound = std::min(c->h_flag[2 * slot + 1], nz);
        }
        pass += ch;
    }
    c->last_passes = done_pass;
    if (!bd) {
        hipLaunchKernelGGL(zone_ye_kernel, dim3(nz), dim3(64), c->sort_lds, s, a);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(c->ev[3], s));
    if (c->kn.integration_method == 1) {
        // the state the reference would carry into its next call: the last zone's (exact for
        // one zone per call; a batch runs its zones from the same entry state)
        HIP_OK(hipMemcpyAsync(&c->bd, c->d_bd + (nz - 1), sizeof(BDState), hipMemcpyDeviceToHost,
                              s));
        HIP_OK(hipStreamSynchronize(s));
    }
    return NUPPN_OK;
}

extern "C" {

int nuppn_integrate_batch_dev(nuppn_ctx *c, int nz, double *yps, const double *t9_0,
                              const double *t9_1, const double *rho_0, const double *rho_1,
                              double *ye, const double *dtwant, int *ierr,
                              nuppn_zone_stats *stats, void *stream) {
    if (!c || nz <= 0 || !yps || !t9_0 || !t9_1 || !rho_0 || !rho_1 || !ye || !dtwant || !ierr)
        return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    return run_batch(c, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, nullptr,
                     (hipStream_t) stream);
}

int nuppn_last_timing(nuppn_ctx *c, double *out, int nout) {
    if (!c || !out || nout <= 0) return NUPPN_ERR_ARG;
    HIP_OK(hipEventSynchronize(c->ev[3]));
    float r = 0.f, t = 0.f;
    HIP_OK(hipEventElapsedTime(&r, c->ev[0], c->ev[1]));
    HIP_OK(hipEventElapsedTime(&t, c->ev[2], c->ev[3]));
    const double nms = c->newton_ms < 0.0 ? (double) t : c->newton_ms;
    const double vals[7] = {r, t, nms, c->control_ms, (double) c->newton_launches,
                            (double) c->last_passes, (double) c->graph_launches};
    for (int i = 0; i < nout && i < 7; ++i) out[i] = vals[i];
    return NUPPN_OK;
}

// the mppnp cycle's device buffers for m zones (kept between calls)
static int ensure_mp(nuppn_ctx *c, int m) {
    if (m <= c->mp_cap) return 0;
    for (void *p : {(void *) c->mp_zidx, (void *) c->mp_ierrb, (void *) c->mp_small,
                    (void *) c->mp_yb, (void *) c->mp_t9b, (void *) c->mp_rhob, (void *) c->mp_dtb,
                    (void *) c->mp_yeb, (void *) c->mp_dl, (void *) c->mp_a, (void *) c->mp_b,
                    (void *) c->mp_c, (void *) c->mp_bet, (void *) c->mp_gam, (void *) c->mp_res,
                    (void *) c->mp_max, (void *) c->mp_stb})
        dev_free(p);
    c->mp_cap = 0;
    const size_t nsp = c->net.nsp;
    HIP_OK(hipMalloc((void **) &c->mp_zidx, sizeof(int) * m));
    HIP_OK(hipMalloc((void **) &c->mp_ierrb, sizeof(int) * m));
    HIP_OK(hipMalloc((void **) &c->mp_small, sizeof(int) * 4));
    HIP_OK(hipMalloc((void **) &c->mp_yb, sizeof(double) * nsp * m));
    for (double **p : {&c->mp_t9b, &c->mp_rhob, &c->mp_dtb, &c->mp_yeb, &c->mp_dl, &c->mp_a,
                       &c->mp_b, &c->mp_c, &c->mp_bet, &c->mp_gam, &c->mp_res})
        HIP_OK(hipMalloc((void **) p, sizeof(double) * m));
    HIP_OK(hipMalloc((void **) &c->mp_max, sizeof(double)));
    HIP_OK(hipMalloc((void **) &c->mp_stb, sizeof(ZoneStats) * m));
    c->mp_cap = m;
    return 0;
}

// One mppnp cycle on device arrays (frames/mppnp/source/mppnp.F90:637-781): zoneselect, the
// selected zones integrated as one batch (bo null: the plain FP64 path, else the batched path),
// renorm_1d above 5e-7 per zone, then, unless a zone failed (the reference stops there), the
// mixing of all m zones, renorm_2d when the largest residual reaches SUM_TOL, and the clip.
static int run_mppnp(nuppn_ctx *c, int m, double *yps, const double *t9, const double *rho,
                     const double *xm, const double *dq, const double *rad, const double *d_eul,
                     double dzeit, const nuppn_mppnp_opts &mo, const nuppn_batch_opts *bo,
                     int *ierr, nuppn_zone_stats *stats, hipStream_t s) {
    const int nsp = c->net.nsp;
    if ((size_t) nsp * 8 > 64 * 1024) return NUPPN_ERR_UNSUPPORTED;
    if (ensure_mp(c, m)) return NUPPN_ERR_HIP;
    ZoneStats *st = (ZoneStats *) stats;
    int *cnt = c->mp_small, *bad = c->mp_small + 1;
    hipLaunchKernelGGL(mp_select_kernel, dim3(1), dim3(64), 0, s, m, t9, mo.t9threshold,
                       c->mp_zidx, cnt);
    const int g = std::max(1, std::min(4096, (int) (((size_t) m * nsp + 255) / 256)));
    hipLaunchKernelGGL(mp_gather_kernel, dim3(g), dim3(256), 0, s, m, nsp, c->mp_zidx, cnt, yps,
                       t9, rho, dzeit, c->mp_yb, c->mp_t9b, c->mp_rhob, c->mp_dtb, c->mp_yeb,
                       ierr, st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(c->h_flag, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const int n = c->h_flag[0];
    const size_t lds = (size_t) nsp * 8;
    if (n > 0) {
        const int rc = run_batch(c, n, c->mp_yb, c->mp_t9b, c->mp_t9b, c->mp_rhob, c->mp_rhob,
                                 c->mp_yeb, c->mp_dtb, c->mp_ierrb,
                                 (nuppn_zone_stats *) c->mp_stb, bo, s);
        if (rc != NUPPN_OK) return rc;
        hipLaunchKernelGGL(mp_scatter_kernel, dim3(std::min(n, 4096)), dim3(64), lds, s, nsp,
                           c->mp_zidx, cnt, c->mp_yb, yps, c->mp_ierrb, c->mp_stb, ierr, st);
        HIP_OK(hipGetLastError());
        // mppnp stops on any integration error (mppnp.F90:726-735): no mixing then
        std::vector<int> eb(n);
        HIP_OK(hipMemcpyAsync(eb.data(), c->mp_ierrb, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        for (int k = 0; k < n; ++k)
            if (eb[k] != 0) return NUPPN_OK;
    }
    // do_mixing (mixing.F90:30-64)
    HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), s));
    const int gm = std::max(1, std::min(1024, (m + 255) / 256));
    hipLaunchKernelGGL(mp_dlagr_kernel, dim3(gm), dim3(256), 0, s, m, rho, rad, d_eul, c->mp_dl);
    hipLaunchKernelGGL(mp_matrix_kernel, dim3(gm), dim3(256), 0, s, m, dzeit, xm, dq, c->mp_dl,
                       mo.sig_term_limit, mo.code_source, c->mp_a, c->mp_b, c->mp_c, bad);
    hipLaunchKernelGGL(mp_thomas_kernel, dim3(1), dim3(64), 0, s, m, c->mp_a, c->mp_b, c->mp_c,
                       c->mp_bet, c->mp_gam, bad);
    hipLaunchKernelGGL(mp_species_kernel, dim3((nsp + 255) / 256), dim3(256), 0, s, m, nsp,
                       c->net.igamma, yps, c->mp_a, c->mp_bet, c->mp_gam, bad);
    hipLaunchKernelGGL(mp_residual_kernel, dim3(std::min(m, 4096)), dim3(64), lds, s, m, nsp, yps,
                       c->mp_res);
    hipLaunchKernelGGL(mp_max_kernel, dim3(1), dim3(64), 0, s, m, c->mp_res, c->mp_max);
    hipLaunchKernelGGL(mp_renorm_clip_kernel, dim3(std::min(m, 4096)), dim3(64), lds, s, m, nsp,
                       yps, c->mp_max);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(c->h_flag + 1, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return c->h_flag[1] ? NUPPN_ERR_MIXING : NUPPN_OK;
}

// A host-array batch call: the staging buffers are filled, the batch run and the results read
// back under the context's lock, so that concurrent calls on one context serialise whole.
static int host_batch(nuppn_ctx *c, int nz, double *yps, const double *t9_0, const double *t9_1,
                      const double *rho_0, const double *rho_1, double *ye, const double *dtwant,
                      int *ierr, nuppn_zone_stats *stats, const nuppn_batch_opts *bo) {
    std::lock_guard<std::mutex> g(c->mu);
    if (ensure_zone_buffers(c, nz)) return NUPPN_ERR_HIP;
    const size_t nsp = c->net.nsp;
    HIP_OK(hipMemcpy(c->d_yps, yps, sizeof(double) * nsp * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_t9_0, t9_0, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_t9_1, t9_1, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_0, rho_0, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_1, rho_1, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_dt, dtwant, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_ye, ye, sizeof(double) * nz, hipMemcpyHostToDevice));
    const int rc = run_batch(c, nz, c->d_yps, c->d_t9_0, c->d_t9_1, c->d_rho_0, c->d_rho_1,
                             c->d_ye, c->d_dt, c->d_ierr, (nuppn_zone_stats *) c->d_stats, bo,
                             nullptr);
    if (rc) return rc;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(yps, c->d_yps, sizeof(double) * nsp * nz, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ye, c->d_ye, sizeof(double) * nz, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ierr, c->d_ierr, sizeof(int) * nz, hipMemcpyDeviceToHost));
    if (stats)
        HIP_OK(hipMemcpy(stats, c->d_stats, sizeof(ZoneStats) * nz, hipMemcpyDeviceToHost));
    return NUPPN_OK;
}

// the batched path's options, checked (defaults for null: the autotuner's choice once it ran)
static int mixed_opts(const nuppn_batch_opts *opts, nuppn_batch_opts *o,
                      const nuppn_ctx *c = nullptr) {
    if (opts) *o = *opts;
    else if (c && c->tuned) *o = c->tuned_opts;
    else nuppn_default_batch_opts(o);
    if ((o->factor_precision != 32 && o->factor_precision != 64) || o->max_refine < 1 ||
        (o->team_waves != 0 && o->team_waves != 2 && o->team_waves != 4) ||
        o->queue_passes < 1 || o->queue_passes > 16 || (o->use_graph != 0 && o->use_graph != 1))
        return NUPPN_ERR_ARG;
    return NUPPN_OK;
}

int nuppn_integrate_batch(nuppn_ctx *c, int nz, double *yps, const double *t9_0,
                          const double *t9_1, const double *rho_0, const double *rho_1, double *ye,
                          const double *dtwant, int *ierr, nuppn_zone_stats *stats) {
    if (!c || nz <= 0 || !yps || !t9_0 || !t9_1 || !rho_0 || !rho_1 || !ye || !dtwant || !ierr)
        return NUPPN_ERR_ARG;
    return host_batch(c, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, nullptr);
}

void nuppn_default_batch_opts(nuppn_batch_opts *o) {
    o->factor_precision = 32;
    o->max_refine = 12;
    o->refine_tol = 1.0e-17;
    o->accept_tol = 0.0;
    o->team_zones = -1;
    o->team_waves = 0;
    o->queue_passes = 16;
    o->use_graph = 1;
}

void nuppn_default_mppnp_opts(nuppn_mppnp_opts *o) {
    o->t9threshold = 0.0;
    o->sig_term_limit = 1.0e10;
    o->code_source = 0;
    o->path = 1;
}

static int mppnp_checked(nuppn_ctx *c, int m, const nuppn_mppnp_opts *mo_in,
                         const nuppn_batch_opts *bo_in, nuppn_mppnp_opts *mo,
                         nuppn_batch_opts *bo) {
    if (!c || m < 2) return NUPPN_ERR_ARG;
    if (mo_in) *mo = *mo_in;
    else nuppn_default_mppnp_opts(mo);
    if ((mo->code_source != 0 && mo->code_source != 1) || (mo->path != 0 && mo->path != 1))
        return NUPPN_ERR_ARG;
    if (mo->path == 1) {
        if (mixed_opts(bo_in, bo, c) != NUPPN_OK) return NUPPN_ERR_ARG;
        if (!c->bnb || !batched_method_enabled(c)) return NUPPN_ERR_UNSUPPORTED;
    }
    return NUPPN_OK;
}

int nuppn_mppnp_cycle_dev(nuppn_ctx *c, int m, double *yps, const double *t9, const double *rho,
                          const double *xm, const double *dq, const double *rad,
                          const double *d_eul, double dzeit, const nuppn_mppnp_opts *opts,
                          const nuppn_batch_opts *bopts, int *ierr, nuppn_zone_stats *stats,
                          void *stream) {
    nuppn_mppnp_opts mo;
    nuppn_batch_opts bo;
    const int rc = mppnp_checked(c, m, opts, bopts, &mo, &bo);
    if (rc != NUPPN_OK) return rc;
    if (!yps || !t9 || !rho || !xm || !dq || !rad || !d_eul || !ierr) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    return run_mppnp(c, m, yps, t9, rho, xm, dq, rad, d_eul, dzeit, mo,
                     mo.path == 1 ? &bo : nullptr, ierr, stats, (hipStream_t) stream);
}

int nuppn_mppnp_cycle(nuppn_ctx *c, int m, double *yps, const double *t9, const double *rho,
                      const double *xm, const double *dq, const double *rad, const double *d_eul,
                      double dzeit, const nuppn_mppnp_opts *opts, const nuppn_batch_opts *bopts,
                      int *ierr, nuppn_zone_stats *stats) {
    nuppn_mppnp_opts mo;
    nuppn_batch_opts bo;
    int rc = mppnp_checked(c, m, opts, bopts, &mo, &bo);
    if (rc != NUPPN_OK) return rc;
    if (!yps || !t9 || !rho || !xm || !dq || !rad || !d_eul || !ierr) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    const size_t nsp = c->net.nsp, md = (size_t) m * sizeof(double);
    // one staging block: yps, t9, rho, xm, dq, rad, d_eul, then ierr and the statistics
    char *d = nullptr;
    const size_t bytes = nsp * md + 6 * md + m * (sizeof(int) + sizeof(ZoneStats)) + 64;
    HIP_OK(hipMalloc((void **) &d, bytes));
    double *dy = (double *) d, *dv = dy + nsp * m;
    const double *src[6] = {t9, rho, xm, dq, rad, d_eul};
    int *di = (int *) (dv + 6 * (size_t) m);
    ZoneStats *ds = (ZoneStats *) (((uintptr_t) (di + m) + 15) & ~(uintptr_t) 15);
    rc = NUPPN_ERR_HIP;
    if (hipMemcpy(dy, yps, nsp * md, hipMemcpyHostToDevice) == hipSuccess) {
        bool ok = true;
        for (int k = 0; k < 6 && ok; ++k)
            ok = hipMemcpy(dv + k * (size_t) m, src[k], md, hipMemcpyHostToDevice) == hipSuccess;
        if (ok) {
            rc = run_mppnp(c, m, dy, dv, dv + m, dv + 2 * (size_t) m, dv + 3 * (size_t) m,
                           dv + 4 * (size_t) m, dv + 5 * (size_t) m, dzeit, mo,
                           mo.path == 1 ? &bo : nullptr, di, (nuppn_zone_stats *) ds, nullptr);
            if ((rc == NUPPN_OK || rc == NUPPN_ERR_MIXING) &&
                (hipMemcpy(yps, dy, nsp * md, hipMemcpyDeviceToHost) != hipSuccess ||
                 hipMemcpy(ierr, di, sizeof(int) * m, hipMemcpyDeviceToHost) != hipSuccess ||
                 (stats && hipMemcpy(stats, ds, sizeof(ZoneStats) * m,
                                     hipMemcpyDeviceToHost) != hipSuccess)))
                rc = NUPPN_ERR_HIP;
        }
    }
    (void) hipFree(d);
    return rc;
}

// a line of the autotuner's cache file: the network and device it was tuned for, the choice
struct TuneKey {
    int nsp, nre, nnz, ncu;
    bool operator==(const TuneKey &o) const {
        return nsp == o.nsp && nre == o.nre && nnz == o.nnz && ncu == o.ncu;
    }
};

static bool tune_cache_read(const char *path, const TuneKey &k, int *nb, int *tz, int *tw) {
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) return false;
    TuneKey r;
    int a, b, d;
    bool hit = false;
    while (fscanf(f, "%d %d %d %d %d %d %d", &r.nsp, &r.nre, &r.nnz, &r.ncu, &a, &b, &d) == 7)
        if (r == k) {
            *nb = a;
            *tz = b;
            *tw = d;
            hit = true;
        }
    fclose(f);
    return hit;
}

int nuppn_autotune(nuppn_ctx *c, int nz, const double *yps, const double *t9_0,
                   const double *t9_1, const double *rho_0, const double *rho_1,
                   const double *dtwant, const char *cache_path, nuppn_tune_result *res) {
    if (!c || nz <= 0 || !yps || !t9_0 || !t9_1 || !rho_0 || !rho_1 || !dtwant || !res)
        return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    // (the band form has one shape: nothing to choose)
    if (!c->bnb || c->bnb == kBandForm || !batched_method_enabled(c)) return NUPPN_ERR_UNSUPPORTED;
    memset(res, 0, sizeof(*res));
    const TuneKey key{c->net.nsp, c->net.nre, c->sym.nnzA, c->ncu};
    int nb = 0, tz = 0, tw = 0;
    auto apply = [&](int nb_, int tz_, int tw_) -> int {
        if (nb_ != c->bnb) {
            free_batched(c);
            if (setup_batched(c, nb_) != 0 || c->bnb != nb_) return NUPPN_ERR_HIP;
        }
        nuppn_default_batch_opts(&c->tuned_opts);
        c->tuned_opts.team_zones = tz_;
        c->tuned_opts.team_waves = tw_;
        c->tuned = 1;
        return NUPPN_OK;
    };
    if (tune_cache_read(cache_path, key, &nb, &tz, &tw)) {
        const int rc = apply(nb, tz, tw);
        res->block_size = nb;
        res->team_zones = tz;
        res->team_waves = tw;
        res->from_cache = 1;
        res->consistent = 1;
        return rc;
    }
    // stage the sample once; every candidate integrates a fresh copy of it
    if (ensure_zone_buffers(c, nz)) return NUPPN_ERR_HIP;
    const size_t nsp = c->net.nsp;
    double *d_y0 = nullptr;
    HIP_OK(hipMalloc((void **) &d_y0, sizeof(double) * nsp * nz));
    std::unique_ptr<double, void (*)(double *)> guard(d_y0, [](double *p) { (void) hipFree(p); });
    HIP_OK(hipMemcpy(d_y0, yps, sizeof(double) * nsp * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_t9_0, t9_0, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_t9_1, t9_1, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_0, rho_0, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_rho_1, rho_1, sizeof(double) * nz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->d_dt, dtwant, sizeof(double) * nz, hipMemcpyHostToDevice));
    std::vector<double> first(nsp * nz), cur(nsp * nz);
    const int team_choices[3][2] = {{-1, 0}, {0, 0}, {2 * c->ncu, 2}};
    const int nb0 = c->bnb;
    double best = 1e300;
    res->consistent = 1;
    for (int nbc : {32, 64}) {
        if (nbc != c->bnb) {
            BlockSym probe;
            if (build_block_symbolic(c->net, c->sym, nbc, probe) != 0) continue;
            free_batched(c);
            if (setup_batched(c, nbc) != 0 || c->bnb != nbc) {
                free_batched(c);
                if (setup_batched(c, nb0) != 0) return NUPPN_ERR_HIP;
                continue;
            }
        }
        for (const auto &tc : team_choices) {
            if (res->ncand >= 8) break;
            nuppn_batch_opts o;
            nuppn_default_batch_opts(&o);
            o.team_zones = tc[0];
            o.team_waves = tc[1];
            HIP_OK(hipMemcpy(c->d_yps, d_y0, sizeof(double) * nsp * nz, hipMemcpyDeviceToDevice));
            HIP_OK(hipDeviceSynchronize());
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = run_batch(c, nz, c->d_yps, c->d_t9_0, c->d_t9_1, c->d_rho_0,
                                     c->d_rho_1, c->d_ye, c->d_dt, c->d_ierr, nullptr, &o, nullptr);
            HIP_OK(hipDeviceSynchronize());
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
            if (rc != NUPPN_OK) return rc;
            HIP_OK(hipMemcpy(cur.data(), c->d_yps, sizeof(double) * nsp * nz,
                             hipMemcpyDeviceToHost));
            if (res->ncand == 0) first = cur;
            else if (memcmp(first.data(), cur.data(), sizeof(double) * nsp * nz) != 0)
                res->consistent = 0;
            const int k = res->ncand++;
            res->ms[k] = ms;
            res->cand_nb[k] = nbc;
            res->cand_tz[k] = o.team_zones;
            res->cand_tw[k] = o.team_waves;
            if (ms < best) {
                best = ms;
                nb = nbc;
                tz = o.team_zones;
                tw = o.team_waves;
            }
        }
    }
    const int rc = apply(nb, tz, tw);
    if (rc != NUPPN_OK) return rc;
    res->block_size = nb;
    res->team_zones = tz;
    res->team_waves = tw;
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%d %d %d %d %d %d %d\n", key.nsp, key.nre, key.nnz, key.ncu, nb, tz, tw);
            fclose(f);
        }
    }
    return NUPPN_OK;
}

int nuppn_set_block_form(nuppn_ctx *c, int nb) {
    if (!c || (nb != 32 && nb != 64 && nb != kBandForm)) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    BlockSym probe;
    if (build_block_symbolic(c->net, c->sym, nb == kBandForm ? 0 : nb, probe) != 0)
        return NUPPN_ERR_UNSUPPORTED;
    if (nb == c->bnb) return NUPPN_OK;
    const int nb0 = c->bnb;
    free_batched(c);
    if (setup_batched(c, nb) != 0 || c->bnb != nb) {
        free_batched(c);
        if (nb0) (void) setup_batched(c, nb0);
        return NUPPN_ERR_HIP;
    }
    c->tuned = 0;
    return NUPPN_OK;
}

int nuppn_batch_block_info(nuppn_ctx *c, int *info, int ninfo) {
    if (!c || !info || ninfo <= 0) return NUPPN_ERR_ARG;
    const BlockSym &b = c->bsym;
    const int vals[7] = {c->bnb, b.nblk, b.nbd, b.ni, b.bandwidth, c->bmax_slots, (int) c->blds64};
    for (int i = 0; i < ninfo && i < 7; ++i) info[i] = vals[i];
    return c->bnb ? NUPPN_OK : NUPPN_ERR_UNSUPPORTED;
}

int nuppn_integrate_batch_mixed_dev(nuppn_ctx *c, int nz, double *yps, const double *t9_0,
                                    const double *t9_1, const double *rho_0, const double *rho_1,
                                    double *ye, const double *dtwant, int *ierr,
                                    nuppn_zone_stats *stats, const nuppn_batch_opts *opts,
                                    void *stream) {
    if (!c || nz <= 0 || !yps || !t9_0 || !t9_1 || !rho_0 || !rho_1 || !ye || !dtwant || !ierr)
        return NUPPN_ERR_ARG;
    if (!c->bnb || !batched_method_enabled(c)) return NUPPN_ERR_UNSUPPORTED;
    nuppn_batch_opts o;
    if (mixed_opts(opts, &o, c)) return NUPPN_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    return run_batch(c, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, &o,
                     (hipStream_t) stream);
}

int nuppn_integrate_batch_mixed(nuppn_ctx *c, int nz, double *yps, const double *t9_0,
                                const double *t9_1, const double *rho_0, const double *rho_1,
                                double *ye, const double *dtwant, int *ierr,
                                nuppn_zone_stats *stats, const nuppn_batch_opts *opts) {
    if (!c || nz <= 0 || !yps || !t9_0 || !t9_1 || !rho_0 || !rho_1 || !ye || !dtwant || !ierr)
        return NUPPN_ERR_ARG;
    if (!c->bnb || !batched_method_enabled(c)) return NUPPN_ERR_UNSUPPORTED;
    nuppn_batch_opts o;
    if (mixed_opts(opts, &o, c)) return NUPPN_ERR_ARG;
    return host_batch(c, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, &o);
}

int nuppn_integrate_zone(nuppn_ctx *c, double *yps, double t9_0, double t9_1, double rho_0,
                         double rho_1, double *ye, double dtwant, int *ierr,
                         nuppn_zone_stats *stats) {
    if (!c || !yps || !ye || !ierr) return NUPPN_ERR_ARG;
    return nuppn_integrate_batch(c, 1, yps, &t9_0, &t9_1, &rho_0, &rho_1, ye, &dtwant, ierr,
                                 stats);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// SuperLU link-level drop-ins: the column order is SuperLU's (COLAMD + etree post-order), the
// numeric factorisation and the solves run on the device.

namespace {

struct DropinFactors {
    int n = 0;
    int *d_Acp = nullptr, *d_Ari = nullptr, *d_perm_c = nullptr, *d_iperm_c = nullptr;
    int *d_perm_r = nullptr, *d_Lp = nullptr, *d_info = nullptr;
    double *d_val = nullptr, *d_b = nullptr;
    LUStore st{};
    size_t lds = 0;
    int nb_cap = 0;
    int cap_n = 0, cap_nnz = 0;   // the largest system the buffers hold
};

void dropin_free(DropinFactors *f) {
    if (!f) return;
    for (void *p : {(void *) f->d_Acp, (void *) f->d_Ari, (void *) f->d_perm_c,
                    (void *) f->d_iperm_c, (void *) f->d_perm_r, (void *) f->d_Lp,
                    (void *) f->d_info, (void *) f->d_val, (void *) f->d_b, (void *) f->st.Li,
                    (void *) f->st.Lx, (void *) f->st.Ui, (void *) f->st.Ux, (void *) f->st.Up})
        dev_free(p);
    delete f;
}

// factor a 0-based CSC matrix (rows sorted or not; COLAMD handles both); returns SuperLU info,
// or a negative library error.  reuse: buffers of an earlier call, taken when they hold this
// system (n and nnz within their capacities; *out is then reuse), otherwise new ones are
// allocated (and reuse is left to the caller).  On a negative error *out is not set and only
// new buffers are freed.
int dropin_factor(int n, const int *colptr, const int *rowind, const double *val,
                  DropinFactors **out, DropinFactors *reuse = nullptr) {
    std::vector<int> perm_c;
    if (superlu_column_order(n, colptr, rowind, perm_c) != 0) return NUPPN_ERR_ARG;
    std::vector<int> iperm(n);
    for (int i = 0; i < n; ++i) iperm[perm_c[i]] = i;
    const int nnz = colptr[n];
    const bool take = reuse && reuse->cap_n >= n && reuse->cap_nnz >= nnz;
    DropinFactors *f = take ? reuse : new DropinFactors;
    auto fail = [&](int code) {
        if (f != reuse) dropin_free(f);
        return code;
    };
    f->n = n;
    if (take) {
        const hipMemcpyKind h2d = hipMemcpyHostToDevice;
        if (hipMemcpy(f->d_Acp, colptr, sizeof(int) * (n + 1), h2d) != hipSuccess ||
            hipMemcpy(f->d_Ari, rowind, sizeof(int) * nnz, h2d) != hipSuccess ||
            hipMemcpy(f->d_val, val, sizeof(double) * nnz, h2d) != hipSuccess ||
            hipMemcpy(f->d_perm_c, perm_c.data(), sizeof(int) * n, h2d) != hipSuccess ||
            hipMemcpy(f->d_iperm_c, iperm.data(), sizeof(int) * n, h2d) != hipSuccess)
            return fail(NUPPN_ERR_HIP);
    } else {
        const size_t capL = (size_t) n * (n - 1) / 2 + 1, capU = (size_t) n * (n + 1) / 2 + 1;
        if (dev_upload(&f->d_Acp, colptr, n + 1) || dev_upload(&f->d_Ari, rowind, nnz) ||
            dev_upload(&f->d_val, val, nnz) || dev_upload(&f->d_perm_c, perm_c) ||
            dev_upload(&f->d_iperm_c, iperm) ||
            dev_upload(&f->d_perm_r, (const int *) nullptr, n) ||
            dev_upload(&f->d_Lp, (const int *) nullptr, n + 1) ||
            dev_upload(&f->d_info, (const int *) nullptr, 1) ||
            dev_upload(&f->st.Li, (const int *) nullptr, capL) ||
            dev_upload(&f->st.Lx, (const double *) nullptr, capL) ||
            dev_upload(&f->st.Ui, (const int *) nullptr, capU) ||
            dev_upload(&f->st.Ux, (const double *) nullptr, capU) ||
            dev_upload(&f->st.Up, (const int *) nullptr, n + 1))
            return fail(NUPPN_ERR_HIP);
        f->cap_n = n;
        f->cap_nnz = nnz;
    }
    f->lds = lds_bytes(n);
    if (f->lds > 64 * 1024)
        (void) hipFuncSetAttribute((const void *) lu_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int) f->lds);
    if (ensure_stack((const void *) lu_kernel)) return fail(NUPPN_ERR_HIP);
    hipLaunchKernelGGL(lu_kernel, dim3(1), dim3(64), f->lds, 0, n, f->d_Acp, f->d_Ari, f->d_val,
                       f->d_perm_c, f->d_iperm_c, 0, (double *) nullptr, f->st, f->d_perm_r,
                       f->d_Lp, f->d_info, 1);
    int info = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(&info, f->d_info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NUPPN_ERR_HIP);
    *out = f;
    return info;
}

// slu_solve's buffers, per device, kept from call to call (until the process ends)
std::mutex g_slu_mu;
std::vector<DropinFactors *> g_slu_ws;

int dropin_solve(DropinFactors *f, int nrhs, double *b, int ldb) {
    const int n = f->n;
    if (nrhs <= 0) return 0;
    if (nrhs > f->nb_cap) {   // room for nrhs vectors of the largest system the buffers hold
        dev_free(f->d_b);
        f->d_b = nullptr;
        f->nb_cap = 0;
        if (dev_upload(&f->d_b, (const double *) nullptr, (size_t) f->cap_n * nrhs))
            return NUPPN_ERR_HIP;
        f->nb_cap = nrhs;
    }
    for (int r = 0; r < nrhs; ++r)
        if (hipMemcpy(f->d_b + (size_t) r * n, b + (size_t) r * ldb, sizeof(double) * n,
                      hipMemcpyHostToDevice) != hipSuccess)
            return NUPPN_ERR_HIP;
    hipLaunchKernelGGL(lu_kernel, dim3(1), dim3(64), f->lds, 0, n, f->d_Acp, f->d_Ari, f->d_val,
                       f->d_perm_c, f->d_iperm_c, nrhs, f->d_b, f->st, f->d_perm_r, f->d_Lp,
                       f->d_info, 0);
    if (hipGetLastError() != hipSuccess) return NUPPN_ERR_HIP;
    for (int r = 0; r < nrhs; ++r)
        if (hipMemcpy(b + (size_t) r * ldb, f->d_b + (size_t) r * n, sizeof(double) * n,
                      hipMemcpyDeviceToHost) != hipSuccess)
            return NUPPN_ERR_HIP;
    return 0;
}

}  // namespace

extern "C" {

void slu_solve(int n, int nnz, int isrow, int shiftindex, double val[], int colind[],
               int rowptr[], double rhs[], int *ierr) {
    if (!ierr) return;
    if (n <= 0 || nnz < 0 || !val || !colind || !rowptr || !rhs) {
        *ierr = -1;
        return;
    }
    if (shiftindex) {   // in place, as slu_f90.c:11-15 (the Fortran caller restores them)
        for (int i = 0; i <= n; ++i) rowptr[i] -= 1;
        for (int i = 0; i < nnz; ++i) colind[i] -= 1;
    }
    std::vector<int> cp(n + 1, 0), ri(nnz);
    std::vector<double> cv(nnz);
    if (isrow) {   // dCompRow_to_CompCol (dutil.c:154): rows ascending within each column
        for (int p = 0; p < nnz; ++p) cp[colind[p] + 1]++;
        for (int i = 0; i < n; ++i) cp[i + 1] += cp[i];
        std::vector<int> pos(cp.begin(), cp.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
                const int q = pos[colind[p]]++;
                ri[q] = r;
                cv[q] = val[p];
            }
    } else {
        std::copy(rowptr, rowptr + n + 1, cp.begin());
        std::copy(colind, colind + nnz, ri.begin());
        std::copy(val, val + nnz, cv.begin());
    }
    // the device buffers of the last call on this device are reused while the system fits them
    // (a larger one replaces them); calls are serialised
    std::lock_guard<std::mutex> lock(g_slu_mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
        *ierr = NUPPN_ERR_HIP;
        return;
    }
    if ((int) g_slu_ws.size() <= dev) g_slu_ws.resize(dev + 1, nullptr);
    DropinFactors *ws = g_slu_ws[dev], *f = nullptr;
    const int info = dropin_factor(n, cp.data(), ri.data(), cv.data(), &f, ws);
    if (f && f != ws) {
        dropin_free(ws);
        g_slu_ws[dev] = f;
    }
    if (info != 0) {
        *ierr = info;
        return;
    }
    *ierr = dropin_solve(f, 1, rhs, n);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Allocation-free slu_solve: every device buffer of the call (the CSC pattern and values, the
// column order, the factors at dense capacity, the right-hand side) is carved out of one
// file-scope device array.  Its device address differs per GPU and is resolved once per device
// with hipGetSymbolAddress; host code uses only that address, never the array's name.

constexpr size_t kNoallocScratchBytes = 48ull << 20;   // 48 MiB per GPU, held until exit
static __device__ __align__(256) char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_noalloc_mu;            // one call at a time per process (the array is shared)
std::vector<char *> g_noalloc_ptr;  // device address of `scratch`, per device

int noalloc_scratch(char **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return NUPPN_ERR_HIP;
    if ((size_t) dev >= g_noalloc_ptr.size()) g_noalloc_ptr.resize(dev + 1, nullptr);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "nuppn_hip: hipGetSymbolAddress(scratch) failed: %s\n",
                    hipGetErrorString(e));
            return NUPPN_ERR_HIP;
        }
        g_noalloc_ptr[dev] = (char *) p;
    }
    *out = g_noalloc_ptr[dev];
    return NUPPN_OK;
}

// byte offsets of the call's buffers in the scratch array (each 256-aligned)
struct NoallocLayout {
    size_t cp, ri, val, pc, ipc, pr, lp, info, b, li, lx, ui, ux, up, bytes;
};

NoallocLayout noalloc_layout(int n, int nnz) {
    const size_t capL = (size_t) n * (n - 1) / 2 + 1, capU = (size_t) n * (n + 1) / 2 + 1;
    const size_t len[14] = {4ull * (n + 1), 4ull * nnz, 8ull * nnz, 4ull * n, 4ull * n, 4ull * n,
                            4ull * (n + 1), 16, 8ull * n, 4 * capL, 8 * capL, 4 * capU,
                            8 * capU, 4ull * (n + 1)};
    size_t off[14], o = 0;
    for (int i = 0; i < 14; ++i) {
        off[i] = o;
        o += (len[i] + 255) & ~(size_t) 255;
    }
    return {off[0], off[1], off[2], off[3], off[4], off[5], off[6], off[7], off[8], off[9],
            off[10], off[11], off[12], off[13], o};
}

}  // namespace

extern "C" {

void slu_solve_noalloc(int n, int nnz, int isrow, int shiftindex, double val[], int colind[],
                       int rowptr[], double rhs[], int *ierr) {
    if (!ierr) return;
    if (n <= 0 || nnz < 0 || !val || !colind || !rowptr || !rhs) {
        *ierr = -1;
        return;
    }
    const NoallocLayout L = noalloc_layout(n, nnz);
    if (L.bytes > kNoallocScratchBytes || lds_bytes(n) > 160 * 1024) {
        fprintf(stderr, "nuppn_hip: slu_solve_noalloc: n = %d, nnz = %d needs %zu bytes of "
                "device scratch, more than its fixed %zu (n up to about 2,000)\n", n, nnz,
                L.bytes, kNoallocScratchBytes);
        *ierr = NUPPN_ERR_UNSUPPORTED;
        return;
    }
    if (shiftindex) {   // in place, as slu_f90.c:11-15 (the Fortran caller restores them)
        for (int i = 0; i <= n; ++i) rowptr[i] -= 1;
        for (int i = 0; i < nnz; ++i) colind[i] -= 1;
    }
    std::vector<int> cp(n + 1, 0), ri(nnz);
    std::vector<double> cv(nnz);
    if (isrow) {   // dCompRow_to_CompCol (dutil.c:154): rows ascending within each column
        for (int p = 0; p < nnz; ++p) cp[colind[p] + 1]++;
        for (int i = 0; i < n; ++i) cp[i + 1] += cp[i];
        std::vector<int> pos(cp.begin(), cp.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
                const int q = pos[colind[p]]++;
                ri[q] = r;
                cv[q] = val[p];
            }
    } else {
        std::copy(rowptr, rowptr + n + 1, cp.begin());
        std::copy(colind, colind + nnz, ri.begin());
        std::copy(val, val + nnz, cv.begin());
    }
    std::vector<int> perm_c;
    if (superlu_column_order(n, cp.data(), ri.data(), perm_c) != 0) {
        *ierr = -1;
        return;
    }
    std::vector<int> iperm(n);
    for (int i = 0; i < n; ++i) iperm[perm_c[i]] = i;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    char *sp = nullptr;
    if (noalloc_scratch(&sp)) {
        *ierr = NUPPN_ERR_HIP;
        return;
    }
    auto at = [&](size_t off) { return (void *) (sp + off); };
    const size_t lds = lds_bytes(n);
    if ((lds > 64 * 1024 &&
         hipFuncSetAttribute((const void *) lu_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int) lds) != hipSuccess) ||
        hipMemcpy(at(L.cp), cp.data(), 4ull * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(at(L.ri), ri.data(), 4ull * nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        (nnz && hipMemcpy(at(L.val), cv.data(), 8ull * nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(at(L.pc), perm_c.data(), 4ull * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(at(L.ipc), iperm.data(), 4ull * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(at(L.b), rhs, 8ull * n, hipMemcpyHostToDevice) != hipSuccess) {
        *ierr = NUPPN_ERR_HIP;
        return;
    }
    LUStore st;
    st.Li = (int *) at(L.li);
    st.Lx = (double *) at(L.lx);
    st.Ui = (int *) at(L.ui);
    st.Ux = (double *) at(L.ux);
    st.Up = (int *) at(L.up);
    hipLaunchKernelGGL(lu_kernel, dim3(1), dim3(64), lds, 0, n, (const int *) at(L.cp),
                       (const int *) at(L.ri), (const double *) at(L.val), (const int *) at(L.pc),
                       (const int *) at(L.ipc), 1, (double *) at(L.b), st, (int *) at(L.pr),
                       (int *) at(L.lp), (int *) at(L.info), 1);
    int info = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(&info, at(L.info), sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
        *ierr = NUPPN_ERR_HIP;
        return;
    }
    if (info == 0 && hipMemcpy(rhs, at(L.b), 8ull * n, hipMemcpyDeviceToHost) != hipSuccess) {
        *ierr = NUPPN_ERR_HIP;
        return;
    }
    *ierr = info;
}

void c_fortran_dgssv_(int *iopt, int *n, int *nnz, int *nrhs, double *values, int *rowind,
                      int *colptr, double *b, int *ldb, long long *f_factors, int *info) {
    if (!iopt || !info) return;
    if (*iopt == 1) {   // factor (1-based CSC copied to 0-based, c_fortran_dgssv.c:90-93)
        const int nn = *n, nz = *nnz;
        std::vector<int> cp(nn + 1), ri(nz);
        for (int i = 0; i <= nn; ++i) cp[i] = colptr[i] - 1;
        for (int i = 0; i < nz; ++i) ri[i] = rowind[i] - 1;
        DropinFactors *f = nullptr;
        *info = dropin_factor(nn, cp.data(), ri.data(), values, &f);
        *f_factors = (long long) (uintptr_t) f;
    } else if (*iopt == 2) {
        DropinFactors *f = (DropinFactors *) (uintptr_t) *f_factors;
        *info = f ? dropin_solve(f, *nrhs, b, *ldb) : -1;
    } else if (*iopt == 3) {
        dropin_free((DropinFactors *) (uintptr_t) *f_factors);
        *f_factors = 0;
        *info = 0;
    } else {
        // the reference prints and exits(-1) (c_fortran_dgssv.c:186-188); errors stay codes here
        *info = -1;
    }
}

}  // extern "C"
