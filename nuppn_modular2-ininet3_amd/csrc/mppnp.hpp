// The mppnp cycle around the integration, on the device (frames/mppnp/source/mppnp.F90:637-781):
// zone selection by temperature, the gather of the selected zones into a batch and the scatter
// of its results with the per-zone mass check and renormalisation, and the diffusive mixing of
// every species over the zones (mixing.F90) followed by the 2-d mass check, renormalisation and
// clip.  Every kernel follows its reference routine's operation order, so that the results equal
// a sequential evaluation bit for bit (given the same D_lagr: that one goes through log / exp).
#pragma once
#include "integrate.hpp"
#include "zone.hpp"

namespace nuppn {

// zoneselect (mppnp.F90:1080-1097): the zones with T9 above the threshold, in zone order
// (one wavefront: a ballot compaction per 64 zones)
__global__ void __launch_bounds__(64) mp_select_kernel(int m, const double *t9, double thr,
                                                       int *zidx, int *count) {
    const int l = lane_id();
    int base = 0;
    for (int j0 = 0; j0 < m; j0 += WAVE) {
        const int j = j0 + l;
        const bool on = j < m && t9[j] > thr;
        const unsigned long long mask = __ballot(on);
        const int pos = __popcll(mask & ((1ull << l) - 1ull));
        if (on) zidx[base + pos] = j;
        base += __popcll(mask);
    }
    if (l == 0) *count = base;
}

// the selected zones' inputs as a batch: yps rows, T9 and rho at both ends of the cycle (constant
// over it, mppnp.F90:722), dtwant = dzeit; ierr and statistics of every zone cleared
__global__ void __launch_bounds__(256) mp_gather_kernel(int m, int nsp, const int *zidx,
                                                        const int *count, const double *yps,
                                                        const double *t9, const double *rho,
                                                        double dzeit, double *yb, double *t9b,
                                                        double *rhob, double *dtb, double *yeb,
                                                        int *ierr, ZoneStats *stats) {
    const int n = *count;
    const size_t tot = (size_t) n * nsp;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < tot;
         i += (size_t) gridDim.x * blockDim.x) {
        const size_t k = i / nsp, s = i % nsp;
        yb[i] = yps[(size_t) zidx[k] * nsp + s];
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        if (k < n) {
            const int z = zidx[k];
            t9b[k] = t9[z];
            rhob[k] = rho[z];
            dtb[k] = dzeit;
            yeb[k] = 0.5;
        }
        ierr[k] = 0;
        if (stats) stats[k] = ZoneStats{};
    }
}

// ksum (utils.F90:670-684): Kahan summation in index order, every lane folding the LDS copy
__device__ inline double ksum_lds(NUPPN_LDS const double *a, int n) {
    double c = 0.0, s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double y = a[i] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    return s;
}

// |1 - sum(y)| (check_1d_mass_cons, utils.F90:387-397) and, when `renorm` says so, renorm_1d
// (utils.F90:433-444: y / ksum(y)) of one zone's row on one wavefront; y staged in LDS
__device__ inline double mp_check_renorm(const double *y, double *out, int nsp, double limit,
                                         bool strict, NUPPN_LDS double *buf) {
    for (int i = lane_id(); i < nsp; i += WAVE) buf[i] = y[i];
    wave_sync();
    const double res = fabs(1.0 - seq_sum(buf, nsp));
    const bool renorm = strict ? res > limit : res >= limit;
    if (renorm) {
        const double s = ksum_lds(buf, nsp);
        for (int i = lane_id(); i < nsp; i += WAVE) out[i] = buf[i] / s;
    } else if (out != y) {
        for (int i = lane_id(); i < nsp; i += WAVE) out[i] = buf[i];
    }
    wave_sync();
    return res;
}

// each integrated zone: the mass check with renorm_1d above 5e-7 (mppnp.F90:737-738), the
// abundances back into the zone's row, its error code and statistics
__global__ void __launch_bounds__(64) mp_scatter_kernel(int nsp, const int *zidx, const int *count,
                                                        const double *yb, double *yps,
                                                        const int *ierrb, const ZoneStats *stb,
                                                        int *ierr, ZoneStats *stats) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS double *buf = (NUPPN_LDS double *) smem;
    const int n = *count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int z = zidx[k];
        mp_check_renorm(yb + (size_t) k * nsp, yps + (size_t) z * nsp, nsp, 5.0e-7, true, buf);
        if (lane_id() == 0) {
            ierr[z] = ierrb[k];
            if (stats) stats[z] = stb[k];
        }
    }
}

// mix_calculate_D_Lagrangian (mixing.F90:68-110); dconst ** 2 as the constant dconst * dconst
__global__ void __launch_bounds__(256) mp_dlagr_kernel(int m, const double *rho, const double *rad,
                                                       const double *d_eul, double *dl) {
    const double dconst = 3.0589703419947736e-11;
    const double dc2 = dconst * dconst;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (d_eul[j] > 0.0) {
            double lnrho, lnrad;
            if (j == 0) {
                lnrho = log(rho[j]);
                lnrad = log(rad[j]);
            } else {
                lnrho = (log(rho[j]) + log(rho[j - 1])) / 2.0;
                lnrad = (log(rad[j]) + log(rad[j - 1])) / 2.0;
            }
            v = d_eul[j] * dc2 * exp(2.0 * (2.0 * lnrad + lnrho));
        }
        dl[j] = v;
    }
}

// D(k) after matrixdiffus's limit in row k-1 (mixing.F90:214-224): rows 2..m-1 limit the
// coefficient of their outer interface, each from its own geometry, so every limited value can
// be formed on its own (Fortran index k = 3..m; D(1), D(2) are never limited)
__device__ inline double mp_d_limited(int k, double h, const double *xm, const double *D,
                                      double sig) {
    double d = D[k - 1];
    if (k >= 3) {
        const int j = k - 1;   // the row that limits it
        const double xr = xm[j] - xm[j - 1];
        const double xmm = 0.5 * (xm[j] - xm[j - 2]);
        const double alpha = h / xmm;
        if (sig < d * alpha / xr) d = sig / alpha * xr;
    }
    return d;
}

// d_reduce(k) of matrixdiffus_mesa (mixing.F90:270-300, reduce_term 325-340), k = 2..m
__device__ inline double mp_d_reduced(int k, int m, double h, const double *dq, const double *d,
                                      double sig) {
    if (k == 1) return d[0];
    double alpha, onedxr;
    if (k == 2) {
        const double xr = dq[0];
        alpha = h / xr;
        onedxr = 1.0 / xr;
    } else {
        const int j = k - 1;
        const double xr = dq[j - 1];
        const double xmm = 0.5 * (dq[j - 1] + dq[j - 2]);
        alpha = h / xmm;
        onedxr = 1.0 / xr;
    }
    double dc = d[k - 1];
    if (alpha * onedxr * dc > sig) dc = sig / (alpha * onedxr);
    return dc;
}

// the diffusion matrix, one row per thread: matrixdiffus (mixing.F90:180-250, code_source GNV
// / EVL) or matrixdiffus_mesa (254-322, MES); *bad set where the reference stops
__global__ void __launch_bounds__(256) mp_matrix_kernel(int m, double h, const double *xm,
                                                        const double *dq, const double *D,
                                                        double sig, int mesa, double *a,
                                                        double *b, double *c, int *bad) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
        const int j = r + 1;   // Fortran row
        double aj, bj, cj;
        if (!mesa) {
            if (j == 1) {
                const double xr = xm[1] - xm[0];
                double alpha = D[1] * h / (xr * xr);
                if (sig < alpha) alpha = sig;
                aj = 0.0;
                bj = 1.0 + alpha;
                cj = -1.0 * alpha;
            } else if (j < m) {
                const double xl = xm[j - 1] - xm[j - 2];
                const double xr = xm[j] - xm[j - 1];
                const double xmm = 0.5 * (xm[j] - xm[j - 2]);
                const double alpha = h / xmm;
                const double dj = mp_d_limited(j, h, xm, D, sig);
                const double dj1 = mp_d_limited(j + 1, h, xm, D, sig);
                aj = -1.0 * dj * alpha / xl;
                bj = 1.0 + (dj1 / xr + dj / xl) * alpha;
                cj = -1.0 * dj1 * alpha / xr;
                if (xl == 0.0 || xr == 0.0) *bad = 1;
            } else {
                const double xl = xm[m - 1] - xm[m - 2];
                const double alpha = mp_d_limited(m, h, xm, D, sig) * h / (xl * xl);
                aj = -1.0 * alpha;
                bj = 1.0 + alpha;
                cj = 0.0;
                if (xl == 0.0) *bad = 1;
            }
        } else {
            if (j == 1) {
                const double xr = dq[0];
                const double alpha = h / xr;
                const double term_c = 1.0 / xr;
                const double d2 = mp_d_reduced(2, m, h, dq, D, sig);
                aj = 0.0;
                bj = 1.0 + term_c * alpha * d2;
                cj = -1.0 * term_c * alpha * d2;
                if (!(dq[0] > 0.0)) *bad = 1;
            } else if (j < m) {
                const double xl = dq[j - 2];
                const double xr = dq[j - 1];
                const double xmm = 0.5 * (dq[j - 1] + dq[j - 2]);
                const double alpha = h / xmm;
                const double term_a = 1.0 / xl;
                const double term_c = 1.0 / xr;
                const double dj = mp_d_reduced(j, m, h, dq, D, sig);
                const double dj1 = mp_d_reduced(j + 1, m, h, dq, D, sig);
                const double term_b = term_c * dj1 + term_a * dj;
                aj = -1.0 * term_a * dj * alpha;
                bj = 1.0 + term_b * alpha;
                cj = -1.0 * term_c * dj1 * alpha;
            } else {
                const double xl = dq[m - 1];
                const double alpha = h / (xl * xl);
                aj = -1.0 * alpha * D[m - 1];
                bj = 1.0 + alpha * D[m - 1];
                cj = 0.0;
            }
        }
        a[r] = aj;
        b[r] = bj;
        c[r] = cj;
    }
}

// tridag's elimination (mixing.F90:140-176) is the same for every species (one matrix): bet(j)
// and gam(j) once, by one thread; *bad set at a zero pivot, where the reference stops
__global__ void mp_thomas_kernel(int m, const double *a, const double *b, const double *c,
                                 double *bet, double *gam, int *bad) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (b[0] == 0.0) {
        *bad = 1;
        return;
    }
    double bt = b[0];
    bet[0] = bt;
    gam[0] = 0.0;
    for (int j = 1; j < m; ++j) {
        const double g = c[j - 1] / bt;
        bt = b[j] - a[j] * g;
        if (bt == 0.0) *bad = 1;
        gam[j] = g;
        bet[j] = bt;
    }
}

// tridag's substitutions, one considered species per thread (mixdiff08's loop over species
// 1..NSP-1, mixing.F90:118-130), in place on yps[m][nsp] (rows of neighbouring species adjacent)
__global__ void __launch_bounds__(256) mp_species_kernel(int m, int nsp, int igamma,
                                                         double *yps, const double *a,
                                                         const double *bet, const double *gam,
                                                         const int *bad) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsp - 1 || s == igamma || *bad) return;
    NUPPN_GLB double *y = gptr(yps);
    double u = y[s] / bet[0];
    y[s] = u;
    for (int j = 1; j < m; ++j) {
        u = (y[(size_t) j * nsp + s] - a[j] * u) / bet[j];
        y[(size_t) j * nsp + s] = u;
    }
    for (int j = m - 2; j >= 0; --j) {
        u = y[(size_t) j * nsp + s] - gam[j + 1] * u;
        y[(size_t) j * nsp + s] = u;
    }
}

// the 2-d mass check after mixing (check_2d_mass_cons, utils.F90:409-422): |1 - sum| per zone
__global__ void __launch_bounds__(64) mp_residual_kernel(int m, int nsp, const double *yps,
                                                         double *res) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS double *buf = (NUPPN_LDS double *) smem;
    for (int z = blockIdx.x; z < m; z += gridDim.x) {
        const double *y = yps + (size_t) z * nsp;
        for (int i = lane_id(); i < nsp; i += WAVE) buf[i] = y[i];
        wave_sync();
        const double r = fabs(1.0 - seq_sum(buf, nsp));
        if (lane_id() == 0) res[z] = r;
        wave_sync();
    }
}

// the largest residual (one wavefront; max is exact in any order)
__global__ void __launch_bounds__(64) mp_max_kernel(int m, const double *res, double *out) {
    double v = 0.0;
    for (int z = lane_id(); z < m; z += WAVE) v = fmax(v, res[z]);
    v = wave_max(v);
    if (lane_id() == 0) *out = v;
}

// renorm_2d when the largest residual reaches SUM_TOL (mppnp.F90:768-773; every zone divided by
// its ksum), then the clip max(yps, 1e-99) (mppnp.F90:781)
__global__ void __launch_bounds__(64) mp_renorm_clip_kernel(int m, int nsp, double *yps,
                                                            const double *maxres) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    NUPPN_LDS double *buf = (NUPPN_LDS double *) smem;
    const bool renorm = *maxres >= 1.0e-6;
    for (int z = blockIdx.x; z < m; z += gridDim.x) {
        double *y = yps + (size_t) z * nsp;
        if (renorm) {
            for (int i = lane_id(); i < nsp; i += WAVE) buf[i] = y[i];
            wave_sync();
            const double s = ksum_lds(buf, nsp);
            for (int i = lane_id(); i < nsp; i += WAVE) y[i] = fmax(buf[i] / s, 1.0e-99);
            wave_sync();
        } else {
            for (int i = lane_id(); i < nsp; i += WAVE) y[i] = fmax(y[i], 1.0e-99);
        }
    }
}

}  // namespace nuppn
