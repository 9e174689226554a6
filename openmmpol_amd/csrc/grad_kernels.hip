// Geometric gradient of the polarization energy (polelec_geomgrad,
// mod_geomgrad.F90:113-273, and rotation_geomgrad, rotate_multipoles.F90:3-230),
// FP64 throughout.
//
// Below 5,000 polarizable sites, two pair sweeps, one thread per target, sources
// staged through LDS in tiles of 256 and split over gridDim.y (from 5,000 the
// static part of both runs once per pair, k_gstat_both, and the dipole-dipole
// part once per unordered pair, k_gpol_dd_sym):
//   polarizable target j   sum_b w_b Egrd_ab of the static multipoles, with
//                          w = 1/2 (mu_D + mu_P) (AMOEBA) or mu (AMBER), and
//                          1/2 [mu_D,j . Egrd(mu_P) + mu_P,j . Egrd(mu_D)] of
//                          the other dipoles (mu_j . Egrd(mu) for AMBER).  The
//                          contraction is done per pair, so a thread keeps three
//                          accumulators instead of two 6-component gradients.
//   MM target m            field E, field gradient Egrd (kept for the rotation
//                          term) and Q_m : EHes of the dipoles
//                          nu = 1/2 (mu_P + mu_D) (AMOEBA) or mu (AMBER).
// Sign conventions are the reference's: Egrd_ab = -dE_b/dr_a, EHes_abc =
// d2E_c/dr_a dr_b, with d = r_target - r_source.  Damped kernels k_n =
// lambda_{2n-1} / r^{2n-1} obey dk_n/dd_a = -(2n-1) k_{n+1} d_a for both damping
// schemes, which is what makes these the derivatives of the damped fields.
// As in the static-field kernel, the sweeps use scale 1 for every pair and the
// finalize kernels add the per-split partial sums in a fixed order and subtract
// (1 - s) times the listed pairs.
// The rotation term differentiates the frame construction with forward-mode
// dual numbers, one thread per (site, frame atom); an assemble kernel gathers
// the per-atom sum through a fixed index list.  No atomics: results are
// bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "grad_math.h"
#include "pair_math.h"
#include "tmv_sym.h"

namespace ommp {


// ------------------------------------------------- polarizable targets
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gpol_static(int nmm, int npol, int chunk, const double4 *__restrict__ ms,
                                                        const double *__restrict__ ma, const double *__restrict__ mq,
                                                        const double4 *__restrict__ ps, const double *__restrict__ pa,
                                                        const int *__restrict__ polar_mm,
                                                        const double *__restrict__ ipd, double *__restrict__ part) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_a[kBlock];
    __shared__ double s_q[LD][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < npol;
    const long n = 3L * npol;
    const double4 pj = valid ? ps[j] : make_double4(0.0, 0.0, 0.0, 1.0);
    const double aj = valid ? pa[j] : 0.39;
    const int jm = valid ? polar_mm[j] : -1;
    double wx = 0.0, wy = 0.0, wz = 0.0;
    if (valid) {
        wx = ipd[3L * j];
        wy = ipd[3L * j + 1];
        wz = ipd[3L * j + 2];
        if (AMOEBA) {
            wx = 0.5 * (wx + ipd[n + 3L * j]);
            wy = 0.5 * (wy + ipd[n + 3L * j + 1]);
            wz = 0.5 * (wz + ipd[n + 3L * j + 2]);
        }
    }
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(nmm, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            s_pos[tid] = ms[base + tid];
            s_a[tid] = ma[base + tid];
#pragma unroll
            for (int c = 0; c < LD; c++) s_q[c][tid] = mq[(long)LD * (base + tid) + c];
        }
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const double4 pi = s_pos[k];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double ir = (base + k) == jm ? 0.0 : fast_rsqrt(r2);
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, ir, pi.w, pj.w, s_a[k], aj, k2, k3, k4, k5);
            double q[LD];
#pragma unroll
            for (int c = 0; c < LD; c++) q[c] = s_q[c][k];
            static_grad_pair<AMOEBA>(dx, dy, dz, k2, k3, k4, k5, q, wx, wy, wz, gx, gy, gz);
        }
    }
    if (valid) {
        double *o = part + (long)blockIdx.y * 3 * npol;
        o[j] = gx;
        o[npol + j] = gy;
        o[2L * npol + j] = gz;
    }
}

// Dipole-dipole part: AMOEBA 1/2 [mu_D,j . Egrd(mu_P,i) + mu_P,j . Egrd(mu_D,i)],
// AMBER mu_j . Egrd(mu_i).
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gpol_dd(int npol, int chunk, const double4 *__restrict__ ps,
                                                    const double *__restrict__ pa, const double *__restrict__ ipd,
                                                    double *__restrict__ part) {
    constexpr int NS = AMOEBA ? 6 : 3;
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_a[kBlock];
    __shared__ double s_m[NS][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < npol;
    const long n = 3L * npol;
    const double4 pj = valid ? ps[j] : make_double4(0.0, 0.0, 0.0, 1.0);
    const double aj = valid ? pa[j] : 0.39;
    // a = mu_D,j (contracted with the field of mu_P), b = mu_P,j (with that of mu_D)
    double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
    if (valid) {
        ax = ipd[3L * j];
        ay = ipd[3L * j + 1];
        az = ipd[3L * j + 2];
        if (AMOEBA) {
            bx = ipd[n + 3L * j];
            by = ipd[n + 3L * j + 1];
            bz = ipd[n + 3L * j + 2];
        }
    }
    const double f = AMOEBA ? 0.5 : 1.0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(npol, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const long i = base + tid;
            s_pos[tid] = ps[i];
            s_a[tid] = pa[i];
            // slots 0-2: mu_P (AMOEBA) / mu (AMBER); 3-5: mu_D
#pragma unroll
            for (int c = 0; c < 3; c++) {
                s_m[c][tid] = AMOEBA ? ipd[n + 3 * i + c] : ipd[3 * i + c];
                if (AMOEBA) s_m[3 + c][tid] = ipd[3 * i + c];
            }
        }
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const double4 pi = s_pos[k];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double ir = (base + k) == j ? 0.0 : fast_rsqrt(r2);
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, ir, pi.w, pj.w, s_a[k], aj, k2, k3, k4, k5);
            dip_grad_pair(dx, dy, dz, k3, k4, ax, ay, az, s_m[0][k], s_m[1][k], s_m[2][k], f, gx, gy, gz);
            if (AMOEBA)
                dip_grad_pair(dx, dy, dz, k3, k4, bx, by, bz, s_m[3][k], s_m[4][k], s_m[5][k], f, gx, gy, gz);
        }
    }
    if (valid) {
        double *o = part + (long)blockIdx.y * 3 * npol;
        o[j] = gx;
        o[npol + j] = gy;
        o[2L * npol + j] = gz;
    }
}

// The same dipole-dipole part with each unordered pair evaluated once, from
// 5,000 polarizable sites: the pair's term is odd in d and symmetric in the two
// dipoles it contracts, so the term at i equals minus the term at j; it is
// formed once and added to j (target side) and subtracted from i (source
// side), over the block pairs of the pair-once matvec (tmv_sym.h: sources of J
// in LDS, source-side sums carried by the DPP wave rotate).  Target-side sums
// of component c go to part_i[p][c][t], source-side ones to part_j[p][c][t].
// Padding sites sit far away with zero dipoles, which add exactly 0.
template <bool AMOEBA>
__global__ __launch_bounds__(kSymBlock) void k_gpol_dd_sym(int npol, const int2 *__restrict__ pairs,
                                                           const double4 *__restrict__ ps,
                                                           const double *__restrict__ pa,
                                                           const double *__restrict__ ipd,
                                                           double *__restrict__ part_i, double *__restrict__ part_j) {
    constexpr int NS = AMOEBA ? 6 : 3;  // mu_P then mu_D (AMOEBA), mu (AMBER)
    __shared__ double4 s_pos[kSymBlock];
    __shared__ double s_a[kSymBlock];
    __shared__ double s_m[NS][kSymBlock];
    __shared__ double s_acc[kSymWaves][3][64];
    const int2 pr = pairs[blockIdx.x];
    const int I = pr.x, J = pr.y;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long n = 3L * npol;
    const double far = 1e15, gap = 1e9, f = AMOEBA ? 0.5 : 1.0;
    {
        const int jj = J * kSymBlock + tid;
        if (jj < npol) {
            s_pos[tid] = ps[jj];
            s_a[tid] = pa[jj];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                s_m[c][tid] = AMOEBA ? ipd[n + 3L * jj + c] : ipd[3L * jj + c];
                if (AMOEBA) s_m[3 + c][tid] = ipd[3L * jj + c];
            }
        } else {
            s_pos[tid] = make_double4(far + gap * (double)tid, far + gap * (double)tid, far + gap * (double)tid, 1.0);
            s_a[tid] = 0.39;
#pragma unroll
            for (int c = 0; c < NS; c++) s_m[c][tid] = 0.0;
        }
    }
    // the thread's target: mu_D (a) and mu_P (b) for AMOEBA, mu (a) for AMBER
    const int i = I * kSymBlock + tid;
    double4 pt;
    double at = 0.39, ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
    if (i < npol) {
        pt = ps[i];
        at = pa[i];
        ax = ipd[3L * i];
        ay = ipd[3L * i + 1];
        az = ipd[3L * i + 2];
        if (AMOEBA) {
            bx = ipd[n + 3L * i];
            by = ipd[n + 3L * i + 1];
            bz = ipd[n + 3L * i + 2];
        }
    } else {
        pt = make_double4(far + gap * (double)tid, far + gap * (double)tid, far + gap * (double)tid, 1.0);
    }
    // term of the pair at the target (d = r_t - r_s)
    auto term = [&](int js, double &gx, double &gy, double &gz) {
        const double4 psr = s_pos[js];
        const double dx = pt.x - psr.x, dy = pt.y - psr.y, dz = pt.z - psr.z;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        double k2, k3, k4, k5;
        damp_k5<AMOEBA>(r2, fast_rsqrt(r2), psr.w, pt.w, s_a[js], at, k2, k3, k4, k5);
        gx = gy = gz = 0.0;
        dip_grad_pair(dx, dy, dz, k3, k4, ax, ay, az, s_m[0][js], s_m[1][js], s_m[2][js], f, gx, gy, gz);
        if (AMOEBA)
            dip_grad_pair(dx, dy, dz, k3, k4, bx, by, bz, s_m[3][js], s_m[4][js], s_m[5][js], f, gx, gy, gz);
    };
    double acc[3] = {0.0, 0.0, 0.0};
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < kSymWaves; k++) {
        const bool mirror = I == J && k < w;  // wave-uniform
        const bool diag = I == J && k == w;
        double jacc[3] = {0.0, 0.0, 0.0};
        if (diag) {
            // one-sided over the other 63 sites of the sub-tile
            for (int st = 1; st < 64; st++) {
                double gx, gy, gz;
                term(k * 64 + ((l + st) & 63), gx, gy, gz);
                acc[0] += gx;
                acc[1] += gy;
                acc[2] += gz;
            }
        } else if (!mirror) {
            for (int st = 0; st < 64; st++) {
                double gx, gy, gz;
                term(k * 64 + ((l + st) & 63), gx, gy, gz);
                acc[0] += gx;
                acc[1] += gy;
                acc[2] += gz;
                jacc[0] -= gx;
                jacc[1] -= gy;
                jacc[2] -= gz;
#pragma unroll
                for (int c = 0; c < 3; c++) jacc[c] = lane_next(jacc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) s_acc[w][c][l] = jacc[c];
        __syncthreads();
        if (tid < 3 * 64) {
            const int c = tid >> 6, jl = tid & 63;
            double sum = s_acc[0][c][jl];
#pragma unroll
            for (int q = 1; q < kSymWaves; q++) sum += s_acc[q][c][jl];
            part_j[((long)blockIdx.x * 3 + c) * kSymBlock + k * 64 + jl] = sum;
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 3; c++) part_i[((long)blockIdx.x * 3 + c) * kSymBlock + tid] = acc[c];
}

// Sum of a site's row (pairs (b, J >= b): target sides) and column (pairs
// (I <= b, b): source sides) block-pair partials into out[c * npol + j], in a
// fixed order: kQ groups of kSymBlock threads per site block, group q taking
// every kQ-th partial, the groups added in order.
__global__ __launch_bounds__(4 * kSymBlock) void k_sym3_reduce(int npol, int nb, const double *__restrict__ part,
                                                               double *__restrict__ out) {
    constexpr int kQ = 4;
    __shared__ double sh[kQ - 1][3][kSymBlock];
    const int q = threadIdx.x / kSymBlock, li = threadIdx.x % kSymBlock;
    const int b = blockIdx.x, j = b * kSymBlock + li;
    const long np = (long)nb * (nb + 1) / 2, stride = 3L * kSymBlock;
    double sum[3] = {0.0, 0.0, 0.0};
    if (j < npol) {
        const long row = (long)b * nb - (long)b * (b - 1) / 2;
        for (int J = b + q; J < nb; J += kQ) {
            const double *p = part + (row + J - b) * stride + li;
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
        const double *part_j = part + np * stride;
        for (int I = q; I <= b; I += kQ) {
            const double *p = part_j + ((long)I * nb - (long)I * (I - 1) / 2 + b - I) * stride + li;
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
    }
    if (q > 0)
#pragma unroll
        for (int c = 0; c < 3; c++) sh[q - 1][c][li] = sum[c];
    __syncthreads();
    if (q != 0 || j >= npol) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double v = sum[c];
#pragma unroll
        for (int g = 0; g < kQ - 1; g++) v += sh[g][c][li];
        out[(long)c * npol + j] = v;
    }
}

// The finalizes take 64 sites per workgroup with kGfQ groups of 64 threads:
// group q sums every kGfQ-th partial and every kGfQ-th listed pair of its
// site, and group 0 adds the group results in a fixed order (one thread per
// site waited on the latency of its whole chain of partials and pairs).
constexpr int kGfQ = 4;
constexpr int kGfSites = kBlock / kGfQ;
static_assert(kGfSites == 64, "one wave of sites per group");

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gpol_finalize(int npol, int nparts, const double *__restrict__ part,
                                                          const double4 *__restrict__ ms,
                                                          const double *__restrict__ ma,
                                                          const double *__restrict__ mq,
                                                          const double4 *__restrict__ ps,
                                                          const double *__restrict__ pa,
                                                          const double *__restrict__ ipd, const int *__restrict__ sp_ri,
                                                          const int *__restrict__ sp_ci,
                                                          const double2 *__restrict__ sp_f,
                                                          const int *__restrict__ pp_ri,
                                                          const int *__restrict__ pp_ci,
                                                          const double *__restrict__ pp_f,
                                                          double *__restrict__ gpol) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double sh[kGfQ - 1][3][kGfSites];
    const int q = threadIdx.x / kGfSites, li = threadIdx.x % kGfSites;
    const int j = blockIdx.x * kGfSites + li;
    const bool valid = j < npol;
    const long n = 3L * npol;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (valid) {
        for (int s = q; s < nparts; s += kGfQ) {
            const double *p = part + (long)s * 3 * npol;
            gx += p[j];
            gy += p[npol + j];
            gz += p[2L * npol + j];
        }
        const double4 pj = ps[j];
        const double aj = pa[j];
        const double dX = ipd[3L * j], dY = ipd[3L * j + 1], dZ = ipd[3L * j + 2];
        const double pX = AMOEBA ? ipd[n + 3L * j] : dX, pY = AMOEBA ? ipd[n + 3L * j + 1] : dY,
                     pZ = AMOEBA ? ipd[n + 3L * j + 2] : dZ;
        // listed static sources: remove (1 - s_p) of the mu_D contraction and
        // (1 - s_d) of the mu_P one (AMBER: both slots hold 1 - s)
        for (int k = sp_ri[j] + q; k < sp_ri[j + 1]; k += kGfQ) {
            const int i = sp_ci[k];
            const double2 f = sp_f[k];
            const double wx = AMOEBA ? 0.5 * (f.y * dX + f.x * pX) : f.y * dX;
            const double wy = AMOEBA ? 0.5 * (f.y * dY + f.x * pY) : f.y * dY;
            const double wz = AMOEBA ? 0.5 * (f.y * dZ + f.x * pZ) : f.y * dZ;
            const double4 pi = ms[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, ma[i], aj, k2, k3, k4, k5);
            double qm[LD];
            for (int c = 0; c < LD; c++) qm[c] = mq[(long)LD * i + c];
            double cx = 0.0, cy = 0.0, cz = 0.0;
            static_grad_pair<AMOEBA>(dx, dy, dz, k2, k3, k4, k5, qm, wx, wy, wz, cx, cy, cz);
            gx -= cx;
            gy -= cy;
            gz -= cz;
        }
        // listed dipole sources: remove (1 - s_u) of the pair
        for (int k = pp_ri[j] + q; k < pp_ri[j + 1]; k += kGfQ) {
            const int i = pp_ci[k];
            const double4 pi = ps[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, pa[i], aj, k2, k3, k4, k5);
            const double f = pp_f[k] * (AMOEBA ? 0.5 : 1.0);
            double cx = 0.0, cy = 0.0, cz = 0.0;
            if (AMOEBA) {
                dip_grad_pair(dx, dy, dz, k3, k4, dX, dY, dZ, ipd[n + 3L * i], ipd[n + 3L * i + 1],
                              ipd[n + 3L * i + 2], f, cx, cy, cz);
                dip_grad_pair(dx, dy, dz, k3, k4, pX, pY, pZ, ipd[3L * i], ipd[3L * i + 1], ipd[3L * i + 2], f, cx,
                              cy, cz);
            } else {
                dip_grad_pair(dx, dy, dz, k3, k4, dX, dY, dZ, ipd[3L * i], ipd[3L * i + 1], ipd[3L * i + 2], f, cx,
                              cy, cz);
            }
            gx -= cx;
            gy -= cy;
            gz -= cz;
        }
    }
    if (q > 0) {
        sh[q - 1][0][li] = gx;
        sh[q - 1][1][li] = gy;
        sh[q - 1][2][li] = gz;
    }
    __syncthreads();
    if (q != 0 || !valid) return;
    for (int g = 0; g < kGfQ - 1; g++) {
        gx += sh[g][0][li];
        gy += sh[g][1][li];
        gz += sh[g][2][li];
    }
    gpol[3L * j] = gx;
    gpol[3L * j + 1] = gy;
    gpol[3L * j + 2] = gz;
}

// ------------------------------------------------------------ MM targets
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gmm_pairs(int nmm, int npol, int chunk, const double4 *__restrict__ ms,
                                                      const double *__restrict__ ma, const double *__restrict__ mq,
                                                      const double4 *__restrict__ ps,
                                                      const double *__restrict__ pa,
                                                      const int *__restrict__ mm_polar,
                                                      const double *__restrict__ ipd, double *__restrict__ part) {
    constexpr int NA = AMOEBA ? 12 : 3;
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_a[kBlock];
    __shared__ double s_n[3][kBlock];
    const int tid = threadIdx.x;
    const int m = blockIdx.x * kBlock + tid;
    const bool valid = m < nmm;
    const long n = 3L * npol;
    const double4 pm = valid ? ms[m] : make_double4(0.0, 0.0, 0.0, 1.0);
    const double am = valid ? ma[m] : 0.39;
    const int mp = valid ? mm_polar[m] : -1;
    double Q[6] = {0, 0, 0, 0, 0, 0};
    if (AMOEBA && valid)
        for (int c = 0; c < 6; c++) Q[c] = mq[10L * m + 4 + c];
    const double trQ = Q[0] + Q[2] + Q[5];
    double acc[NA];
#pragma unroll
    for (int c = 0; c < NA; c++) acc[c] = 0.0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(npol, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const long i = base + tid;
            s_pos[tid] = ps[i];
            s_a[tid] = pa[i];
#pragma unroll
            for (int c = 0; c < 3; c++)
                s_n[c][tid] = AMOEBA ? 0.5 * (ipd[3 * i + c] + ipd[n + 3 * i + c]) : ipd[3 * i + c];
        }
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const double4 pi = s_pos[k];
            const double ex = pm.x - pi.x, ey = pm.y - pi.y, ez = pm.z - pi.z;
            const double r2 = ex * ex + ey * ey + ez * ez;
            const double ir = (base + k) == mp ? 0.0 : fast_rsqrt(r2);
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, ir, pi.w, pm.w, s_a[k], am, k2, k3, k4, k5);
            dip_at_mm<AMOEBA>(ex, ey, ez, k2, k3, k4, k5, s_n[0][k], s_n[1][k], s_n[2][k], Q, trQ, acc);
        }
    }
    if (valid) {
        double *o = part + (long)blockIdx.y * NA * nmm;
#pragma unroll
        for (int c = 0; c < NA; c++) o[(long)c * nmm + m] = acc[c];
    }
}

// ------------------------------------- static-multipole pairs, both ends
// Each (polarizable site i, MM site m) pair evaluated once instead of once per
// end.  The pair's energy -nu_i . E_m(r_i) (nu = 1/2 (mu_D + mu_P) for AMOEBA,
// mu for AMBER; every pair at scale 1 as in the per-target sweeps) depends on
// r_i - r_m alone at fixed multipole orientation, so its gradient at m is minus
// its gradient at i, sum_b nu_b Egrd_ab of m's multipoles (static_grad_pair);
// the MM end also takes the field and field gradient of nu for the rotation
// term.  One workgroup per (MM block J, polarizable block I) of kBlock sites:
// wave w holds the MM sites of sub-tile w of J as targets (position,
// multipoles, field, field gradient and translational gradient in registers),
// the polarizable sites of I are staged in LDS; at step s lane l meets source
// (l + s) mod 64 of sub-tile k, and the source's gradient sum passes one lane
// down per step by the DPP wave rotate (after 64 steps it is back at its lane).
// Outputs: polarizable-end sums part_pol[J][3][npol], MM-end sums
// part_mm[I][NA][nmm]: E, Egrd and the translational gradient (AMOEBA), E (AMBER).
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gstat_both(int nmm, int npol, const double4 *__restrict__ ms,
                                                       const double *__restrict__ ma, const double *__restrict__ mq,
                                                       const double4 *__restrict__ ps,
                                                       const double *__restrict__ pa,
                                                       const int *__restrict__ mm_polar,
                                                       const double *__restrict__ ipd, double *__restrict__ part_pol,
                                                       double *__restrict__ part_mm) {
    constexpr int LD = AMOEBA ? 10 : 1;
    constexpr int NA = AMOEBA ? 12 : 3;
    static_assert(kBlock == 256, "four 64-site sub-tiles per block");
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_a[kBlock];
    __shared__ double s_n[3][kBlock];
    __shared__ double s_acc[kBlock / 64][3][64];
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int J = blockIdx.x, I = blockIdx.y;
    const long n = 3L * npol;
    // sources: polarizable sites of I; padding far away with zero dipoles
    {
        const int i = I * kBlock + tid;
        if (i < npol) {
            s_pos[tid] = ps[i];
            s_a[tid] = pa[i];
#pragma unroll
            for (int c = 0; c < 3; c++)
                s_n[c][tid] = AMOEBA ? 0.5 * (ipd[3L * i + c] + ipd[n + 3L * i + c]) : ipd[3L * i + c];
        } else {
            s_pos[tid] = make_double4(-1e15 - 1e9 * tid, -1e15, -1e15, 1.0);
            s_a[tid] = 0.39;
#pragma unroll
            for (int c = 0; c < 3; c++) s_n[c][tid] = 0.0;
        }
    }
    // target: MM site m of J; padding far away with zero multipoles
    const int m = J * kBlock + tid;
    const bool valid = m < nmm;
    const double4 pm = valid ? ms[m] : make_double4(1e15 + 1e9 * tid, 1e15, 1e15, 1.0);
    const double am = valid ? ma[m] : 0.39;
    const int self = valid ? mm_polar[m] - I * kBlock : -1;  // the polarizable index of m within I
    double q[LD];
#pragma unroll
    for (int c = 0; c < LD; c++) q[c] = valid ? mq[(long)LD * m + c] : 0.0;
    double acc[NA];
#pragma unroll
    for (int c = 0; c < NA; c++) acc[c] = 0.0;
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < kBlock / 64; k++) {
        double jg[3] = {0.0, 0.0, 0.0};  // gradient at the source now at this lane
        for (int s = 0; s < 64; s++) {
            const int js = k * 64 + ((l + s) & 63);
            const double4 pi = s_pos[js];
            const double ex = pm.x - pi.x, ey = pm.y - pi.y, ez = pm.z - pi.z;
            const double r2 = ex * ex + ey * ey + ez * ez;
            const double ir = js == self ? 0.0 : fast_rsqrt(r2);
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, ir, pi.w, pm.w, s_a[js], am, k2, k3, k4, k5);
            const double nx = s_n[0][js], ny = s_n[1][js], nz = s_n[2][js];
            // field (and gradient) of nu at the MM end; nu.e is shared with the
            // gradient at the polarizable end, d = r_i - r_m = -e
            const double ne = dip_field_at<AMOEBA>(ex, ey, ez, k2, k3, k4, nx, ny, nz, acc);
            double tx = 0.0, ty = 0.0, tz = 0.0;
            static_grad_pair_wd<AMOEBA>(-ex, -ey, -ez, k2, k3, k4, k5, q, nx, ny, nz, -ne, tx, ty, tz);
            jg[0] += tx;
            jg[1] += ty;
            jg[2] += tz;
            if (AMOEBA) {
                acc[9] -= tx;
                acc[10] -= ty;
                acc[11] -= tz;
            }
#pragma unroll
            for (int c = 0; c < 3; c++) jg[c] = lane_next(jg[c]);
        }
        // sources k * 64 + lane: the four waves' rows added in wave order
#pragma unroll
        for (int c = 0; c < 3; c++) s_acc[w][c][l] = jg[c];
        __syncthreads();
        if (tid < 3 * 64) {
            const int c = tid >> 6, jl = tid & 63, i = I * kBlock + k * 64 + jl;
            const double sum = (s_acc[0][c][jl] + s_acc[1][c][jl]) + (s_acc[2][c][jl] + s_acc[3][c][jl]);
            if (i < npol) part_pol[((long)J * 3 + c) * npol + i] = sum;
        }
        __syncthreads();
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < NA; c++) part_mm[((long)I * NA + c) * nmm + m] = acc[c];
    }
}

// E_D2M -> emm[m][3], Egrd_D2M -> egrd[m][6] (AMOEBA), and the MM-site gradient
// -q E + mu . Egrd - Q : EHes -> gtr[m][3].  direct (AMOEBA): the partials'
// entries 9-11 hold the translational gradient itself (k_gstat_both), and the
// formula applies to the listed-pair corrections only.
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_gmm_finalize(int nmm, int npol, int nparts, int direct,
                                                         const double *__restrict__ part,
                                                         const double4 *__restrict__ ms,
                                                         const double *__restrict__ ma, const double *__restrict__ mq,
                                                         const double4 *__restrict__ ps,
                                                         const double *__restrict__ pa,
                                                         const double *__restrict__ ipd,
                                                         const int *__restrict__ mp_ri, const int *__restrict__ mp_ci,
                                                         const double2 *__restrict__ mp_f, double *__restrict__ emm,
                                                         double *__restrict__ egrd, double *__restrict__ gtr) {
    constexpr int NA = AMOEBA ? 12 : 3;
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double sh[kGfQ - 1][2 * NA][kGfSites];
    const int qg = threadIdx.x / kGfSites, li = threadIdx.x % kGfSites;
    const int m = blockIdx.x * kGfSites + li;
    const bool valid = m < nmm;
    const long n = 3L * npol;
    double acc[NA], corr[NA];
    for (int c = 0; c < NA; c++) acc[c] = corr[c] = 0.0;
    double q[LD];
    for (int c = 0; c < LD; c++) q[c] = valid ? mq[(long)LD * m + c] : 0.0;
    const double *Q = q + (AMOEBA ? 4 : 0);
    const double trQ = AMOEBA ? Q[0] + Q[2] + Q[5] : 0.0;
    if (valid) {
        for (int s = qg; s < nparts; s += kGfQ) {
            const double *p = part + (long)s * NA * nmm;
            for (int c = 0; c < NA; c++) acc[c] += p[(long)c * nmm + m];
        }
        const double4 pm = ms[m];
        const double am = ma[m];
        // listed sources: remove 1/2 ((1 - s_d) mu_P + (1 - s_p) mu_D) (AMBER: (1 - s) mu)
        for (int k = mp_ri[m] + qg; k < mp_ri[m + 1]; k += kGfQ) {
            const int i = mp_ci[k];
            const double2 f = mp_f[k];
            double nx, ny, nz;
            if (AMOEBA) {
                nx = -0.5 * (f.x * ipd[n + 3L * i] + f.y * ipd[3L * i]);
                ny = -0.5 * (f.x * ipd[n + 3L * i + 1] + f.y * ipd[3L * i + 1]);
                nz = -0.5 * (f.x * ipd[n + 3L * i + 2] + f.y * ipd[3L * i + 2]);
            } else {
                nx = -f.x * ipd[3L * i];
                ny = -f.x * ipd[3L * i + 1];
                nz = -f.x * ipd[3L * i + 2];
            }
            const double4 pi = ps[i];
            const double ex = pm.x - pi.x, ey = pm.y - pi.y, ez = pm.z - pi.z;
            const double r2 = ex * ex + ey * ey + ez * ez;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pm.w, pa[i], am, k2, k3, k4, k5);
            dip_at_mm<AMOEBA>(ex, ey, ez, k2, k3, k4, k5, nx, ny, nz, Q, trQ, corr);
        }
    }
    if (qg > 0) {
        for (int c = 0; c < NA; c++) {
            sh[qg - 1][c][li] = acc[c];
            sh[qg - 1][NA + c][li] = corr[c];
        }
    }
    __syncthreads();
    if (qg != 0 || !valid) return;
    for (int g = 0; g < kGfQ - 1; g++)
        for (int c = 0; c < NA; c++) {
            acc[c] += sh[g][c][li];
            corr[c] += sh[g][NA + c][li];
        }
    // the formula's terms: of the corrections alone (direct), else of the totals
    double f[NA];
    for (int c = 0; c < NA; c++) f[c] = AMOEBA && direct ? corr[c] : acc[c] + corr[c];
    double g[3];
    for (int c = 0; c < 3; c++) {
        g[c] = -q[0] * f[c];
        emm[3L * m + c] = acc[c] + corr[c];
    }
    if (AMOEBA) {
        const double *G = f + 3;  // xx xy yy xz yz zz
        g[0] += q[1] * G[0] + q[2] * G[1] + q[3] * G[3] - f[9];
        g[1] += q[1] * G[1] + q[2] * G[2] + q[3] * G[4] - f[10];
        g[2] += q[1] * G[3] + q[2] * G[4] + q[3] * G[5] - f[11];
        if (direct)
            for (int c = 0; c < 3; c++) g[c] += acc[9 + c];
        for (int c = 0; c < 6; c++) egrd[6L * m + c] = acc[3 + c] + corr[3 + c];
    }
    for (int c = 0; c < 3; c++) gtr[3L * m + c] = g[c];
}

// ------------------------------------------------------------- rotation
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double iv = 1.0 / b.v;
    return {a.v * iv, (a.d - a.v * iv * b.d) * iv};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
    const double s = sqrt(a.v);
    return {s, 0.5 * a.d / s};
}
__device__ __forceinline__ Dual dot3(const Dual *a, const Dual *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// rotation_matrix (rotate_multipoles.F90:298-463): columns of r are the frame
// axes (ex, ey, ez) of a site at p[0] with x, y, z reference atoms at p[1..3].
__device__ void frame_axes(int conv, const Dual (&p)[4][3], Dual (&r)[3][3]) {
    Dual xi[3], eta[3], zeta[3], u[3], v[3];
    for (int k = 0; k < 3; k++) {
        xi[k] = p[3][k] - p[0][k];
        eta[k] = p[1][k] - p[0][k];
        zeta[k] = p[2][k] - p[0][k];
    }
    const Dual xn = dsqrt(dot3(xi, xi)), en = dsqrt(dot3(eta, eta)), zn = dsqrt(dot3(zeta, zeta));
    for (int k = 0; k < 3; k++) {
        switch (conv) {
            case 1:  // z-then-x
                u[k] = xi[k];
                v[k] = eta[k];
                break;
            case 2:  // bisector
                u[k] = en * xi[k] + xn * eta[k];
                v[k] = eta[k];
                break;
            case 3:  // z-only: v is a fixed lab axis
                u[k] = xi[k];
                v[k] = {0.0, 0.0};
                break;
            case 4:  // z-bisector
                u[k] = xi[k];
                v[k] = zn * eta[k] + en * zeta[k];
                break;
            default:  // 5: 3-fold
                u[k] = en * zn * xi[k] + xn * zn * eta[k] + xn * en * zeta[k];
                v[k] = eta[k];
                break;
        }
    }
    if (conv == 3) {
        if (u[2].v / xn.v <= 0.866 && fabs(u[1].v) > kEpsRpDev) v[0].v = 1.0;
        else v[1].v = 1.0;
    }
    const Dual un = dsqrt(dot3(u, u));
    Dual ez[3], ex[3], ey[3];
    for (int k = 0; k < 3; k++) ez[k] = u[k] / un;
    const Dual vez = dot3(v, ez);
    for (int k = 0; k < 3; k++) ex[k] = v[k] - vez * ez[k];
    const Dual exn = dsqrt(dot3(ex, ex));
    for (int k = 0; k < 3; k++) ex[k] = ex[k] / exn;
    ey[0] = ez[1] * ex[2] - ez[2] * ex[1];
    ey[1] = ez[2] * ex[0] - ez[0] * ex[2];
    ey[2] = ez[0] * ex[1] - ez[1] * ex[0];
    for (int k = 0; k < 3; k++) {
        r[k][0] = ex[k];
        r[k][1] = ey[k];
        r[k][2] = ez[k];
    }
}

// One thread per (site j, frame atom role: 0 self, 1 x, 2 y, 3 z); rotc[(4 j + role)][t]
// = d/dr_t of  -mu_lab . E_j + sum_ab Q_lab,ab Egrd_j,ab  through that atom.
__global__ __launch_bounds__(kBlock) void k_grot(int nmm, const double4 *__restrict__ ms,
                                                 const double *__restrict__ q0, const int4 *__restrict__ frames,
                                                 const double *__restrict__ emm, const double *__restrict__ egrd,
                                                 double *__restrict__ rotc) {
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4L * nmm) return;
    const int j = (int)(t >> 2), role = (int)(t & 3);
    const int4 f = frames[j];
    double *o = rotc + 3 * t;
    if (f.x == 0) {
        o[0] = o[1] = o[2] = 0.0;
        return;
    }
    // frames = (convention, z atom, x atom, y atom), -1 = the site itself
    const int at[4] = {j, f.z >= 0 ? f.z : j, f.w >= 0 ? f.w : j, f.y >= 0 ? f.y : j};
    double pos[4][3];
    for (int a = 0; a < 4; a++) {
        const double4 c = ms[at[a]];
        pos[a][0] = c.x;
        pos[a][1] = c.y;
        pos[a][2] = c.z;
    }
    const double *mu0 = q0 + 10L * j + 1, *Qs = q0 + 10L * j + 4;
    const double Q0[3][3] = {{Qs[0], Qs[1], Qs[3]}, {Qs[1], Qs[2], Qs[4]}, {Qs[3], Qs[4], Qs[5]}};
    const double *g6 = egrd + 6L * j;
    const double G[3][3] = {{g6[0], g6[1], g6[3]}, {g6[1], g6[2], g6[4]}, {g6[3], g6[4], g6[5]}};
    const double E[3] = {emm[3L * j], emm[3L * j + 1], emm[3L * j + 2]};
    for (int tc = 0; tc < 3; tc++) {
        Dual p[4][3], R[3][3];
        for (int a = 0; a < 4; a++)
            for (int k = 0; k < 3; k++) p[a][k] = {pos[a][k], (a == role && k == tc) ? 1.0 : 0.0};
        frame_axes(f.x, p, R);
        // -(dR mu0) . E + 2 sum_ab (dR Q0 R^T)_ab G_ab  (G symmetric)
        double acc = 0.0;
        for (int a = 0; a < 3; a++) {
            const double dmu = R[a][0].d * mu0[0] + R[a][1].d * mu0[1] + R[a][2].d * mu0[2];
            acc -= dmu * E[a];
            double M[3];
            for (int c = 0; c < 3; c++) M[c] = R[a][0].d * Q0[0][c] + R[a][1].d * Q0[1][c] + R[a][2].d * Q0[2][c];
            for (int b = 0; b < 3; b++) {
                const double mr = M[0] * R[b][0].v + M[1] * R[b][1].v + M[2] * R[b][2].v;
                acc += 2.0 * mr * G[a][b];
            }
        }
        o[tc] = acc;
    }
}

// out[a] = gtr[a] + gpol[mm_polar[a]] + sum of the rotation terms listed for a.
__global__ __launch_bounds__(kBlock) void k_gassemble(int nmm, const double *__restrict__ gtr,
                                                      const double *__restrict__ gpol,
                                                      const int *__restrict__ mm_polar,
                                                      const int *__restrict__ rev_ri, const int *__restrict__ rev_ci,
                                                      const double *__restrict__ rotc, double *__restrict__ out) {
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= nmm) return;
    double g[3] = {0.0, 0.0, 0.0};
    if (gtr)
        for (int c = 0; c < 3; c++) g[c] = gtr[3L * a + c];
    if (gpol) {
        const int p = mm_polar[a];
        if (p >= 0)
            for (int c = 0; c < 3; c++) g[c] += gpol[3L * p + c];
    }
    if (rotc)
        for (int k = rev_ri[a]; k < rev_ri[a + 1]; k++) {
            const long e = rev_ci[k];
            for (int c = 0; c < 3; c++) g[c] += rotc[3 * e + c];
        }
    for (int c = 0; c < 3; c++) out[3L * a + c] = g[c];
}

// ------------------------------------------------------------- launchers
// From this size each (polarizable, MM) pair of the static part is evaluated
// once for both ends (k_gstat_both), while its partial sums (one per block of
// the other end) stay under 4 GiB; below it the two per-target sweeps with
// source splits fill the chip better.
bool grad_both_ends(int npol, int nmm, bool amoeba) {
    if (npol < 5000) return false;
    const size_t pb = (npol + kBlock - 1) / kBlock, mb = (nmm + kBlock - 1) / kBlock;
    return sizeof(double) * (mb * 3 * (size_t)npol + pb * (amoeba ? 12 : 3) * (size_t)nmm) <= (size_t(4) << 30);
}

template <bool AMOEBA>
static void geomgrad_sweeps(hipStream_t st, const GradArgs &a, int &nparts_pol, int &nparts_mm) {
    const int nmm = a.nmm, npol = a.npol;
    const int pb = (npol + kBlock - 1) / kBlock, mb = (nmm + kBlock - 1) / kBlock;
    const int cs = (nmm + a.split_s - 1) / a.split_s, cp = (npol + a.split_p - 1) / a.split_p;
    const int cm = (npol + a.split_m - 1) / a.split_m;
    // static part: both ends at once (mb polarizable-end partials, pb MM-end
    // ones) or per target (split_s and split_m partials)
    const int nstat = a.fused ? mb : a.split_s;
    if (a.fused) {
        hipLaunchKernelGGL(k_gstat_both<AMOEBA>, dim3(mb, pb), dim3(kBlock), 0, st, nmm, npol, a.ms, a.ma, a.mq,
                           a.ps, a.pa, a.mm_polar, a.ipd, a.part_pol, a.part_mm);
        nparts_mm = pb;
    } else {
        hipLaunchKernelGGL(k_gpol_static<AMOEBA>, dim3(pb, a.split_s), dim3(kBlock), 0, st, nmm, npol, cs, a.ms, a.ma,
                           a.mq, a.ps, a.pa, a.polar_mm, a.ipd, a.part_pol);
        hipLaunchKernelGGL(k_gmm_pairs<AMOEBA>, dim3(mb, a.split_m), dim3(kBlock), 0, st, nmm, npol, cm, a.ms, a.ma,
                           a.mq, a.ps, a.pa, a.mm_polar, a.ipd, a.part_mm);
        nparts_mm = a.split_m;
    }
    // the dipole-dipole part pair-once into one partial after the static ones,
    // or per target
    double *part_p = a.part_pol + (long)nstat * 3 * npol;
    if (a.sym_pairs) {
        const int nb = tmv_sym_blocks(npol);
        const long np = (long)nb * (nb + 1) / 2;
        double *sym_i = a.sympart, *sym_j = a.sympart + np * 3 * kSymBlock;
        hipLaunchKernelGGL(k_gpol_dd_sym<AMOEBA>, dim3((unsigned)np), dim3(kSymBlock), 0, st, npol, a.sym_pairs,
                           a.ps, a.pa, a.ipd, sym_i, sym_j);
        hipLaunchKernelGGL(k_sym3_reduce, dim3(nb), dim3(4 * kSymBlock), 0, st, npol, nb, a.sympart, part_p);
        nparts_pol = nstat + 1;
    } else {
        hipLaunchKernelGGL(k_gpol_dd<AMOEBA>, dim3(pb, a.split_p), dim3(kBlock), 0, st, npol, cp, a.ps, a.pa, a.ipd,
                           part_p);
        nparts_pol = nstat + a.split_p;
    }
    HIP_CHECK(hipGetLastError());
}

void launch_geomgrad(hipStream_t st, const GradArgs &a, hipEvent_t evb, hipEvent_t eve) {
    int nparts_pol = 0, nparts_mm = 0;
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    if (a.npol > 0) {
        if (a.amoeba) geomgrad_sweeps<true>(st, a, nparts_pol, nparts_mm);
        else geomgrad_sweeps<false>(st, a, nparts_pol, nparts_mm);
    }
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
    launch_geomgrad_finish(st, a, nparts_pol, nparts_mm, a.fused);
}

void launch_geomgrad_finish(hipStream_t st, const GradArgs &a, int nparts_pol, int nparts_mm, bool direct) {
    const int nmm = a.nmm, npol = a.npol;
    const int mb = (nmm + kBlock - 1) / kBlock;
    const int pbf = (npol + kGfSites - 1) / kGfSites, mbf = (nmm + kGfSites - 1) / kGfSites;
    if (npol > 0) {
        if (a.amoeba) {
            hipLaunchKernelGGL(k_gpol_finalize<true>, dim3(pbf), dim3(kBlock), 0, st, npol, nparts_pol, a.part_pol,
                               a.ms, a.ma, a.mq, a.ps, a.pa, a.ipd, a.sp.ri, a.sp.ci, a.sp.f2, a.pp.ri, a.pp.ci,
                               a.pp.f, a.gpol);
            hipLaunchKernelGGL(k_gmm_finalize<true>, dim3(mbf), dim3(kBlock), 0, st, nmm, npol, nparts_mm,
                               (int)direct, a.part_mm, a.ms, a.ma, a.mq, a.ps, a.pa, a.ipd, a.mp.ri, a.mp.ci,
                               a.mp.f2, a.emm, a.egrd, a.gtr);
        } else {
            hipLaunchKernelGGL(k_gpol_finalize<false>, dim3(pbf), dim3(kBlock), 0, st, npol, nparts_pol, a.part_pol,
                               a.ms, a.ma, a.mq, a.ps, a.pa, a.ipd, a.sp.ri, a.sp.ci, a.sp.f2, a.pp.ri, a.pp.ci,
                               a.pp.f, a.gpol);
            hipLaunchKernelGGL(k_gmm_finalize<false>, dim3(mbf), dim3(kBlock), 0, st, nmm, npol, nparts_mm,
                               (int)direct, a.part_mm, a.ms, a.ma, a.mq, a.ps, a.pa, a.ipd, a.mp.ri, a.mp.ci,
                               a.mp.f2, a.emm, a.egrd, a.gtr);
        }
        HIP_CHECK(hipGetLastError());
    }
    if (a.amoeba) {
        const long nt = 4L * nmm;
        hipLaunchKernelGGL(k_grot, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, nmm, a.ms,
                           a.q0, a.frames, a.emm, a.egrd, a.rotc);
        HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gassemble, dim3(mb), dim3(kBlock), 0, st, nmm, npol > 0 ? a.gtr : nullptr,
                       npol > 0 ? a.gpol : nullptr, a.mm_polar, a.rev_ri, a.rev_ci, a.amoeba ? a.rotc : nullptr,
                       a.out);
    HIP_CHECK(hipGetLastError());
}

void launch_rotation_grad(hipStream_t st, const GradArgs &a) {
    const int nmm = a.nmm;
    const long nt = 4L * nmm;
    hipLaunchKernelGGL(k_grot, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, nmm, a.ms, a.q0,
                       a.frames, a.emm, a.egrd, a.rotc);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_gassemble, dim3((nmm + kBlock - 1) / kBlock), dim3(kBlock), 0, st, nmm, nullptr, nullptr,
                       a.mm_polar, a.rev_ri, a.rev_ci, a.rotc, a.out);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
