// HIP kernels of the induced-dipole path for gfx950 (CDNA4, wave64).
//
// Pair kernels: one thread per polarizable target, sources staged through LDS
// in tiles of 256 and broadcast to the wave, the source range split over
// gridDim.y so that a 15k-site system launches ~1000 workgroups.  Partial sums
// per split go to a scratch buffer and are added in a fixed order by the
// finalize kernel, which also applies the sparse exclusion corrections: the
// dense sweep uses scale 1 for every pair and the finalize subtracts
// (1 - s) * E_ij for the listed pairs (s = 0 removes the pair).  Results do not
// depend on the batch or launch composition.
//
// Vector kernels: FP64, one element per thread, deterministic reductions
// (fixed tree per block, fixed-order sum of the per-block partials).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "device.h"
#include "pair_math.h"

namespace ommp {


int vec_blocks(long n) { return (int)((n + kBlock - 1) / kBlock); }
int site_blocks(int npol) { return (npol + kBlock - 1) / kBlock; }

// ====================================================================== RHS
// Field of sources [blockIdx.y * chunk, +chunk) of the nsrc sources at every
// polarizable target: the MM sites themselves (src == nullptr, the self pair
// skipped) or the MM sites src[0..nsrc) (non-polarizable ones: never a target).
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_rhs_pairs(int nsrc, int npol, int chunk, const int *__restrict__ src,
                                                      const double4 *__restrict__ ms,
                                                      const double *__restrict__ ma,
                                                      const double *__restrict__ mq,
                                                      const double4 *__restrict__ ps,
                                                      const double *__restrict__ pa,
                                                      const int *__restrict__ polar_mm,
                                                      double *__restrict__ part) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_a[kBlock];
    __shared__ double s_q[LD][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < npol;
    const double4 pj = valid ? ps[j] : make_double4(0.0, 0.0, 0.0, 1.0);
    const double aj = valid ? pa[j] : 0.39;
    const int jm = valid && !src ? polar_mm[j] : -1;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(nsrc, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const int si = src ? src[base + tid] : base + tid;
            s_pos[tid] = ms[si];
            s_a[tid] = ma[si];
#pragma unroll
            for (int c = 0; c < LD; c++) s_q[c][tid] = mq[(long)LD * si + c];
        }
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const double4 pi = s_pos[k];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool self = (base + k) == jm;
            const double ir = self ? 0.0 : fast_rsqrt(r2);
            double q[LD];
#pragma unroll
            for (int c = 0; c < LD; c++) q[c] = s_q[c][k];
            double fx, fy, fz;
            m2d_pair<AMOEBA>(dx, dy, dz, ir, pi.w, pj.w, s_a[k], aj, q, fx, fy, fz);
            ex += fx;
            ey += fy;
            ez += fz;
        }
    }
    if (valid) {
        double *o = part + (long)blockIdx.y * 3 * npol + 3L * j;
        o[0] = ex;
        o[1] = ey;
        o[2] = ez;
    }
}

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_rhs_finalize(int npol, int nsplit, const double *__restrict__ part,
                                                         const double4 *__restrict__ ms,
                                                         const double *__restrict__ ma,
                                                         const double *__restrict__ mq,
                                                         const double4 *__restrict__ ps,
                                                         const double *__restrict__ pa,
                                                         const int *__restrict__ sp_ri,
                                                         const int *__restrict__ sp_ci,
                                                         const double2 *__restrict__ sp_f,
                                                         double *__restrict__ e_m2d) {
    constexpr int LD = AMOEBA ? 10 : 1;
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= npol) return;
    double e[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < nsplit; s++) {
        const double *p = part + (long)s * 3 * npol + 3L * j;
        e[0] += p[0];
        e[1] += p[1];
        e[2] += p[2];
    }
    double ed[3] = {e[0], e[1], e[2]}, ep[3] = {e[0], e[1], e[2]};
    const double4 pj = ps[j];
    const double aj = pa[j];
    for (int k = sp_ri[j]; k < sp_ri[j + 1]; k++) {
        const int i = sp_ci[k];
        const double2 f = sp_f[k];
        const double4 pi = ms[i];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double ir = rsqrt(dx * dx + dy * dy + dz * dz);
        double q[LD];
        for (int c = 0; c < LD; c++) q[c] = mq[(long)LD * i + c];
        double fx, fy, fz;
        m2d_pair<AMOEBA>(dx, dy, dz, ir, pi.w, pj.w, ma[i], aj, q, fx, fy, fz);
        ed[0] -= f.x * fx;
        ed[1] -= f.x * fy;
        ed[2] -= f.x * fz;
        ep[0] -= f.y * fx;
        ep[1] -= f.y * fy;
        ep[2] -= f.y * fz;
    }
    if (AMOEBA) {
        for (int c = 0; c < 3; c++) {
            e_m2d[3L * j + c] = ed[c];
            e_m2d[3L * npol + 3L * j + c] = ep[c];
        }
    } else {
        for (int c = 0; c < 3; c++) e_m2d[3L * j + c] = ep[c];
    }
}

void launch_rhs_pairs_list(hipStream_t st, bool amoeba, int nsrc, const int *src, int npol, const double4 *ms64,
                           const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                           double *part, int nsplit) {
    if (npol == 0 || nsrc == 0) return;
    const int chunk = (nsrc + nsplit - 1) / nsplit;
    dim3 grid(site_blocks(npol), nsplit);
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_pairs<true>, grid, dim3(kBlock), 0, st, nsrc, npol, chunk, src, ms64, ma64, mq, ps64,
                           pa64, nullptr, part);
    else
        hipLaunchKernelGGL(k_rhs_pairs<false>, grid, dim3(kBlock), 0, st, nsrc, npol, chunk, src, ms64, ma64, mq,
                           ps64, pa64, nullptr, part);
    HIP_CHECK(hipGetLastError());
}

void launch_rhs_m2d(hipStream_t st, bool amoeba, int nmm, int npol, const double4 *ms64,
                    const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                    const int *polar_mm, const CsrDev &sp, double *part, int nsplit, double *e_m2d,
                    hipEvent_t evb, hipEvent_t eve) {
    if (npol == 0) return;
    const int chunk = (nmm + nsplit - 1) / nsplit;
    dim3 grid(site_blocks(npol), nsplit);
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_pairs<true>, grid, dim3(kBlock), 0, st, nmm, npol, chunk, nullptr, ms64, ma64, mq,
                           ps64, pa64, polar_mm, part);
    else
        hipLaunchKernelGGL(k_rhs_pairs<false>, grid, dim3(kBlock), 0, st, nmm, npol, chunk, nullptr, ms64, ma64, mq,
                           ps64, pa64, polar_mm, part);
    HIP_CHECK(hipGetLastError());
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_finalize<true>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nsplit,
                           part, ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    else
        hipLaunchKernelGGL(k_rhs_finalize<false>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nsplit,
                           part, ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    HIP_CHECK(hipGetLastError());
}

// ============================================================ T on the fly
// One target per thread, sources through LDS two at a time.  Tiles are padded
// to an even count with far-away sources of zero dipole, which add exactly 0.
template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_tmv_pairs(int npol, int chunk,
                                                      const typename V4<R>::t *__restrict__ site,
                                                      const R *__restrict__ adamp,
                                                      const double *__restrict__ x, long xstride,
                                                      R *__restrict__ part, const CgState *__restrict__ skip,
                                                      int skipM) {
    using R4 = typename V4<R>::t;
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ R4 s_pos[kBlock];
    __shared__ R4 s_x[M][kBlock];  // x, y, z of column m; w of column 0 holds the damping a
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < npol;
    R4 pj;
    R aj;
    if (valid) {
        pj = site[j];
        aj = adamp[j];
    } else {
        pj.x = pj.y = pj.z = R(0);
        pj.w = R(1);
        aj = R(0.39);
    }
    R acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = R(0);
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(npol, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const int i = base + tid;
            s_pos[tid] = site[i];
#pragma unroll
            for (int m = 0; m < M; m++) {
                R4 v;
                v.x = (R)x[m * xstride + 3L * i + 0];
                v.y = (R)x[m * xstride + 3L * i + 1];
                v.z = (R)x[m * xstride + 3L * i + 2];
                v.w = m == 0 ? adamp[i] : R(0);
                s_x[m][tid] = v;
            }
        } else if (tid == cnt) {
            R4 far;
            far.x = far.y = far.z = R(1e15);
            far.w = R(1);
            s_pos[tid] = far;
#pragma unroll
            for (int m = 0; m < M; m++) {
                R4 z;
                z.x = z.y = z.z = R(0);
                z.w = R(0.39);
                s_x[m][tid] = z;
            }
        }
        __syncthreads();
        const int cnt2 = (cnt + 1) & ~1;
        for (int k = 0; k < cnt2; k += 2) {
            R4 xa[M], xb[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                xa[m] = s_x[m][k];
                xb[m] = s_x[m][k + 1];
            }
            R dxa, dya, dza, k2a, k3a, arga, dxb, dyb, dzb, k2b, k3b, argb;
            bool neara, nearb;
            pair_geom<R, AMOEBA>(pj, aj, s_pos[k], xa[0].w, dxa, dya, dza, k2a, k3a, arga, neara);
            pair_geom<R, AMOEBA>(pj, aj, s_pos[k + 1], xb[0].w, dxb, dyb, dzb, k2b, k3b, argb, nearb);
            if (__ballot(neara || nearb)) {
                pair_damp<R, AMOEBA>(arga, neara, k2a, k3a);
                pair_damp<R, AMOEBA>(argb, nearb, k2b, k3b);
            }
            pair_accum<R, M>(acc, xa, dxa, dya, dza, k2a, k3a);
            pair_accum<R, M>(acc, xb, dxb, dyb, dzb, k2b, k3b);
        }
    }
    if (valid) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            R *o = part + ((long)blockIdx.y * M + m) * 3 * npol + 3L * j;
            o[0] = acc[m][0];
            o[1] = acc[m][1];
            o[2] = acc[m][2];
        }
    }
}

// y = -(sum of partial fields part[split][m][3 npol]) + corrections (+ x / alpha);
// optional y.x partials.
template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_tmv_finalize(int npol, int nsplit, const R *__restrict__ part,
                                                         const double4 *__restrict__ ps,
                                                         const double *__restrict__ pa,
                                                         const double *__restrict__ alpha,
                                                         const int *__restrict__ pp_ri,
                                                         const int *__restrict__ pp_ci,
                                                         const double *__restrict__ pp_f,
                                                         const double *__restrict__ x, long xstride,
                                                         double *__restrict__ y, long ystride, int dodiag,
                                                         double *__restrict__ dotpart,
                                                         const CgState *__restrict__ skip, int skipM) {
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ double sh[4];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = j < npol;
    double acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    if (valid) {
        for (int s = 0; s < nsplit; s++)
#pragma unroll
            for (int m = 0; m < M; m++) {
                const R *p = part + ((long)s * M + m) * 3 * npol + 3L * j;
                acc[m][0] -= (double)p[0];
                acc[m][1] -= (double)p[1];
                acc[m][2] -= (double)p[2];
            }
    }
    if (valid) {
        const double4 pj = ps[j];
        const double aj = pa[j];
        for (int k = pp_ri[j]; k < pp_ri[j + 1]; k++) {
            const int i = pp_ci[k];
            const double f = pp_f[k];
            const double4 pi = ps[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3;
            dip_kernel<double, AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, pa[i], aj, k2, k3);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const double *xi = x + m * xstride + 3L * i;
                const double c = k3 * (xi[0] * dx + xi[1] * dy + xi[2] * dz);
                acc[m][0] += f * (c * dx - k2 * xi[0]);
                acc[m][1] += f * (c * dy - k2 * xi[1]);
                acc[m][2] += f * (c * dz - k2 * xi[2]);
            }
        }
        if (dodiag) {
            const double ia = 1.0 / alpha[j];
#pragma unroll
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 3; c++) acc[m][c] += x[m * xstride + 3L * j + c] * ia;
        }
#pragma unroll
        for (int m = 0; m < M; m++)
            for (int c = 0; c < 3; c++) y[m * ystride + 3L * j + c] = acc[m][c];
    }
    if (dotpart) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            double v = 0.0;
            if (valid)
                for (int c = 0; c < 3; c++) v += acc[m][c] * x[m * xstride + 3L * j + c];
            v = block_sum_all(v, sh);
            if (threadIdx.x == 0) dotpart[m * gridDim.x + blockIdx.x] = v;
        }
    }
}

// Source splits for a pair kernel with tb target blocks: enough workgroups to
// fill every CU several times, and a count that divides evenly over the CUs
// (a partial last round of workgroups idles the rest of the chip).
int balanced_split(int tb, int nsrc, int num_cu) {
    const int max_s = std::max(1, (nsrc + kBlock - 1) / kBlock);  // >= one LDS tile per split
    int best = 1;
    double best_cost = 1e30;
    for (int s = 1; s <= max_s; s++) {
        const long blocks = (long)tb * s;
        const double rounds = (double)blocks / num_cu;
        if (rounds > 16.0) break;
        const double eff = rounds / std::ceil(rounds);  // utilisation of the last round
        // prefer 2..8 workgroups per CU, then balance
        const double fill = rounds < 2.0 ? (2.0 - rounds) : 0.0;
        const double cost = (1.0 - eff) + 0.5 * fill + 0.003 * rounds;
        if (cost < best_cost) {
            best_cost = cost;
            best = s;
        }
    }
    return best;
}

int tmv_split(int npol) {
    static int num_cu = 0;
    if (num_cu == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            num_cu = prop.multiProcessorCount;
        if (num_cu <= 0) num_cu = 256;
    }
    return balanced_split(site_blocks(npol), npol, num_cu);
}

size_t tmv_part_bytes(int npol, int M, int prec, int nsplit) {
    return (size_t)nsplit * M * 3 * (size_t)npol * (prec == 32 ? 4 : 8);
}

template <typename R, int M>
static void tmv_launch_t(hipStream_t st, bool amoeba, int npol, const typename V4<R>::t *site, const R *ad,
                         const double4 *ps64, const double *pa64, const double *alpha, const CsrDev &pp,
                         const double *x, long xs, double *y, long ys, bool dodiag, double *dotpart,
                         R *part, int nsplit, const CgState *skip, int skipM, hipEvent_t evb, hipEvent_t eve) {
    const int chunk = (npol + nsplit - 1) / nsplit;
    dim3 grid(site_blocks(npol), nsplit);
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    if (amoeba)
        hipLaunchKernelGGL((k_tmv_pairs<R, M, true>), grid, dim3(kBlock), 0, st, npol, chunk, site, ad, x, xs,
                           part, skip, skipM);
    else
        hipLaunchKernelGGL((k_tmv_pairs<R, M, false>), grid, dim3(kBlock), 0, st, npol, chunk, site, ad, x, xs,
                           part, skip, skipM);
    HIP_CHECK(hipGetLastError());
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
    if (amoeba)
        hipLaunchKernelGGL((k_tmv_finalize<R, M, true>), dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol,
                           nsplit, part, ps64, pa64, alpha, pp.ri, pp.ci, pp.f, x, xs, y, ys, (int)dodiag,
                           dotpart, skip, skipM);
    else
        hipLaunchKernelGGL((k_tmv_finalize<R, M, false>), dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol,
                           nsplit, part, ps64, pa64, alpha, pp.ri, pp.ci, pp.f, x, xs, y, ys, (int)dodiag,
                           dotpart, skip, skipM);
    HIP_CHECK(hipGetLastError());
}

void launch_tmv_otf(hipStream_t st, int prec, bool amoeba, int npol, int M, const double4 *ps64,
                    const double *pa64, const float4 *ps32, const float *pa32, const double *alpha,
                    const CsrDev &pp, const double *x, long xstride, double *y, long ystride, bool dodiag,
                    double *dotpart, void *part, int nsplit, const int *skip_flag, hipEvent_t evb,
                    hipEvent_t eve) {
    if (npol == 0) return;
    const CgState *skip = reinterpret_cast<const CgState *>(skip_flag);
    const int skipM = M;
    if (prec == 32) {
        if (M == 1)
            tmv_launch_t<float, 1>(st, amoeba, npol, ps32, pa32, ps64, pa64, alpha, pp, x, xstride, y, ystride,
                                   dodiag, dotpart, (float *)part, nsplit, skip, skipM, evb, eve);
        else
            tmv_launch_t<float, 2>(st, amoeba, npol, ps32, pa32, ps64, pa64, alpha, pp, x, xstride, y, ystride,
                                   dodiag, dotpart, (float *)part, nsplit, skip, skipM, evb, eve);
    } else {
        if (M == 1)
            tmv_launch_t<double, 1>(st, amoeba, npol, ps64, pa64, ps64, pa64, alpha, pp, x, xstride, y, ystride,
                                    dodiag, dotpart, (double *)part, nsplit, skip, skipM, evb, eve);
        else
            tmv_launch_t<double, 2>(st, amoeba, npol, ps64, pa64, ps64, pa64, alpha, pp, x, xstride, y, ystride,
                                    dodiag, dotpart, (double *)part, nsplit, skip, skipM, evb, eve);
    }
}

// Finalize of externally computed FP64 partial fields (the fast multipole
// path: near and far field as two splits).
void launch_tmv_finalize64(hipStream_t st, bool amoeba, int npol, int M, const double *part, int nsplit,
                           const double4 *ps64, const double *pa64, const double *alpha, const CsrDev &pp,
                           const double *x, long xstride, double *y, long ystride, bool dodiag, double *dotpart,
                           const int *skip_flag) {
    if (npol == 0) return;
    const CgState *skip = reinterpret_cast<const CgState *>(skip_flag);
    const dim3 g(site_blocks(npol)), b(kBlock);
    if (amoeba && M == 1)
        hipLaunchKernelGGL((k_tmv_finalize<double, 1, true>), g, b, 0, st, npol, nsplit, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xstride, y, ystride, (int)dodiag, dotpart, skip, M);
    else if (amoeba)
        hipLaunchKernelGGL((k_tmv_finalize<double, 2, true>), g, b, 0, st, npol, nsplit, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xstride, y, ystride, (int)dodiag, dotpart, skip, M);
    else if (M == 1)
        hipLaunchKernelGGL((k_tmv_finalize<double, 1, false>), g, b, 0, st, npol, nsplit, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xstride, y, ystride, (int)dodiag, dotpart, skip, M);
    else
        hipLaunchKernelGGL((k_tmv_finalize<double, 2, false>), g, b, 0, st, npol, nsplit, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xstride, y, ystride, (int)dodiag, dotpart, skip, M);
    HIP_CHECK(hipGetLastError());
}

void launch_rhs_finalize(hipStream_t st, bool amoeba, int npol, int nsplit, const double *part, const double4 *ms64,
                         const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                         const CsrDev &sp, double *e_m2d) {
    if (npol == 0) return;
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_finalize<true>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nsplit, part,
                           ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    else
        hipLaunchKernelGGL(k_rhs_finalize<false>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nsplit, part,
                           ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    HIP_CHECK(hipGetLastError());
}

// ===================================================== packed in-core T
// Upper tile pairs (I <= J) of 64x64 sites, 6 tensor planes (xx, xy, xz, yy,
// yz, zz).  Thread t of the 256-thread block owns the 4x4 sub-block of rows
// 4*(t>>4).. and columns 4*(t&15)..  Inside a tile pair the entries are ordered
// [k][t][4], k = 6 * (row within the sub-block) + plane: one 16-byte (FP32)
// vector per thread and step, so every wave-wide load of a step reads one
// contiguous 1 KB (FP32) or 2 KB (FP64) run.
constexpr int kTile = 64;
constexpr int kTileEntries = kTile * kTile;
constexpr int kTpackGroups = 16;  // source-tile groups of the first reduction stage

__host__ __device__ inline long tile_pair_index(long I, long J, long nt) { return I * nt - I * (I - 1) / 2 + (J - I); }

__device__ inline void tile_pair_from_index(long p, long nt, int &I, int &J) {
    // largest I with start(I) <= p, start(I) = I*nt - I*(I-1)/2
    long lo = 0, hi = nt - 1;
    while (lo < hi) {
        long mid = (lo + hi + 1) / 2;
        if (tile_pair_index(mid, mid, nt) <= p) lo = mid;
        else hi = mid - 1;
    }
    I = (int)lo;
    J = (int)(lo + (p - tile_pair_index(lo, lo, nt)));
}

__host__ __device__ inline long tpack_offset(int a, int b, int e) {
    const int t = (a >> 2) * 16 + (b >> 2);
    return ((long)((a & 3) * 6 + e) * kBlock + t) * 4 + (b & 3);
}

size_t tpack_bytes(int npol, int prec) {
    const long nt = (npol + kTile - 1) / kTile;
    return (size_t)(nt * (nt + 1) / 2) * 6 * kTileEntries * (prec == 32 ? 4 : 8);
}

template <typename R, bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_tpack_build(int npol, long nt, const double4 *__restrict__ ps,
                                                        const double *__restrict__ pa, R *__restrict__ tp) {
    using R4 = typename V4<R>::t;
    int I, J;
    tile_pair_from_index(blockIdx.x, nt, I, J);
    const int t = threadIdx.x;
    const int rg = t >> 4, cg = t & 15;
    R4 *base = reinterpret_cast<R4 *>(tp + (long)blockIdx.x * 6 * kTileEntries) + t;
    for (int ai = 0; ai < 4; ai++) {
        R v[6][4];
        for (int bj = 0; bj < 4; bj++) {
            const int i = I * kTile + rg * 4 + ai;
            const int j = J * kTile + cg * 4 + bj;
            for (int e = 0; e < 6; e++) v[e][bj] = R(0);
            if (i < npol && j < npol && i != j) {
                const double4 pi = ps[i], pj = ps[j];
                const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                double k2, k3;
                dip_kernel<double, AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, pa[i], pa[j], k2, k3);
                // T_ij = k2 I - 3 k3 d d^T (k3 already carries the factor 3)
                v[0][bj] = (R)(k2 - k3 * dx * dx);
                v[1][bj] = (R)(-k3 * dx * dy);
                v[2][bj] = (R)(-k3 * dx * dz);
                v[3][bj] = (R)(k2 - k3 * dy * dy);
                v[4][bj] = (R)(-k3 * dy * dz);
                v[5][bj] = (R)(k2 - k3 * dz * dz);
            }
        }
        for (int e = 0; e < 6; e++) {
            R4 w;
            w.x = v[e][0];
            w.y = v[e][1];
            w.z = v[e][2];
            w.w = v[e][3];
            base[(ai * 6 + e) * kBlock] = w;
        }
    }
}

// Scale listed pairs (dense screening rule, create_tmat semantics).
template <typename R>
__global__ void k_tpack_scale(int npol, long nt, const int *__restrict__ ri, const int *__restrict__ ci,
                              const double *__restrict__ sv, R *__restrict__ tp) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npol) return;
    for (int k = ri[i]; k < ri[i + 1]; k++) {
        const int j = ci[k];
        const int I = i / kTile, J = j / kTile;
        if (I > J) continue;  // stored once, in tile (J, I), handled by row j
        R *p = tp + tile_pair_index(I, J, nt) * 6 * kTileEntries;
        for (int e = 0; e < 6; e++) {
            const long o = tpack_offset(i % kTile, j % kTile, e);
            p[o] = (R)((double)p[o] * sv[k]);
        }
    }
}

size_t tpack_part_bytes(int npol, int M) {
    const long nt = (npol + kTile - 1) / kTile;
    // per-source-tile partials (sized for FP64) + the grouped FP64 first stage
    return (size_t)nt * M * 3 * (size_t)(nt * kTile) * 8 + (size_t)kTpackGroups * M * 3 * (size_t)npol * 8;
}

// Four packed values read once per matvec: nontemporal loads, which keep T
// out of the caches (6.0-6.3 TB/s for a plain read stream against 5.3-5.8 TB/s
// with ordinary loads on MI355X).
template <typename R>
__device__ __forceinline__ typename V4<R>::t tpack_load(const typename V4<R>::t *p) {
    typedef R v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(p));
    typename V4<R>::t r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = v.w;
    return r;
}

// One tile pair per block.  Row sums go to part[J][.][rows of I], column sums
// to part[I][.][rows of J]: every (source tile, target) slot is written by
// exactly one block, so the reduction is deterministic.  The next sub-block
// row of T is loaded while the current one is used.
template <typename R, int M>
__global__ __launch_bounds__(kBlock) void k_tpack_mv(int npol, long nt, const R *__restrict__ tp,
                                                     const double *__restrict__ x, long xstride,
                                                     R *__restrict__ part, const CgState *__restrict__ skip,
                                                     int skipM) {
    using R4 = typename V4<R>::t;
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    int I, J;
    tile_pair_from_index(blockIdx.x, nt, I, J);
    __shared__ R sxI[M][3][kTile], sxJ[M][3][kTile];
    __shared__ R scol[4][M * 3][kTile];
    const int t = threadIdx.x;
    const long npad = nt * kTile;
    const R4 *tv = reinterpret_cast<const R4 *>(tp + (long)blockIdx.x * 6 * kTileEntries) + t;
    R4 cur[6];
#pragma unroll
    for (int e = 0; e < 6; e++) cur[e] = tpack_load<R>(tv + e * kBlock);
    if (t < kTile) {
        const int i = I * kTile + t, j = J * kTile + t;
        for (int m = 0; m < M; m++)
            for (int c = 0; c < 3; c++) {
                sxI[m][c][t] = i < npol ? (R)x[m * xstride + 3L * i + c] : R(0);
                sxJ[m][c][t] = j < npol ? (R)x[m * xstride + 3L * j + c] : R(0);
            }
    }
    __syncthreads();
    const int rg = t >> 4, cg = t & 15;
    // FP32 keeps the four source dipoles of the thread's columns in registers;
    // FP64 reads them from LDS (register budget).
    constexpr bool kCacheX = sizeof(R) == 4;
    R xb[kCacheX ? 4 : 1][M][3];
    if constexpr (kCacheX) {
#pragma unroll
        for (int bj = 0; bj < 4; bj++)
#pragma unroll
            for (int m = 0; m < M; m++)
#pragma unroll
                for (int c = 0; c < 3; c++) xb[bj][m][c] = sxJ[m][c][cg * 4 + bj];
    }
    R col[4][M][3];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) col[b][m][c] = R(0);
    // FP32: fully unrolled (all loads of the tile issued early, ~160 VGPRs);
    // FP64: one row at a time with the next row in flight (bounded registers).
    constexpr int kUnroll = sizeof(R) == 4 ? 4 : 1;
#pragma unroll kUnroll
    for (int ai = 0; ai < 4; ai++) {
        R4 nxt[6];
        if (ai < 3) {
#pragma unroll
            for (int e = 0; e < 6; e++) nxt[e] = tpack_load<R>(tv + ((ai + 1) * 6 + e) * kBlock);
        }
        const int a = rg * 4 + ai;
        R xa[M][3];
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) xa[m][c] = sxI[m][c][a];
        R row[M][3];
#pragma unroll
        for (int m = 0; m < M; m++) row[m][0] = row[m][1] = row[m][2] = R(0);
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
            const R txx = (&cur[0].x)[bj], txy = (&cur[1].x)[bj], txz = (&cur[2].x)[bj];
            const R tyy = (&cur[3].x)[bj], tyz = (&cur[4].x)[bj], tzz = (&cur[5].x)[bj];
#pragma unroll
            for (int m = 0; m < M; m++) {
                R xb0, xb1, xb2;
                if constexpr (kCacheX) {
                    xb0 = xb[bj][m][0];
                    xb1 = xb[bj][m][1];
                    xb2 = xb[bj][m][2];
                } else {
                    xb0 = sxJ[m][0][cg * 4 + bj];
                    xb1 = sxJ[m][1][cg * 4 + bj];
                    xb2 = sxJ[m][2][cg * 4 + bj];
                }
                row[m][0] = fmaa(txz, xb2, fmaa(txy, xb1, fmaa(txx, xb0, row[m][0])));
                row[m][1] = fmaa(tyz, xb2, fmaa(tyy, xb1, fmaa(txy, xb0, row[m][1])));
                row[m][2] = fmaa(tzz, xb2, fmaa(tyz, xb1, fmaa(txz, xb0, row[m][2])));
                col[bj][m][0] = fmaa(txz, xa[m][2], fmaa(txy, xa[m][1], fmaa(txx, xa[m][0], col[bj][m][0])));
                col[bj][m][1] = fmaa(tyz, xa[m][2], fmaa(tyy, xa[m][1], fmaa(txy, xa[m][0], col[bj][m][1])));
                col[bj][m][2] = fmaa(tzz, xa[m][2], fmaa(tyz, xa[m][1], fmaa(txz, xa[m][0], col[bj][m][2])));
            }
        }
        // row sum over the 16 contiguous lanes of the row group
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                R v = row[m][c];
                v += __shfl_xor(v, 8, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 1, 64);
                row[m][c] = v;
            }
        if (cg == 0) {
            const long tgt = (long)I * kTile + a;
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 3; c++) part[((long)J * M + m) * 3 * npad + 3 * tgt + c] = row[m][c];
        }
        if (ai < 3) {
#pragma unroll
            for (int e = 0; e < 6; e++) cur[e] = nxt[e];
        }
    }
    if (I == J) return;  // diagonal tiles hold both orders: rows only (uniform per block)
    // column sums: lanes cg, cg+16, cg+32, cg+48 of each wave, then across the 4 waves
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                R v = col[b][m][c];
                v += __shfl_xor(v, 16, 64);
                v += __shfl_xor(v, 32, 64);
                col[b][m][c] = v;
            }
    const int w = t >> 6, lane = t & 63;
    if (lane < 16)
        for (int b = 0; b < 4; b++)
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 3; c++) scol[w][m * 3 + c][lane * 4 + b] = col[b][m][c];
    __syncthreads();
    for (int idx = t; idx < M * 3 * kTile; idx += kBlock) {
        const int mc = idx / kTile, b = idx % kTile;
        const R v = (scol[0][mc][b] + scol[1][mc][b]) + (scol[2][mc][b] + scol[3][mc][b]);
        const int m = mc / 3, c = mc % 3;
        const long tgt = (long)J * kTile + b;
        part[((long)I * M + m) * 3 * npad + 3 * tgt + c] = v;
    }
}

// First reduction stage: source tiles of group blockIdx.y summed in order, in FP64.
template <typename R, int M>
__global__ __launch_bounds__(kBlock) void k_tpack_reduce(int npol, long nt, const R *__restrict__ part,
                                                         double *__restrict__ red, const CgState *__restrict__ skip,
                                                         int skipM) {
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= npol) return;
    const long npad = nt * kTile;
    const long chunk = (nt + kTpackGroups - 1) / kTpackGroups;
    const long s0 = blockIdx.y * chunk, s1 = s0 + chunk < nt ? s0 + chunk : nt;
    double acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    for (long s = s0; s < s1; s++)
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) acc[m][c] += (double)part[(s * M + m) * 3 * npad + 3L * j + c];
#pragma unroll
    for (int m = 0; m < M; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) red[((long)blockIdx.y * M + m) * 3 * npol + 3L * j + c] = acc[m][c];
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_tpack_finalize(int npol, const double *__restrict__ red,
                                                           const double *__restrict__ alpha,
                                                           const double *__restrict__ x, long xstride,
                                                           double *__restrict__ y, long ystride, int dodiag,
                                                           double *__restrict__ dotpart,
                                                           const CgState *__restrict__ skip, int skipM) {
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ double sh[4];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = j < npol;
    double acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    if (valid) {
        for (int g = 0; g < kTpackGroups; g++)
#pragma unroll
            for (int m = 0; m < M; m++)
#pragma unroll
                for (int c = 0; c < 3; c++) acc[m][c] += red[((long)g * M + m) * 3 * npol + 3L * j + c];
        if (dodiag) {
            const double ia = 1.0 / alpha[j];
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 3; c++) acc[m][c] += x[m * xstride + 3L * j + c] * ia;
        }
        for (int m = 0; m < M; m++)
            for (int c = 0; c < 3; c++) y[m * ystride + 3L * j + c] = acc[m][c];
    }
    if (dotpart) {
        for (int m = 0; m < M; m++) {
            double v = 0.0;
            if (valid)
                for (int c = 0; c < 3; c++) v += acc[m][c] * x[m * xstride + 3L * j + c];
            v = block_sum_all(v, sh);
            if (threadIdx.x == 0) dotpart[m * gridDim.x + blockIdx.x] = v;
        }
    }
}

void launch_tpack_build(hipStream_t st, int prec, bool amoeba, int npol, const double4 *ps64, const double *pa64,
                        void *tpack, const int *pd_ri, const int *pd_ci, const double *pd_s) {
    if (npol == 0) return;
    const long nt = (npol + kTile - 1) / kTile;
    const long np = nt * (nt + 1) / 2;
    if (prec == 32) {
        if (amoeba)
            hipLaunchKernelGGL((k_tpack_build<float, true>), dim3(np), dim3(kBlock), 0, st, npol, nt, ps64, pa64,
                               (float *)tpack);
        else
            hipLaunchKernelGGL((k_tpack_build<float, false>), dim3(np), dim3(kBlock), 0, st, npol, nt, ps64, pa64,
                               (float *)tpack);
        hipLaunchKernelGGL(k_tpack_scale<float>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nt, pd_ri,
                           pd_ci, pd_s, (float *)tpack);
    } else {
        if (amoeba)
            hipLaunchKernelGGL((k_tpack_build<double, true>), dim3(np), dim3(kBlock), 0, st, npol, nt, ps64, pa64,
                               (double *)tpack);
        else
            hipLaunchKernelGGL((k_tpack_build<double, false>), dim3(np), dim3(kBlock), 0, st, npol, nt, ps64,
                               pa64, (double *)tpack);
        hipLaunchKernelGGL(k_tpack_scale<double>, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, nt, pd_ri,
                           pd_ci, pd_s, (double *)tpack);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename R, int M>
static void tpack_mv_t(hipStream_t st, int npol, long nt, const void *tpack, const double *alpha, const double *x,
                       long xs, double *y, long ys, bool dodiag, double *dotpart, void *part,
                       const CgState *skip, hipEvent_t evb, hipEvent_t eve) {
    const long np = nt * (nt + 1) / 2;
    double *red = reinterpret_cast<double *>(static_cast<char *>(part) +
                                             (size_t)nt * M * 3 * (size_t)(nt * kTile) * 8);
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    hipLaunchKernelGGL((k_tpack_mv<R, M>), dim3(np), dim3(kBlock), 0, st, npol, nt, (const R *)tpack, x, xs,
                       (R *)part, skip, M);
    HIP_CHECK(hipGetLastError());
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
    hipLaunchKernelGGL((k_tpack_reduce<R, M>), dim3(site_blocks(npol), kTpackGroups), dim3(kBlock), 0, st, npol,
                       nt, (const R *)part, red, skip, M);
    hipLaunchKernelGGL((k_tpack_finalize<M>), dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, red, alpha, x,
                       xs, y, ys, (int)dodiag, dotpart, skip, M);
    HIP_CHECK(hipGetLastError());
}

void launch_tpack_mv(hipStream_t st, int prec, int npol, int M, const void *tpack, const double *alpha,
                     const double *x, long xstride, double *y, long ystride, bool dodiag, double *dotpart,
                     void *part, const int *skip_flag, hipEvent_t evb, hipEvent_t eve) {
    if (npol == 0) return;
    const CgState *skip = reinterpret_cast<const CgState *>(skip_flag);
    const long nt = (npol + kTile - 1) / kTile;
    if (prec == 32) {
        if (M == 1) tpack_mv_t<float, 1>(st, npol, nt, tpack, alpha, x, xstride, y, ystride, dodiag, dotpart, part, skip, evb, eve);
        else tpack_mv_t<float, 2>(st, npol, nt, tpack, alpha, x, xstride, y, ystride, dodiag, dotpart, part, skip, evb, eve);
    } else {
        if (M == 1) tpack_mv_t<double, 1>(st, npol, nt, tpack, alpha, x, xstride, y, ystride, dodiag, dotpart, part, skip, evb, eve);
        else tpack_mv_t<double, 2>(st, npol, nt, tpack, alpha, x, xstride, y, ystride, dodiag, dotpart, part, skip, evb, eve);
    }
}

// ============================================================ CG kernels
// Column m = blockIdx.y; element k of a column belongs to site k / 3.
__global__ __launch_bounds__(kBlock) void k_norm2(long n, const double *__restrict__ x, double *__restrict__ red) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = k < n ? x[m * n + k] * x[m * n + k] : 0.0;
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[m * gridDim.x + blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_guess_fix(long n, double *__restrict__ x, const double *__restrict__ alpha,
                                                      const double *__restrict__ red, int nred) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const double s = reduce_parts(red + m * nred, nred, sh);
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    // zero-norm guess: take the preconditioned guess alpha * x (mod_solvers.F90:151-159)
    if (s < kEpsRpDev && k < n) x[m * n + k] *= alpha[k / 3];
}

void launch_norm2_guess(hipStream_t st, long n, int M, double *x, const double *alpha, double *red) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_norm2, dim3(nb, M), dim3(kBlock), 0, st, n, x, red);
    hipLaunchKernelGGL(k_guess_fix, dim3(nb, M), dim3(kBlock), 0, st, n, x, alpha, red, nb);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_cg_init(long n, const double *__restrict__ b, const double *__restrict__ h,
                                                    double *__restrict__ x, double *__restrict__ r,
                                                    double *__restrict__ z, double *__restrict__ p,
                                                    const double *__restrict__ alpha, double *__restrict__ red,
                                                    int zero_guess) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (k < n) {
        const long o = m * n + k;
        double rr;
        if (zero_guess) {
            x[o] = 0.0;
            rr = b[o];
        } else {
            rr = b[o] - h[o];
        }
        const double zz = alpha[k / 3] * rr;
        r[o] = rr;
        z[o] = zz;
        p[o] = zz;
        v = rr * zz;
    }
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[m * gridDim.x + blockIdx.x] = v;
}

// pre (optional): columns with pre[m].done start done (nothing to solve).
__global__ __launch_bounds__(kBlock) void k_cg_init2(const double *__restrict__ red, int nred, CgState *cs,
                                                     const double *__restrict__ tol,
                                                     const CgState *__restrict__ pre) {
    __shared__ double sh[4];
    const int m = blockIdx.x;
    const double g = reduce_parts(red + m * nred, nred, sh);
    if (threadIdx.x == 0) {
        const bool skip = pre && pre[m].done;
        cs[m].g[0] = g;
        cs[m].g[1] = g;
        cs[m].rms = 0.0;
        cs[m].last_beta = 0.0;
        cs[m].tol = tol[m];
        cs[m].eps = tol[2 + m];
        cs[m].done = skip ? 1 : 0;
        cs[m].status = skip ? 0 : 2;
        cs[m].iters = 0;
    }
}

void launch_cg_init(hipStream_t st, long n, int M, const double *b, const double *h, double *x, double *r,
                    double *z, double *p, const double *alpha, double *red, CgState *cs, bool zero_guess,
                    const double *tol, const CgState *pre) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_cg_init, dim3(nb, M), dim3(kBlock), 0, st, n, b, h, x, r, z, p, alpha, red,
                       (int)zero_guess);
    hipLaunchKernelGGL(k_cg_init2, dim3(M), dim3(kBlock), 0, st, red, nb, cs, tol, pre);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_init2(hipStream_t st, int M, const double *red, int nred, CgState *cs, const double *tol,
                     const CgState *pre) {
    hipLaunchKernelGGL(k_cg_init2, dim3(M), dim3(kBlock), 0, st, red, nred, cs, tol, pre);
    HIP_CHECK(hipGetLastError());
}

// One refinement step's decision on the device: the FP64 residual rms of each
// column (rms[m]) marks the column converged (out[m].done) when below tol; a
// converged column stays converged (later residuals of it are not computed).
// The inner solve's tolerances go to inner_tol[0..1] (max(eta rms, frac tol)),
// its small-step thresholds inner_tol[2..3] = 0.
__global__ void k_refine_decide(int M, const double *__restrict__ rms, CgState *__restrict__ out, double tol,
                                double eta, double frac, double *__restrict__ inner_tol, int outer) {
    if (threadIdx.x != 0) return;
    for (int m = 0; m < M; m++) {
        if (outer == 0) {
            out[m].done = 0;
            out[m].status = 2;
        }
        if (!out[m].done) {
            out[m].rms = rms[m];
            out[m].iters = outer + 1;
            if (rms[m] < tol) {
                out[m].done = 1;
                out[m].status = 0;
            }
        }
        inner_tol[m] = out[m].done ? 0.0 : fmax(eta * out[m].rms, frac * tol);
        inner_tol[2 + m] = 0.0;
    }
    if (M == 1) {
        inner_tol[1] = inner_tol[0];
        inner_tol[3] = 0.0;
    }
}

void launch_refine_decide(hipStream_t st, int M, const double *rms, CgState *out, double tol, double eta,
                          double frac, double *inner_tol, int outer) {
    hipLaunchKernelGGL(k_refine_decide, dim3(1), dim3(64), 0, st, M, rms, out, tol, eta, frac, inner_tol, outer);
    HIP_CHECK(hipGetLastError());
}

// conjugate_gradient_solver loop body, first half (mod_solvers.F90:170-186).
__global__ __launch_bounds__(kBlock) void k_cg_update(long n, int it, const double *__restrict__ h,
                                                      double *__restrict__ x, double *__restrict__ r,
                                                      double *__restrict__ z, const double *__restrict__ p,
                                                      const double *__restrict__ alpha,
                                                      const double *__restrict__ red_hp, int nred_hp,
                                                      double *__restrict__ red_rz, CgState *cs) {
    __shared__ double sh[4];
    __shared__ int s_done;
    const int m = blockIdx.y;
    // one read per block: block 0 may flip the flag while this kernel runs
    if (threadIdx.x == 0) s_done = cs[m].done;
    __syncthreads();
    if (s_done) return;
    const double gama = reduce_parts(red_hp + m * nred_hp, nred_hp, sh);
    if (fabs(gama) < cs[m].eps || gama == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            cs[m].status = 1;
            cs[m].iters = it;
            cs[m].done = 1;
        }
        return;
    }
    const double gold = cs[m].g[(it + 1) & 1];
    const double a = gold / gama;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (k < n) {
        const long o = m * n + k;
        x[o] += a * p[o];
        const double rr = r[o] - a * h[o];
        const double zz = alpha[k / 3] * rr;
        r[o] = rr;
        z[o] = zz;
        v = rr * zz;
    }
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red_rz[m * gridDim.x + blockIdx.x] = v;
}

// second half: convergence test and new direction (mod_solvers.F90:187-204).
__global__ __launch_bounds__(kBlock) void k_cg_direction(long n, int it, const double *__restrict__ z,
                                                         double *__restrict__ p, const double *__restrict__ red_rz,
                                                         int nred, CgState *cs) {
    __shared__ double sh[4];
    __shared__ int s_done;
    const int m = blockIdx.y;
    if (threadIdx.x == 0) s_done = cs[m].done;
    __syncthreads();
    if (s_done) return;
    const double gnew = reduce_parts(red_rz + m * nred, nred, sh);
    const double rms = sqrt(gnew / (double)n);
    if (rms < cs[m].tol) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            cs[m].rms = rms;
            cs[m].status = 0;
            cs[m].iters = it;
            cs[m].done = 1;
        }
        return;
    }
    const double beta = gnew / cs[m].g[(it + 1) & 1];
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) p[m * n + k] = beta * p[m * n + k] + z[m * n + k];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cs[m].g[it & 1] = gnew;
        cs[m].rms = rms;
        cs[m].last_beta = beta;
        cs[m].iters = it;
    }
}

void launch_cg_update(hipStream_t st, long n, int M, int it, const double *h, double *x, double *r, double *z,
                      const double *p, const double *alpha, const double *red_hp, int nred_hp, double *red_rz,
                      CgState *cs) {
    hipLaunchKernelGGL(k_cg_update, dim3(vec_blocks(n), M), dim3(kBlock), 0, st, n, it, h, x, r, z, p, alpha,
                       red_hp, nred_hp, red_rz, cs);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_direction(hipStream_t st, long n, int M, int it, const double *z, double *p, const double *red_rz,
                         CgState *cs, int *) {
    hipLaunchKernelGGL(k_cg_direction, dim3(vec_blocks(n), M), dim3(kBlock), 0, st, n, it, z, p, red_rz,
                       vec_blocks(n), cs);
    HIP_CHECK(hipGetLastError());
}

// ========================================================= element-wise
__global__ void k_add_field(long n, int nsets, const double *__restrict__ e_m2d, const double *__restrict__ ext,
                            double *__restrict__ ef) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double v = ext ? ext[k] : 0.0;
    for (int s = 0; s < nsets; s++) ef[s * n + k] = (e_m2d ? e_m2d[s * n + k] : 0.0) + v;
}

void launch_add_field(hipStream_t st, long n, int nsets, const double *e_m2d, const double *ext, double *ef) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_add_field, dim3(vec_blocks(n)), dim3(kBlock), 0, st, n, nsets, e_m2d, ext, ef);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_copy(long n, const double *__restrict__ s, double *__restrict__ d) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) d[k] = s[k];
}

void launch_copy(hipStream_t st, long n, const double *src, double *dst) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy, dim3(vec_blocks(n)), dim3(kBlock), 0, st, n, src, dst);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_axpy(long n, double a, const double *__restrict__ x, double *__restrict__ y) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    const int m = blockIdx.y;
    if (k < n) y[m * n + k] += a * x[m * n + k];
}

void launch_axpy(hipStream_t st, long n, int M, double a, const double *x, double *y) {
    hipLaunchKernelGGL(k_axpy, dim3(vec_blocks(n), M), dim3(kBlock), 0, st, n, a, x, y);
    HIP_CHECK(hipGetLastError());
}

// r = b - tx; partial sums of alpha r^2 (the CG convergence norm)
__global__ __launch_bounds__(kBlock) void k_resid(long n, const double *__restrict__ b, const double *__restrict__ tx,
                                                  double *__restrict__ r, const double *__restrict__ alpha,
                                                  double *__restrict__ red) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (k < n) {
        const double rr = b[m * n + k] - tx[m * n + k];
        r[m * n + k] = rr;
        v = alpha[k / 3] * rr * rr;
    }
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[m * gridDim.x + blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_rms_final(long n, const double *__restrict__ red, int nred,
                                                      double *__restrict__ out) {
    __shared__ double sh[4];
    const int m = blockIdx.x;
    const double s = reduce_parts(red + m * nred, nred, sh);
    if (threadIdx.x == 0) out[m] = sqrt(s / (double)n);
}

void launch_residual_rms(hipStream_t st, long n, int M, const double *b, const double *tx, double *r,
                         const double *alpha, double *red, double *rms_out) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_resid, dim3(nb, M), dim3(kBlock), 0, st, n, b, tx, r, alpha, red);
    hipLaunchKernelGGL(k_rms_final, dim3(M), dim3(kBlock), 0, st, n, red, nb, rms_out);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_dot_part(long n, const double *__restrict__ a, const double *__restrict__ b,
                                                     double *__restrict__ red) {
    __shared__ double sh[4];
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = k < n ? a[k] * b[k] : 0.0;
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double *__restrict__ red, int nred, double scale,
                                                      double *__restrict__ out) {
    __shared__ double sh[4];
    const double s = reduce_parts(red, nred, sh);
    if (threadIdx.x == 0) *out = scale * s;
}

void launch_dot(hipStream_t st, long n, const double *a, const double *b, double *red, double *out) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kBlock), 0, st, n, a, b, red);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, red, nb, 1.0, out);
    HIP_CHECK(hipGetLastError());
}

// energy_MM_pol: -1/2 sum mu . E  (mod_electrostatics.F90:943-976)
void launch_energy(hipStream_t st, long n, const double *ipd_d, const double *e_p, double *red, double *out) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kBlock), 0, st, n, ipd_d, e_p, red);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, red, nb, -0.5, out);
    HIP_CHECK(hipGetLastError());
}

// ============================================== fixed multipole energy
// Interaction of the multipoles of target j with the potential, field and field
// gradient of source i, undamped (the contraction of energy_MM_MM,
// mod_electrostatics.F90:915-933, applied to the do_V / do_E / do_grdE props of
// q_elec_prop, mu_elec_prop and quad_elec_prop, :1119-1141, 1197-1223,
// 1307-1345): q_j V_i - mu_j . E_i + Theta_j : G_i with the off-diagonal
// gradient entries counted twice.  d = r_j - r_i, ir = 1 / |d|.  The gradient
// of every source term has the form g_dd d d^T + g_I I - 3 k3 (mu d^T + d mu^T)
// - 30 k4 (Theta d d^T + d d^T Theta) + 6 k3 Theta, so its contraction with
// Theta_j needs d.Theta_j.d, tr Theta_j and Theta_j d only.
template <bool AMOEBA>
__device__ __forceinline__ double mm_pair_energy(double dx, double dy, double dz, double ir, const double *qi,
                                                 const double *qj) {
    const double k1 = ir;
    if (!AMOEBA) return qj[0] * qi[0] * k1;
    const double ir2 = ir * ir;
    const double k2 = k1 * ir2, k3 = k2 * ir2, k4 = k3 * ir2, k5 = k4 * ir2;
    const double mdi = qi[1] * dx + qi[2] * dy + qi[3] * dz;
    const double qix = qi[4] * dx + qi[5] * dy + qi[7] * dz;
    const double qiy = qi[5] * dx + qi[6] * dy + qi[8] * dz;
    const double qiz = qi[7] * dx + qi[8] * dy + qi[9] * dz;
    const double qrri = qix * dx + qiy * dy + qiz * dz;
    const double mdj = qj[1] * dx + qj[2] * dy + qj[3] * dz;
    const double qjx = qj[4] * dx + qj[5] * dy + qj[7] * dz;
    const double qjy = qj[5] * dx + qj[6] * dy + qj[8] * dz;
    const double qjz = qj[7] * dx + qj[8] * dy + qj[9] * dz;
    const double qrrj = qjx * dx + qjy * dy + qjz * dz;
    const double trj = qj[4] + qj[6] + qj[9];
    const double V = qi[0] * k1 + mdi * k2 + 3.0 * qrri * k3;
    const double cd = qi[0] * k2 + 3.0 * mdi * k3 + 15.0 * qrri * k4;
    const double muE = cd * mdj - k2 * (qj[1] * qi[1] + qj[2] * qi[2] + qj[3] * qi[3]) -
                       6.0 * k3 * (qj[1] * qix + qj[2] * qiy + qj[3] * qiz);
    const double g_dd = 3.0 * qi[0] * k3 + 15.0 * mdi * k4 + 105.0 * qrri * k5;
    const double g_I = -qi[0] * k2 - 3.0 * mdi * k3 - 15.0 * qrri * k4;
    const double tq = qi[4] * qj[4] + qi[6] * qj[6] + qi[9] * qj[9] +
                      2.0 * (qi[5] * qj[5] + qi[7] * qj[7] + qi[8] * qj[8]);
    const double thG = g_dd * qrrj + g_I * trj - 6.0 * k3 * (qi[1] * qjx + qi[2] * qjy + qi[3] * qjz) -
                       60.0 * k4 * (qix * qjx + qiy * qjy + qiz * qjz) + 6.0 * k3 * tq;
    return qj[0] * V - muE + thG;
}

// One thread per MM target j, the sources of its split staged through LDS;
// part[split][nmm] receives the unscreened sum over i != j.  Two distinct
// sites closer than eps_rp set *bad (the reference's coulomb_kernel stops there).
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_mm_energy_pairs(int nmm, int chunk, const double4 *__restrict__ ms,
                                                            const double *__restrict__ mq, double *__restrict__ part,
                                                            int *__restrict__ bad) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_q[LD][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < nmm;
    const double4 pj = valid ? ms[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    double qj[LD];
#pragma unroll
    for (int c = 0; c < LD; c++) qj[c] = valid ? mq[(long)LD * j + c] : 0.0;
    double acc = 0.0;
    bool clash = false;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(nmm, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            s_pos[tid] = ms[base + tid];
#pragma unroll
            for (int c = 0; c < LD; c++) s_q[c][tid] = mq[(long)LD * (base + tid) + c];
        }
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const double4 pi = s_pos[k];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool self = (base + k) == j;
            const bool close = !self && r2 < kEpsRpDev * kEpsRpDev;
            clash = clash || close;
            const double ir = (self || close) ? 0.0 : rsqrt(r2);
            double qi[LD];
#pragma unroll
            for (int c = 0; c < LD; c++) qi[c] = s_q[c][k];
            acc += mm_pair_energy<AMOEBA>(dx, dy, dz, ir, qi, qj);
        }
    }
    if (valid) {
        part[(long)blockIdx.y * nmm + j] = acc;
        if (clash) *bad = 1;
    }
}

// Sum of the splits in fixed order, the listed pairs' (1 - s) share removed
// (f = 1 - s per entry of the S-S list by target; s = 0 removes the pair), and
// per-block partial sums of the per-site energies.
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_mm_energy_finalize(int nmm, int nsplit, const double *__restrict__ part,
                                                               const double4 *__restrict__ ms,
                                                               const double *__restrict__ mq,
                                                               const int *__restrict__ ri, const int *__restrict__ ci,
                                                               const double *__restrict__ f,
                                                               double *__restrict__ red) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double sh[4];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    double e = 0.0;
    if (j < nmm) {
        for (int s = 0; s < nsplit; s++) e += part[(long)s * nmm + j];
        const double4 pj = ms[j];
        double qj[LD];
        for (int c = 0; c < LD; c++) qj[c] = mq[(long)LD * j + c];
        for (int k = ri[j]; k < ri[j + 1]; k++) {
            const int i = ci[k];
            const double4 pi = ms[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            double qi[LD];
            for (int c = 0; c < LD; c++) qi[c] = mq[(long)LD * i + c];
            e -= f[k] * mm_pair_energy<AMOEBA>(dx, dy, dz, rsqrt(dx * dx + dy * dy + dz * dz), qi, qj);
        }
    }
    e = block_sum_all(e, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = e;
}

void launch_mm_energy(hipStream_t st, bool amoeba, int nmm, const double4 *ms64, const double *mq, const CsrDev &ss,
                      double *part, int nsplit, double *red, int *bad, double *out) {
    if (nmm == 0) return;
    const int chunk = (nmm + nsplit - 1) / nsplit;
    const int nb = site_blocks(nmm);
    dim3 grid(nb, nsplit);
    if (amoeba) {
        hipLaunchKernelGGL(k_mm_energy_pairs<true>, grid, dim3(kBlock), 0, st, nmm, chunk, ms64, mq, part, bad);
        hipLaunchKernelGGL(k_mm_energy_finalize<true>, dim3(nb), dim3(kBlock), 0, st, nmm, nsplit, part, ms64, mq,
                           ss.ri, ss.ci, ss.f, red);
    } else {
        hipLaunchKernelGGL(k_mm_energy_pairs<false>, grid, dim3(kBlock), 0, st, nmm, chunk, ms64, mq, part, bad);
        hipLaunchKernelGGL(k_mm_energy_finalize<false>, dim3(nb), dim3(kBlock), 0, st, nmm, nsplit, part, ms64, mq,
                           ss.ri, ss.ci, ss.f, red);
    }
    // energy_MM_MM: every pair is seen from both ends, hence 1/2
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, st, red, nb, 0.5, out);
    HIP_CHECK(hipGetLastError());
}

// ====================================================== Jacobi / DIIS
// x_new = alpha (rhs - O x); err = x_new - x  (mod_solvers.F90:320-333)
__global__ void k_jacobi(long n, const double *__restrict__ rhs, const double *__restrict__ ox,
                         const double *__restrict__ x, const double *__restrict__ alpha, double *__restrict__ xnew,
                         double *__restrict__ err) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double v = alpha[k / 3] * (rhs[k] - ox[k]);
    xnew[k] = v;
    err[k] = v - x[k];
}

void launch_jacobi(hipStream_t st, long n, const double *rhs, const double *ox, const double *x, const double *alpha,
                   double *xnew, double *err) {
    hipLaunchKernelGGL(k_jacobi, dim3(vec_blocks(n)), dim3(kBlock), 0, st, n, rhs, ox, x, alpha, xnew, err);
    HIP_CHECK(hipGetLastError());
}

struct PtrList {
    const double *p[24];
};

__global__ __launch_bounds__(kBlock) void k_multi_dot(long n, int k, PtrList v, const double *__restrict__ e,
                                                      double *__restrict__ red) {
    __shared__ double sh[4];
    const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
    const double ev = idx < n ? e[idx] : 0.0;
    for (int q = 0; q < k; q++) {
        double s = idx < n ? ev * v.p[q][idx] : 0.0;
        s = block_sum_all(s, sh);
        if (threadIdx.x == 0) red[q * gridDim.x + blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_multi_final(const double *__restrict__ red, int nred, double *__restrict__ out) {
    __shared__ double sh[4];
    const int q = blockIdx.x;
    const double s = reduce_parts(red + q * nred, nred, sh);
    if (threadIdx.x == 0) out[q] = s;
}

void launch_multi_dot(hipStream_t st, long n, int k, const double *const *vecs, const double *e, double *red,
                      double *out) {
    PtrList pl{};
    for (int q = 0; q < k; q++) pl.p[q] = vecs[q];
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_multi_dot, dim3(nb), dim3(kBlock), 0, st, n, k, pl, e, red);
    hipLaunchKernelGGL(k_multi_final, dim3(k), dim3(kBlock), 0, st, red, nb, out);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_lincomb(long n, int k, PtrList v, const double *__restrict__ coef, double *__restrict__ out) {
    const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n) return;
    double s = 0.0;
    for (int q = 0; q < k; q++) s += coef[q] * v.p[q][idx];
    out[idx] = s;
}

void launch_lincomb(hipStream_t st, long n, int k, const double *const *vecs, const double *coef, double *out) {
    PtrList pl{};
    for (int q = 0; q < k; q++) pl.p[q] = vecs[q];
    hipLaunchKernelGGL(k_lincomb, dim3(vec_blocks(n)), dim3(kBlock), 0, st, n, k, pl, coef, out);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_maxabs_part(long n, const double *__restrict__ a, const double *__restrict__ b,
                                                        double *__restrict__ red) {
    __shared__ double sh[4];
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double v = k < n ? fabs(a[k] - b[k]) : 0.0;
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) red[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ __launch_bounds__(kBlock) void k_max_final(const double *__restrict__ red, int nred, double *__restrict__ out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < nred; i += kBlock) v = fmax(v, red[i]);
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

void launch_maxabs_diff(hipStream_t st, long n, const double *a, const double *b, double *red, double *out) {
    const int nb = vec_blocks(n);
    hipLaunchKernelGGL(k_maxabs_part, dim3(nb), dim3(kBlock), 0, st, n, a, b, red);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kBlock), 0, st, red, nb, out);
    HIP_CHECK(hipGetLastError());
}

// ============================================== dense T and LU (inversion)
// T_ii = I / alpha_i, T_ij = k2 I - 3 k3 d d^T (dipole_T, mod_polarization.F90:327-387),
// row-major n x n with n = 3 npol.
template <bool AMOEBA>
__global__ void k_tdense(int npol, const double4 *__restrict__ ps, const double *__restrict__ pa,
                         const double *__restrict__ alpha, double *__restrict__ T) {
    const int j = blockIdx.x * 16 + (threadIdx.x & 15);
    const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= npol || j >= npol) return;
    const long n = 3L * npol;
    double v[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (i == j) {
        for (int c = 0; c < 3; c++) v[c][c] = 1.0 / alpha[i];
    } else {
        const double4 pi = ps[i], pj = ps[j];
        const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double k2, k3;
        dip_kernel<double, AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, pa[i], pa[j], k2, k3);
        const double d[3] = {dx, dy, dz};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) v[a][b] = (a == b ? k2 : 0.0) - k3 * d[a] * d[b];
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) T[(3L * i + a) * n + 3L * j + b] = v[a][b];
}

__global__ void k_tdense_scale(int npol, const int *__restrict__ ri, const int *__restrict__ ci,
                               const double *__restrict__ sv, double *__restrict__ T) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npol) return;
    const long n = 3L * npol;
    for (int k = ri[i]; k < ri[i + 1]; k++) {
        const int j = ci[k];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) T[(3L * i + a) * n + 3L * j + b] *= sv[k];
    }
}

void launch_tdense_build(hipStream_t st, bool amoeba, int npol, const double4 *ps64, const double *pa64,
                         const double *alpha, const int *pd_ri, const int *pd_ci, const double *pd_s, double *T) {
    dim3 grid((npol + 15) / 16, (npol + 15) / 16);
    if (amoeba)
        hipLaunchKernelGGL(k_tdense<true>, grid, dim3(256), 0, st, npol, ps64, pa64, alpha, T);
    else
        hipLaunchKernelGGL(k_tdense<false>, grid, dim3(256), 0, st, npol, ps64, pa64, alpha, T);
    hipLaunchKernelGGL(k_tdense_scale, dim3(site_blocks(npol)), dim3(kBlock), 0, st, npol, pd_ri, pd_ci, pd_s, T);
    HIP_CHECK(hipGetLastError());
}

// LU with partial pivoting, right-looking, one column per step (dgetrf semantics).
__global__ __launch_bounds__(kBlock) void k_lu_pivot(long n, long k, double *__restrict__ A, int *__restrict__ piv,
                                                     int *__restrict__ info) {
    __shared__ double sv[kBlock];
    __shared__ long si[kBlock];
    double best = -1.0;
    long bi = k;
    for (long i = k + threadIdx.x; i < n; i += kBlock) {
        const double v = fabs(A[i * n + k]);
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = sv[threadIdx.x + s];
            const long oi = si[threadIdx.x + s];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const long p = si[0];
    if (threadIdx.x == 0) {
        piv[k] = (int)p;
        if (A[p * n + k] == 0.0 && *info == 0) *info = (int)k + 1;
    }
    if (p != k)
        for (long j = threadIdx.x; j < n; j += kBlock) {
            const double t = A[k * n + j];
            A[k * n + j] = A[p * n + j];
            A[p * n + j] = t;
        }
}

__global__ void k_lu_scale(long n, long k, double *__restrict__ A) {
    const long i = k + 1 + (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double d = A[k * n + k];
    if (d != 0.0) A[i * n + k] /= d;
}

__global__ void k_lu_update(long n, long k, double *__restrict__ A) {
    const long j = k + 1 + (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const long i = k + 1 + (long)blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    A[i * n + j] -= A[i * n + k] * A[k * n + j];
}

// Forward/back substitution on one right-hand side, one block.
__global__ __launch_bounds__(kBlock) void k_lu_solve(long n, const double *__restrict__ A, const int *__restrict__ piv,
                                                     double *__restrict__ b) {
    if (threadIdx.x == 0)
        for (long k = 0; k < n; k++)
            if (piv[k] != k) {
                const double t = b[k];
                b[k] = b[piv[k]];
                b[piv[k]] = t;
            }
    __syncthreads();
    for (long j = 0; j < n; j++) {  // L y = b (unit diagonal)
        const double bj = b[j];
        for (long i = j + 1 + threadIdx.x; i < n; i += kBlock) b[i] -= A[i * n + j] * bj;
        __syncthreads();
    }
    for (long j = n - 1; j >= 0; j--) {  // U x = y
        if (threadIdx.x == 0) b[j] /= A[j * n + j];
        __syncthreads();
        const double bj = b[j];
        for (long i = threadIdx.x; i < j; i += kBlock) b[i] -= A[i * n + j] * bj;
        __syncthreads();
    }
}

int launch_lu_factor(hipStream_t st, long n, double *A, int *piv, double *work) {
    int *info = reinterpret_cast<int *>(work);
    HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int), st));
    for (long k = 0; k < n; k++) {
        hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(kBlock), 0, st, n, k, A, piv, info);
        const long rem = n - k - 1;
        if (rem > 0) {
            hipLaunchKernelGGL(k_lu_scale, dim3((rem + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, k, A);
            hipLaunchKernelGGL(k_lu_update, dim3((rem + 63) / 64, (rem + 3) / 4), dim3(256), 0, st, n, k, A);
        }
    }
    HIP_CHECK(hipGetLastError());
    int h_info = 0;
    HIP_CHECK(hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return h_info;
}

void launch_lu_apply(hipStream_t st, long n, const double *A, const int *piv, double *rhs) {
    hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(kBlock), 0, st, n, A, piv, rhs);
    HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------- vector FMA rate
// Eight independent FMA chains per thread over enough workgroups to fill every
// SIMD, built like every kernel here without packed FP32 instructions: the rate
// that non-packed vector code reaches on this device, the reference point of
// the pair kernels, whose arithmetic is mostly such FMAs.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fma_rate(T *__restrict__ out, int iters, T m, T c) {
    T a0 = (T)threadIdx.x * T(1e-7) + T(0.1), a1 = a0 + T(0.01), a2 = a0 + T(0.02), a3 = a0 + T(0.03);
    T a4 = a0 + T(0.04), a5 = a0 + T(0.05), a6 = a0 + T(0.06), a7 = a0 + T(0.07);
    for (int i = 0; i < iters; i++) {
        a0 = fmaa(a0, m, c);
        a1 = fmaa(a1, m, c);
        a2 = fmaa(a2, m, c);
        a3 = fmaa(a3, m, c);
        a4 = fmaa(a4, m, c);
        a5 = fmaa(a5, m, c);
        a6 = fmaa(a6, m, c);
        a7 = fmaa(a7, m, c);
    }
    out[(long)blockIdx.x * kBlock + threadIdx.x] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

template <typename T>
static double fma_rate_t() {
    constexpr int kBlocks = 8192, kIters = 20000, kReps = 5;
    T *out = nullptr;
    HIP_CHECK(hipMalloc((void **)&out, sizeof(T) * kBlocks * kBlock));
    hipStream_t st;
    HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(k_fma_rate<T>, dim3(kBlocks), dim3(kBlock), 0, st, out, kIters, T(0.999), T(0.001));
    HIP_CHECK(hipEventRecord(a, st));
    for (int r = 0; r < kReps; r++)
        hipLaunchKernelGGL(k_fma_rate<T>, dim3(kBlocks), dim3(kBlock), 0, st, out, kIters, T(0.999), T(0.001));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(b, st));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    HIP_CHECK(hipEventDestroy(a));
    HIP_CHECK(hipEventDestroy(b));
    HIP_CHECK(hipStreamDestroy(st));
    HIP_CHECK(hipFree(out));
    const double flop = 2.0 * 8.0 * kIters * (double)kBlock * kBlocks * kReps;
    return ms > 0.f ? flop / (ms * 1e-3) / 1e12 : 0.0;
}

double vector_fma_tflops(int prec) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        fatal("No GPU available: this library runs the polarization solver on an AMD Instinct GPU only");
    return prec == 32 ? fma_rate_t<float>() : fma_rate_t<double>();
}

}  // namespace ommp
