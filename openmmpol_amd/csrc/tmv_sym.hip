// T matvec on the fly with each unordered pair evaluated once (T_ij = T_ji):
// the geometry and damping of a pair (half of the per-pair work) are shared by
// its two directions.  One workgroup per pair of 256-site blocks (I <= J),
// row-major over the upper triangle (pairs[] lists them); the body is
// sym_block_pair (tmv_sym.h).
//
// Outputs, per block pair p: target side part_i[p][M * 3][256] (sites of I) and
// source side part_j[p][M * 3][256] (sites of J); k_tmv_sym_finalize adds, for
// every site, the nb + 1 partials of its row and column in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "pair_math.h"
#include "tmv_sym.h"

namespace ommp {


static_assert(kSymBlock == kBlock, "pair-once blocks are the blocks of the per-site kernels");

template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kSymBlock) void k_tmv_sym(int npol, const int2 *__restrict__ pairs,
                                                       const typename V4<R>::t *__restrict__ site,
                                                       const R *__restrict__ adamp, const double *__restrict__ x,
                                                       long xstride, R *__restrict__ part_i, R *__restrict__ part_j,
                                                       const CgState *__restrict__ skip, int skipM) {
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ __align__(16) char lds[sym_lds_bytes<R, M>()];
    const int2 pr = pairs[blockIdx.x];
    const double *xcol[M];
    R *oi[M], *oj[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        xcol[m] = x + m * xstride;
        oi[m] = part_i + ((long)blockIdx.x * M * 3 + 3 * m) * kSymBlock;
        oj[m] = part_j + ((long)blockIdx.x * M * 3 + 3 * m) * kSymBlock;
    }
    sym_block_pair_any<R, M, AMOEBA>(npol, pr.x, pr.y, site, adamp, xcol, oi, oj, lds);
}

// y = -(row and column block-pair partials) + list_P_P corrections (+ x / alpha),
// optional y.x partials.  One workgroup per 64 sites (kSymSub per block of
// kSymBlock), kSymQ groups of 64 threads each: group q sums every kSymQ-th
// partial of the site's row (pairs (b, J >= b): target sides) and column (pairs
// (I <= b, b): source sides) and every kSymQ-th entry of its correction list;
// group 0 (one wave) adds the group results in a fixed order.  Spreading the
// 2 nb partials of a site over 8 threads and the blocks over 4 workgroups keeps
// each thread's chain of dependent loads short and puts the reads of a 15,000-site
// system on ~240 CUs (one workgroup of 4 x 256 threads per block used 59).
constexpr int kSymQ = 8;
constexpr int kSymSub = kSymBlock / 64;

template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kSymQ * 64) void k_tmv_sym_finalize(
    int npol, int nb, const R *__restrict__ part, const double4 *__restrict__ ps, const double *__restrict__ pa,
    const double *__restrict__ alpha, const int *__restrict__ pp_ri, const int *__restrict__ pp_ci,
    const double *__restrict__ pp_f, const double *__restrict__ x, long xstride, double *__restrict__ y,
    long ystride, int dodiag, double *__restrict__ dotpart, const CgState *__restrict__ skip, int skipM) {
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ R shs[kSymQ - 1][M * 3][64];
    __shared__ double shc[kSymQ - 1][M * 3][64];
    const int q = threadIdx.x >> 6, li = threadIdx.x & 63;
    const int b = blockIdx.x / kSymSub, t = (blockIdx.x % kSymSub) * 64 + li;  // t: site within block b
    const int j = b * kSymBlock + t;
    const bool valid = j < npol;
    const long np = (long)nb * (nb + 1) / 2, stride = (long)M * 3 * kSymBlock;
    R sum[M * 3];
    double corr[M * 3];
#pragma unroll
    for (int c = 0; c < M * 3; c++) {
        sum[c] = R(0);
        corr[c] = 0.0;
    }
    if (valid) {
        // pair (I, J) has index I nb - I (I - 1) / 2 + J - I
        const long row = (long)b * nb - (long)b * (b - 1) / 2;
        constexpr int kUnroll = sizeof(R) == 4 ? 4 : 2;
#pragma unroll kUnroll
        for (int J = b + q; J < nb; J += kSymQ) {
            const R *p = part + (row + J - b) * stride + t;
#pragma unroll
            for (int c = 0; c < M * 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
        const R *part_j = part + np * stride;
#pragma unroll kUnroll
        for (int I = q; I <= b; I += kSymQ) {
            const R *p = part_j + ((long)I * nb - (long)I * (I - 1) / 2 + b - I) * stride + t;
#pragma unroll
            for (int c = 0; c < M * 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
        const double4 pj = ps[j];
        const double aj = pa[j];
        for (int k = pp_ri[j] + q; k < pp_ri[j + 1]; k += kSymQ) {
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
                corr[3 * m + 0] += f * (c * dx - k2 * xi[0]);
                corr[3 * m + 1] += f * (c * dy - k2 * xi[1]);
                corr[3 * m + 2] += f * (c * dz - k2 * xi[2]);
            }
        }
    }
    if (q > 0) {
#pragma unroll
        for (int c = 0; c < M * 3; c++) {
            shs[q - 1][c][li] = sum[c];
            shc[q - 1][c][li] = corr[c];
        }
    }
    __syncthreads();
    if (q != 0) return;  // wave 0 from here on
    double v[M];
#pragma unroll
    for (int m = 0; m < M; m++) v[m] = 0.0;
    if (valid) {
        double acc[M * 3];
#pragma unroll
        for (int c = 0; c < M * 3; c++) {
            R s = sum[c];
            double u = corr[c];
#pragma unroll
            for (int g = 0; g < kSymQ - 1; g++) {
                s += shs[g][c][li];
                u += shc[g][c][li];
            }
            acc[c] = u - (double)s;
        }
        if (dodiag) {
            const double ia = 1.0 / alpha[j];
#pragma unroll
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 3; c++) acc[3 * m + c] += x[m * xstride + 3L * j + c] * ia;
        }
#pragma unroll
        for (int m = 0; m < M; m++)
            for (int c = 0; c < 3; c++) {
                y[m * ystride + 3L * j + c] = acc[3 * m + c];
                v[m] += acc[3 * m + c] * x[m * xstride + 3L * j + c];
            }
    }
    if (!dotpart) return;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const double s = wave_sum(v[m]);
        if (li == 0) dotpart[m * gridDim.x + blockIdx.x] = s;
    }
}

template <typename R, int M>
static void sym_finalize_t(hipStream_t st, bool amoeba, int npol, const R *part, const double4 *ps64,
                           const double *pa64, const double *alpha, const CsrDev &pp, const double *x, long xs,
                           double *y, long ys, bool dodiag, double *dotpart, const CgState *skip, int skipM) {
    const int nb = tmv_sym_blocks(npol);
    const dim3 grid(nb * kSymSub), block(kSymQ * 64);
    if (amoeba)
        hipLaunchKernelGGL((k_tmv_sym_finalize<R, M, true>), grid, block, 0, st, npol, nb, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xs, y, ys, (int)dodiag, dotpart, skip, skipM);
    else
        hipLaunchKernelGGL((k_tmv_sym_finalize<R, M, false>), grid, block, 0, st, npol, nb, part, ps64, pa64, alpha,
                           pp.ri, pp.ci, pp.f, x, xs, y, ys, (int)dodiag, dotpart, skip, skipM);
    HIP_CHECK(hipGetLastError());
}

int tmv_sym_blocks(int npol) { return (npol + kSymBlock - 1) / kSymBlock; }
int tmv_sym_dot_parts(int npol) { return tmv_sym_blocks(npol) * kSymSub; }

// The pair-once sweep evaluates half the pairs, at the price of a workgroup per
// block pair and of its block-pair partial sums; below a few thousand sites it
// has too few workgroups to fill the chip.  Crossovers against the per-target
// sweep measured on MI355X (whole matvec, two targets per lane): FP32 two
// columns ~11,000 sites (-5% at 12,000, -22% at 15,000), one column ~7,000;
// FP64 two columns ~7,000 (-8% at 8,000), one column ~6,000 (-19% at 8,000).
// Used from those sizes while its partial sums for max(ncols, 2) columns stay
// under max_bytes (2 GiB: up to ~100,000 sites).
bool tmv_pair_once(int npol, int prec, int ncols, size_t max_bytes) {
    const int from = prec == 32 ? (ncols >= 2 ? 11000 : 7000) : (ncols >= 2 ? 7000 : 6000);
    if (npol < from) return false;
    return tmv_sym_part_bytes(npol, ncols > 2 ? ncols : 2, 64) <= max_bytes;
}

size_t tmv_sym_part_bytes(int npol, int M, int prec) {
    const long nb = tmv_sym_blocks(npol), np = nb * (nb + 1) / 2;
    return (size_t)(2 * np * M * 3 * kSymBlock) * (prec == 32 ? 4 : 8);  // part_i + part_j
}

// evb / eve bracket the pair kernel only.
template <typename R, int M>
static void sym_launch_t(hipStream_t st, bool amoeba, int npol, const int2 *pairs, const typename V4<R>::t *site,
                         const R *ad, const double *x, long xs, R *part, const CgState *skip, int skipM,
                         hipEvent_t evb, hipEvent_t eve) {
    const int nb = tmv_sym_blocks(npol);
    const long np = (long)nb * (nb + 1) / 2;
    R *part_i = part, *part_j = part + np * M * 3 * kSymBlock;
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    if (amoeba)
        hipLaunchKernelGGL((k_tmv_sym<R, M, true>), dim3((unsigned)np), dim3(kSymBlock), 0, st, npol, pairs, site,
                           ad, x, xs, part_i, part_j, skip, skipM);
    else
        hipLaunchKernelGGL((k_tmv_sym<R, M, false>), dim3((unsigned)np), dim3(kSymBlock), 0, st, npol, pairs, site,
                           ad, x, xs, part_i, part_j, skip, skipM);
    HIP_CHECK(hipGetLastError());
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
}

void launch_tmv_sym(hipStream_t st, int prec, bool amoeba, int npol, int M, const int2 *pairs, const double4 *ps64,
                    const double *pa64, const float4 *ps32, const float *pa32, const double *alpha,
                    const CsrDev &pp, const double *x, long xstride, double *y, long ystride, bool dodiag,
                    double *dotpart, void *part, const int *skip_flag, hipEvent_t evb, hipEvent_t eve) {
    if (npol == 0) return;
    const CgState *skip = reinterpret_cast<const CgState *>(skip_flag);
    if (prec == 32) {
        float *pf = (float *)part;
        if (M == 1) {
            sym_launch_t<float, 1>(st, amoeba, npol, pairs, ps32, pa32, x, xstride, pf, skip, M, evb, eve);
            sym_finalize_t<float, 1>(st, amoeba, npol, pf, ps64, pa64, alpha, pp, x, xstride, y, ystride, dodiag,
                                     dotpart, skip, M);
        } else {
            sym_launch_t<float, 2>(st, amoeba, npol, pairs, ps32, pa32, x, xstride, pf, skip, M, evb, eve);
            sym_finalize_t<float, 2>(st, amoeba, npol, pf, ps64, pa64, alpha, pp, x, xstride, y, ystride, dodiag,
                                     dotpart, skip, M);
        }
    } else {
        double *pd = (double *)part;
        if (M == 1) {
            sym_launch_t<double, 1>(st, amoeba, npol, pairs, ps64, pa64, x, xstride, pd, skip, M, evb, eve);
            sym_finalize_t<double, 1>(st, amoeba, npol, pd, ps64, pa64, alpha, pp, x, xstride, y, ystride, dodiag,
                                      dotpart, skip, M);
        } else {
            sym_launch_t<double, 2>(st, amoeba, npol, pairs, ps64, pa64, x, xstride, pd, skip, M, evb, eve);
            sym_finalize_t<double, 2>(st, amoeba, npol, pd, ps64, pa64, alpha, pp, x, xstride, y, ystride, dodiag,
                                      dotpart, skip, M);
        }
    }
}

}  // namespace ommp
