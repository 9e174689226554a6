// T matvec on the fly with each unordered pair evaluated once (T_ij = T_ji):
// the geometry and damping of a pair (half of the per-pair work) are shared by
// its two directions.  One workgroup per pair of 256-site blocks (I <= J),
// row-major over the upper triangle (pairs[] lists them); the body is
// sym_block_pair (tmv_sym.h).
//
// Outputs, per block pair p: target side part_i[p][M * 3][256] (sites of I) and
// source side part_j[p][M * 3][256] (sites of J); the finalize kernel adds, for
// every site, the nb partials of its row and column in a fixed order (a separate
// reduction kernel with one thread per component measured slower).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "tmv_sym.h"

namespace ommp {

#define HIP_CHECK(x)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                      \
            std::abort();                                                                \
        }                                                                                \
    } while (0)

static_assert(kSymBlock == kBlock, "the finalize kernel maps one block of sites per workgroup");

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
    sym_block_pair<R, M, AMOEBA>(npol, pr.x, pr.y, site, adamp, xcol, oi, oj, lds);
}

int tmv_sym_blocks(int npol) { return (npol + kSymBlock - 1) / kSymBlock; }

// The pair-once sweep evaluates half the pairs, at the price of a workgroup per
// block pair and of its block-pair partial sums.  Measured on MI355X against the
// per-target sweep (two columns): FP64 faster from 5,000 sites (-7%) to 100,000
// (-25%); FP32 slower up to 10,000 sites (+14%), faster from 15,000 (-14%; -15%
// at 100,000).  Used from those sizes while its partial sums for ncols columns
// stay under max_bytes (2 GiB for two columns: up to ~100,000 sites).
bool tmv_pair_once(int npol, int prec, int ncols, size_t max_bytes) {
    if (npol < (prec == 32 ? 12000 : 4000)) return false;
    return tmv_sym_part_bytes(npol, ncols, 64) <= max_bytes;
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
    const int nb = tmv_sym_blocks(npol);
    if (prec == 32) {
        if (M == 1)
            sym_launch_t<float, 1>(st, amoeba, npol, pairs, ps32, pa32, x, xstride, (float *)part, skip, M, evb, eve);
        else
            sym_launch_t<float, 2>(st, amoeba, npol, pairs, ps32, pa32, x, xstride, (float *)part, skip, M, evb, eve);
    } else {
        if (M == 1)
            sym_launch_t<double, 1>(st, amoeba, npol, pairs, ps64, pa64, x, xstride, (double *)part, skip, M, evb,
                                    eve);
        else
            sym_launch_t<double, 2>(st, amoeba, npol, pairs, ps64, pa64, x, xstride, (double *)part, skip, M, evb,
                                    eve);
    }
    launch_tmv_finalize_prec(st, prec, amoeba, npol, M, part, -nb, ps64, pa64, alpha, pp, x, xstride, y, ystride,
                             dodiag, dotpart, skip_flag);
}

}  // namespace ommp
