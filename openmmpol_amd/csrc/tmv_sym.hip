// T matvec on the fly with each unordered pair evaluated once (T_ij = T_ji):
// the geometry and damping of a pair (half of the per-pair work) are shared by
// its two directions.
//
// Sites are grouped in blocks of kSymBlock = 256 (four 64-site sub-tiles, one per
// wave).  One workgroup per block pair (I <= J), row-major over the upper
// triangle (pairs[] lists them).  Wave w holds the targets of sub-tile w of I in
// registers; the 256 sites of J are staged in LDS.  For sub-tile k of J, lane l
// meets source (l + s) mod 64 at step s = 0..63, so at every step the 64 lanes
// touch 64 distinct sources: the target side accumulates in registers, and so
// does the source side, each lane holding the sum of its current source and
// taking over the next lane's after every step (a DPP wave rotate, on the
// vector ALU); after 64 steps every source-side sum is back at its source's
// lane.  Every sum has a fixed order (results bitwise reproducible).  Measured
// alternatives for the source side, both slower: LDS float atomics into a
// per-wave row (FP32 ones were 14x slower here) and lane shuffles through the
// LDS crossbar.
// In a diagonal block pair (I = J) the sub-tile pairs k < w are the mirror
// images of pairs handled by wave k and are skipped, and the diagonal sub-tile
// (k = w) is swept one-sided over s = 1..63.
//
// Outputs, per block pair p: target side part_i[p][M * 3][256] (sites of I) and
// source side part_j[p][M * 3][256] (sites of J); the finalize kernel adds, for
// every site, the nb partials of its row and column in a fixed order (a separate
// reduction kernel with one thread per component measured slower).  Scale 1 for
// every pair (the finalize applies the screening corrections), self excluded.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "pair_math.h"

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

constexpr int kSymWaves = kSymBlock / 64;
static_assert(kSymBlock == kBlock, "the finalize kernel maps one block of sites per workgroup");


// Lane l + 1's value at lane l (wave rotate by one lane, DPP).
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_next(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x134, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kSymBlock) void k_tmv_sym(int npol, const int2 *__restrict__ pairs,
                                                       const typename V4<R>::t *__restrict__ site,
                                                       const R *__restrict__ adamp, const double *__restrict__ x,
                                                       long xstride, R *__restrict__ part_i, R *__restrict__ part_j,
                                                       const CgState *__restrict__ skip, int skipM) {
    using R4 = typename V4<R>::t;
    if (skip) {
        bool all = true;
        for (int m = 0; m < skipM; m++) all = all && skip[m].done;
        if (all) return;
    }
    __shared__ R4 s_pos[kSymBlock];
    __shared__ R4 s_x[M][kSymBlock];  // x, y, z of column m; w of column 0 holds the damping a
    __shared__ R s_acc[kSymWaves][M * 3][64];
    const int2 pr = pairs[blockIdx.x];
    const int I = pr.x, J = pr.y;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
    // Padding sites (past npol, in the last block only) sit far away at distinct
    // points with zero dipoles, so that no pair but the skipped self pair has a
    // zero distance.
    const R far = R(1e15), gap = R(1e9);
    // sources: the sites of J
    {
        const int jj = J * kSymBlock + tid;
        R4 p, v[M];
        if (jj < npol) {
            p = site[jj];
#pragma unroll
            for (int m = 0; m < M; m++) {
                v[m].x = (R)x[m * xstride + 3L * jj + 0];
                v[m].y = (R)x[m * xstride + 3L * jj + 1];
                v[m].z = (R)x[m * xstride + 3L * jj + 2];
                v[m].w = m == 0 ? adamp[jj] : R(0);
            }
        } else {
            p.x = p.y = p.z = far + gap * (R)tid;
            p.w = R(1);
#pragma unroll
            for (int m = 0; m < M; m++) {
                v[m].x = v[m].y = v[m].z = R(0);
                v[m].w = R(0.39);
            }
        }
        s_pos[tid] = p;
#pragma unroll
        for (int m = 0; m < M; m++) s_x[m][tid] = v[m];
    }
    // targets: this thread's site of I
    const int i = I * kSymBlock + tid;
    R4 pj;
    R aj, xi[M][3];
    if (i < npol) {
        pj = site[i];
        aj = adamp[i];
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) xi[m][c] = (R)x[m * xstride + 3L * i + c];
    } else {
        pj.x = pj.y = pj.z = far + gap * (R)tid;
        pj.w = R(1);
        aj = R(0.39);
#pragma unroll
        for (int m = 0; m < M; m++) xi[m][0] = xi[m][1] = xi[m][2] = R(0);
    }
    R acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = R(0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSymWaves; k++) {
        const bool mirror = I == J && k < w;  // wave-uniform
        const bool diag = I == J && k == w;
        R jacc[M * 3];
#pragma unroll
        for (int c = 0; c < M * 3; c++) jacc[c] = R(0);
        if (diag) {
            // one-sided over the other 63 sites of the sub-tile
            for (int s = 1; s < 64; s++) {
                const int js = k * 64 + ((l + s) & 63);
                R4 xv[M];
#pragma unroll
                for (int m = 0; m < M; m++) xv[m] = s_x[m][js];
                R dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom<R, AMOEBA, false>(pj, aj, s_pos[js], xv[0].w, dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<R, AMOEBA>(arg, near, k2, k3);
                pair_accum<R, M>(acc, xv, dx, dy, dz, k2, k3);
            }
        } else if (!mirror) {
            int jl = l;
            for (int s = 0; s < 64; s++) {
                const int js = k * 64 + jl;
                R4 xv[M];
#pragma unroll
                for (int m = 0; m < M; m++) xv[m] = s_x[m][js];
                R dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom<R, AMOEBA, false>(pj, aj, s_pos[js], xv[0].w, dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<R, AMOEBA>(arg, near, k2, k3);
                pair_accum<R, M>(acc, xv, dx, dy, dz, k2, k3);
                const R nk2 = -k2;
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const R c = k3 * fmaa(xi[m][2], dz, fmaa(xi[m][1], dy, xi[m][0] * dx));
                    jacc[3 * m + 0] = fmaa(c, dx, fmaa(nk2, xi[m][0], jacc[3 * m + 0]));
                    jacc[3 * m + 1] = fmaa(c, dy, fmaa(nk2, xi[m][1], jacc[3 * m + 1]));
                    jacc[3 * m + 2] = fmaa(c, dz, fmaa(nk2, xi[m][2], jacc[3 * m + 2]));
                }
#pragma unroll
                for (int c = 0; c < M * 3; c++) jacc[c] = lane_next(jacc[c]);
                jl = (jl + 1) & 63;
            }
        }
#pragma unroll
        for (int c = 0; c < M * 3; c++) s_acc[w][c][l] = jacc[c];
        __syncthreads();
        // source side of sub-tile k: the waves' rows added in wave order
        for (int e = tid; e < M * 3 * 64; e += kSymBlock) {
            const int c = e >> 6, jl = e & 63;
            R sum = s_acc[0][c][jl];
#pragma unroll
            for (int q = 1; q < kSymWaves; q++) sum += s_acc[q][c][jl];
            part_j[((long)blockIdx.x * M * 3 + c) * kSymBlock + k * 64 + jl] = sum;
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < M; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) part_i[((long)blockIdx.x * M * 3 + 3 * m + c) * kSymBlock + tid] = acc[m][c];
}

int tmv_sym_blocks(int npol) { return (npol + kSymBlock - 1) / kSymBlock; }

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
