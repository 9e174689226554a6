// Pair-once sweep of one block pair: the body shared by the single-system T
// matvec (tmv_sym.hip) and the batched solver (batch_kernels.hip).
//
// Sites are grouped in blocks of kSymBlock = 256 (four 64-site sub-tiles, one per
// wave).  For the block pair (I <= J), wave w holds the targets of sub-tile w of
// I in registers; the 256 sites of J are staged in LDS.  For sub-tile k of J,
// lane l meets source (l + s) mod 64 at step s = 0..63, so at every step the 64
// lanes touch 64 distinct sources: the target side accumulates in registers, and
// so does the source side, each lane holding the sum of its current source and
// taking over the next lane's after every step (a DPP wave rotate, on the vector
// ALU); after 64 steps every source-side sum is back at its source's lane.
// Every sum has a fixed order (results bitwise reproducible).  Measured
// alternatives for the source side, both slower: LDS float atomics into a
// per-wave row (FP32 ones were 14x slower here) and lane shuffles through the
// LDS crossbar.  In a diagonal block pair (I = J) the sub-tile pairs k < w are
// the mirror images of pairs handled by wave k and are skipped, and the
// diagonal sub-tile (k = w) is swept one-sided over s = 1..63.  Scale 1 for every
// pair (the finalize kernels apply the screening corrections), self excluded.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"
#include "pair_math.h"

namespace ommp {

constexpr int kSymWaves = kSymBlock / 64;

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

// LDS of sym_block_pair for up to M columns (bytes): the block's sources
// (position, damping, dipoles) and the per-wave rows of source-side sums.
template <typename R, int M>
constexpr int sym_lds_bytes() {
    return (int)(sizeof(typename V4<R>::t) * kSymBlock * (1 + M) + sizeof(R) * kSymWaves * M * 3 * 64);
}

// Pair geometry with the AMOEBA damping argument from cubed polarizability
// factors (vi = w_i^3, vj = w_j^3): a u^3 = a r^3 w_i^3 w_j^3 with one multiply
// fewer than cubing u = r w_i w_j.  Wang (AMBER) damping keeps u itself.
template <typename R, bool AMOEBA>
__device__ __forceinline__ void pair_geom_v(const typename V4<R>::t &pj, R aj, const typename V4<R>::t &pi, R ai,
                                            R &dx, R &dy, R &dz, R &k2, R &k3, R &arg, bool &near) {
    if (!AMOEBA) {
        pair_geom<R, AMOEBA, false>(pj, aj, pi, ai, dx, dy, dz, k2, k3, arg, near);
        return;
    }
    dx = pj.x - pi.x;
    dy = pj.y - pi.y;
    dz = pj.z - pi.z;
    const R r2 = fmaa(dz, dz, fmaa(dy, dy, dx * dx));
    const R ir = fast_rsqrt(r2);
    const R ir2 = ir * ir;
    k2 = ir * ir2;
    k3 = (R(3) * k2) * ir2;
    const R r3 = (r2 * ir) * r2;
    arg = pos_min(ai, aj) * (r3 * (pi.w * pj.w));
    near = arg < R(50);
}

// Both directions' accumulation of one pair: E_i += k3 (x_j.d) d - k2 x_j at the
// target, E_j += k3 (x_i.d) d - k2 x_i at the source, with q = k3 d formed once.
template <typename R, int M>
__device__ __forceinline__ void pair_accum2(R (&acc)[M][3], R *jacc, const typename V4<R>::t *xv,
                                            const R (&xi)[M][3], R dx, R dy, R dz, R k2, R k3) {
    const R qx = k3 * dx, qy = k3 * dy, qz = k3 * dz, nk2 = -k2;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const R c = fmaa(xv[m].z, qz, fmaa(xv[m].y, qy, xv[m].x * qx));
        acc[m][0] = fmaa(c, dx, fmaa(nk2, xv[m].x, acc[m][0]));
        acc[m][1] = fmaa(c, dy, fmaa(nk2, xv[m].y, acc[m][1]));
        acc[m][2] = fmaa(c, dz, fmaa(nk2, xv[m].z, acc[m][2]));
        const R cj = fmaa(xi[m][2], qz, fmaa(xi[m][1], qy, xi[m][0] * qx));
        jacc[3 * m + 0] = fmaa(cj, dx, fmaa(nk2, xi[m][0], jacc[3 * m + 0]));
        jacc[3 * m + 1] = fmaa(cj, dy, fmaa(nk2, xi[m][1], jacc[3 * m + 1]));
        jacc[3 * m + 2] = fmaa(cj, dz, fmaa(nk2, xi[m][2], jacc[3 * m + 2]));
    }
}

// Pair (I, J) of blocks with a kSymBlock-thread workgroup: xcol[m] = column m's
// dipoles [npol][3] (FP64); the target-side sums of component c of column m go
// to out_i[m][c * kSymBlock + t] (t = the site's index in I), the source-side
// ones to out_j[m][c * kSymBlock + t] (t in J).  lds: sym_lds_bytes<R, M>()
// bytes, 16-byte aligned (the caller's, so that bodies of different widths in
// one kernel share it); the body ends with a workgroup barrier.
template <typename R, int M, bool AMOEBA>
__device__ __forceinline__ void sym_block_pair(int npol, int I, int J, const typename V4<R>::t *__restrict__ site,
                                               const R *__restrict__ adamp, const double *const (&xcol)[M],
                                               R *const (&out_i)[M], R *const (&out_j)[M], char *lds) {
    using R4 = typename V4<R>::t;
    R4 *s_pos = reinterpret_cast<R4 *>(lds);
    R4(*s_x)[kSymBlock] = reinterpret_cast<R4(*)[kSymBlock]>(s_pos + kSymBlock);  // x, y, z; w of column 0: a
    R(*s_acc)[M * 3][64] = reinterpret_cast<R(*)[M * 3][64]>(s_x + M);
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
            if (AMOEBA) p.w = p.w * p.w * p.w;  // w^3 (pair_geom_v)
#pragma unroll
            for (int m = 0; m < M; m++) {
                v[m].x = (R)xcol[m][3L * jj + 0];
                v[m].y = (R)xcol[m][3L * jj + 1];
                v[m].z = (R)xcol[m][3L * jj + 2];
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
        if (AMOEBA) pj.w = pj.w * pj.w * pj.w;
        aj = adamp[i];
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) xi[m][c] = (R)xcol[m][3L * i + c];
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
                pair_geom_v<R, AMOEBA>(pj, aj, s_pos[js], xv[0].w, dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<R, AMOEBA>(arg, near, k2, k3);
                pair_accum<R, M>(acc, xv, dx, dy, dz, k2, k3);
            }
        } else if (!mirror) {
            // byte offset of the current source within the sub-tile (one add
            // and one mask per step)
            constexpr int kRow = 64 * (int)sizeof(R4);
            const char *pos_k = reinterpret_cast<const char *>(s_pos + k * 64);
            const char *x_k = reinterpret_cast<const char *>(&s_x[0][k * 64]);
            int jo = l * (int)sizeof(R4);
            for (int s = 0; s < 64; s++) {
                R4 xv[M];
#pragma unroll
                for (int m = 0; m < M; m++)
                    xv[m] = *reinterpret_cast<const R4 *>(x_k + m * kSymBlock * (int)sizeof(R4) + jo);
                const R4 ps = *reinterpret_cast<const R4 *>(pos_k + jo);
                R dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom_v<R, AMOEBA>(pj, aj, ps, xv[0].w, dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<R, AMOEBA>(arg, near, k2, k3);
                pair_accum2<R, M>(acc, jacc, xv, xi, dx, dy, dz, k2, k3);
#pragma unroll
                for (int c = 0; c < M * 3; c++) jacc[c] = lane_next(jacc[c]);
                jo = (jo + (int)sizeof(R4)) & (kRow - 1);
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
            out_j[c / 3][(c % 3) * kSymBlock + k * 64 + jl] = sum;
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < M; m++)
#pragma unroll
        for (int c = 0; c < 3; c++) out_i[m][c * kSymBlock + tid] = acc[m][c];
}

// Off-diagonal block pair (I < J) with two targets per lane: wave w holds the
// targets of sub-tiles 2 (w & 1) and 2 (w & 1) + 1 of I and sweeps sub-tiles
// 2 (w >> 1) and 2 (w >> 1) + 1 of J, so that every sub-tile pair is met by one
// wave and each source read from LDS, and each source-side sum rotated, serves
// two pairs.  A target's sum is split over the two waves holding it, a source's
// over the two waves sweeping it; both are added in wave order.  Same interface,
// LDS and outputs as sym_block_pair.
template <typename R, int M, bool AMOEBA>
__device__ __forceinline__ void sym_block_pair2(int npol, int I, int J, const typename V4<R>::t *__restrict__ site,
                                                const R *__restrict__ adamp, const double *const (&xcol)[M],
                                                R *const (&out_i)[M], R *const (&out_j)[M], char *lds) {
    using R4 = typename V4<R>::t;
    static_assert(kSymWaves == 4, "two target and two source halves");
    static_assert(M <= 2, "the target-side rows reuse the sources' LDS");
    R4 *s_pos = reinterpret_cast<R4 *>(lds);
    R4(*s_x)[kSymBlock] = reinterpret_cast<R4(*)[kSymBlock]>(s_pos + kSymBlock);
    R(*s_acc)[M * 3][64] = reinterpret_cast<R(*)[M * 3][64]>(s_x + M);
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int th = w & 1, sh = w >> 1;
    const R far = R(1e15), gap = R(1e9);
    {
        const int jj = J * kSymBlock + tid;
        R4 p, v[M];
        if (jj < npol) {
            p = site[jj];
            if (AMOEBA) p.w = p.w * p.w * p.w;  // w^3 (pair_geom_v)
#pragma unroll
            for (int m = 0; m < M; m++) {
                v[m].x = (R)xcol[m][3L * jj + 0];
                v[m].y = (R)xcol[m][3L * jj + 1];
                v[m].z = (R)xcol[m][3L * jj + 2];
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
    // targets: sites 128 th + l and 128 th + 64 + l of I
    R4 pt[2];
    R at[2], xt[2][M][3];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int li = 128 * th + 64 * h + l, i = I * kSymBlock + li;
        if (i < npol) {
            pt[h] = site[i];
            if (AMOEBA) pt[h].w = pt[h].w * pt[h].w * pt[h].w;
            at[h] = adamp[i];
#pragma unroll
            for (int m = 0; m < M; m++)
#pragma unroll
                for (int c = 0; c < 3; c++) xt[h][m][c] = (R)xcol[m][3L * i + c];
        } else {
            pt[h].x = pt[h].y = pt[h].z = far + gap * (R)li;
            pt[h].w = R(1);
            at[h] = R(0.39);
#pragma unroll
            for (int m = 0; m < M; m++) xt[h][m][0] = xt[h][m][1] = xt[h][m][2] = R(0);
        }
    }
    R acc[2][M][3];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int m = 0; m < M; m++) acc[h][m][0] = acc[h][m][1] = acc[h][m][2] = R(0);
    __syncthreads();
    constexpr int kRow = 64 * (int)sizeof(R4);
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const int k = 2 * sh + kk;
        R jacc[M * 3];
#pragma unroll
        for (int c = 0; c < M * 3; c++) jacc[c] = R(0);
        const char *pos_k = reinterpret_cast<const char *>(s_pos + k * 64);
        const char *x_k = reinterpret_cast<const char *>(&s_x[0][k * 64]);
        int jo = l * (int)sizeof(R4);
        for (int s = 0; s < 64; s++) {
            R4 xv[M];
#pragma unroll
            for (int m = 0; m < M; m++)
                xv[m] = *reinterpret_cast<const R4 *>(x_k + m * kSymBlock * (int)sizeof(R4) + jo);
            const R4 ps = *reinterpret_cast<const R4 *>(pos_k + jo);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                R dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom_v<R, AMOEBA>(pt[h], at[h], ps, xv[0].w, dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<R, AMOEBA>(arg, near, k2, k3);
                pair_accum2<R, M>(acc[h], jacc, xv, xt[h], dx, dy, dz, k2, k3);
            }
#pragma unroll
            for (int c = 0; c < M * 3; c++) jacc[c] = lane_next(jacc[c]);
            jo = (jo + (int)sizeof(R4)) & (kRow - 1);
        }
#pragma unroll
        for (int c = 0; c < M * 3; c++) s_acc[w][c][l] = jacc[c];
        __syncthreads();
        // source sides of sub-tiles kk and 2 + kk: rows of waves (0, 1) and (2, 3)
        for (int e = tid; e < 2 * M * 3 * 64; e += kSymBlock) {
            const int hh = e / (M * 3 * 64), r = e - hh * (M * 3 * 64), c = r >> 6, jl = r & 63;
            const R sum = s_acc[2 * hh][c][jl] + s_acc[2 * hh + 1][c][jl];
            out_j[c / 3][(c % 3) * kSymBlock + (2 * hh + kk) * 64 + jl] = sum;
        }
        __syncthreads();
    }
    // target sides: waves th and th + 2 hold the same targets (the sources' LDS
    // is free now): rows [wave][half][component][lane]
    R(*s_t)[2][M * 3][64] = reinterpret_cast<R(*)[2][M * 3][64]>(lds);
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int c = 0; c < 3; c++) s_t[w][h][3 * m + c][l] = acc[h][m][c];
    __syncthreads();
    for (int e = tid; e < 4 * M * 3 * 64; e += kSymBlock) {
        const int ts = e / (M * 3 * 64), r = e - ts * (M * 3 * 64), c = r >> 6, jl = r & 63;
        const int t2 = ts >> 1, h = ts & 1;
        const R sum = s_t[t2][h][c][jl] + s_t[t2 + 2][h][c][jl];
        out_i[c / 3][(c % 3) * kSymBlock + ts * 64 + jl] = sum;
    }
    __syncthreads();
}

// Two targets per lane: FP32 (one or two columns) and FP64 with one or two
// columns, 4-6% faster at 15,000 sites (FP64 two columns: 146 registers, three
// workgroups per CU instead of four, 0.306 -> 0.294 ms per sweep); wider FP64
// groups keep one target per lane.
template <typename R, int M>
constexpr bool kSymTwoTargets = sizeof(R) == 4 || M <= 2;

// The body for block pair (I, J): two targets per lane off the diagonal where
// the width allows it, one otherwise.
template <typename R, int M, bool AMOEBA>
__device__ __forceinline__ void sym_block_pair_any(int npol, int I, int J, const typename V4<R>::t *__restrict__ site,
                                                   const R *__restrict__ adamp, const double *const (&xcol)[M],
                                                   R *const (&out_i)[M], R *const (&out_j)[M], char *lds) {
    if constexpr (kSymTwoTargets<R, M>) {
        if (I != J) {
            sym_block_pair2<R, M, AMOEBA>(npol, I, J, site, adamp, xcol, out_i, out_j, lds);
            return;
        }
    }
    sym_block_pair<R, M, AMOEBA>(npol, I, J, site, adamp, xcol, out_i, out_j, lds);
}

}  // namespace ommp
