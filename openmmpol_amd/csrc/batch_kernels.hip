// Kernels of the batched solver (batch.h): grouped pair kernels whose work
// items span several systems and column groups, and per-column PCG vector
// kernels driven by work lists.  Arithmetic per pair and per element is that of
// the single-system kernels (pair_math.h), so a column gives the same numbers
// in any batch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "batch.h"
#include "pair_math.h"
#include "tmv_sym.h"

namespace ommp {


template <typename R> struct SegSites;
template <> struct SegSites<float> {
    __device__ static const float4 *site(const BSeg &s) { return s.ps32; }
    __device__ static const float *damp(const BSeg &s) { return s.pa32; }
};
template <> struct SegSites<double> {
    __device__ static const double4 *site(const BSeg &s) { return s.ps64; }
    __device__ static const double *damp(const BSeg &s) { return s.pa64; }
};

// Does column c take part in this matvec?  FP32 (use_p): inner PCG columns in
// slot 1.  FP64: outer columns in slot 0 whose residual needs T x.
__device__ __forceinline__ bool mv_active(const BCol &c, int use_p) {
    return use_p ? c.s[1].phase == kPhaseInner : (c.s[0].phase == kPhaseOuter && c.rinit == kResMatvec);
}

// Grouped on-the-fly pair kernel body: one target per thread, sources of one
// split staged through LDS, MM columns cidx[] (as k_tmv_pairs).
template <typename R, int MM, bool AMOEBA>
__device__ __forceinline__ void bpairs_body(const BSeg &S, const int4 w, const double *__restrict__ xin,
                                            const int (&cidx)[MM]) {
    using R4 = typename V4<R>::t;
    const int npol = S.npol, ncols = S.ncols;
    const long n = 3L * npol;
    const R4 *__restrict__ site = SegSites<R>::site(S);
    const R *__restrict__ adamp = SegSites<R>::damp(S);
    __shared__ R4 s_pos[kBlock];
    __shared__ R4 s_x[MM][kBlock];
    const int tid = threadIdx.x;
    const int j = w.y * kBlock + tid;
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
    R acc[MM][3];
#pragma unroll
    for (int m = 0; m < MM; m++) acc[m][0] = acc[m][1] = acc[m][2] = R(0);
    const int i0 = w.z * S.chunk;
    const int i1 = min(npol, i0 + S.chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const int i = base + tid;
            s_pos[tid] = site[i];
#pragma unroll
            for (int m = 0; m < MM; m++) {
                R4 v;
                if (cidx[m] >= 0) {
                    const double *xc = xin + (long)cidx[m] * n + 3L * i;
                    v.x = (R)xc[0];
                    v.y = (R)xc[1];
                    v.z = (R)xc[2];
                } else {
                    v.x = v.y = v.z = R(0);
                }
                v.w = m == 0 ? adamp[i] : R(0);
                s_x[m][tid] = v;
            }
        } else if (tid == cnt) {
            R4 far;
            far.x = far.y = far.z = R(1e15);
            far.w = R(1);
            s_pos[tid] = far;
#pragma unroll
            for (int m = 0; m < MM; m++) {
                R4 z;
                z.x = z.y = z.z = R(0);
                z.w = R(0.39);
                s_x[m][tid] = z;
            }
        }
        __syncthreads();
        const int cnt2 = (cnt + 1) & ~1;
        for (int k = 0; k < cnt2; k += 2) {
            R4 xa[MM], xb[MM];
#pragma unroll
            for (int m = 0; m < MM; m++) {
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
            pair_accum<R, MM>(acc, xa, dxa, dya, dza, k2a, k3a);
            pair_accum<R, MM>(acc, xb, dxb, dyb, dzb, k2b, k3b);
        }
    }
    if (valid) {
        R *part = static_cast<R *>(S.part);
#pragma unroll
        for (int m = 0; m < MM; m++) {
            if (cidx[m] < 0) continue;
            R *o = part + ((long)w.z * ncols + cidx[m]) * n + 3L * j;
            o[0] = acc[m][0];
            o[1] = acc[m][1];
            o[2] = acc[m][2];
        }
    }
}

// Grouped pair kernel over work items {segment, target block, split, column
// group of M}.  A group of two with one active column runs the one-column body
// (the same per-pair arithmetic, so results do not depend on the other column).
template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_bpairs(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                   const int4 *__restrict__ work, int use_p,
                                                   const int *__restrict__ act, const int *__restrict__ act_n) {
    const int4 w = work[blockIdx.x];
    const BSeg &S = segs[w.x];
    const int c0 = w.w * M;
    const double *xin = use_p ? S.p : S.x;
    int cidx[M];
    int nact = 0;
    const int na = act ? act_n[w.x] : 0;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const bool a = act ? c0 + m < na
                           : c0 + m < S.ncols && mv_active(cols[S.col0 + c0 + m], use_p);
        cidx[m] = a ? (act ? act[S.col0 + c0 + m] : c0 + m) : -1;
        nact += a;
    }
    if (nact == 0) return;
    if constexpr (M == 2) {
        if (nact == 1) {
            const int one[1] = {cidx[0] >= 0 ? cidx[0] : cidx[1]};
            bpairs_body<R, 1, AMOEBA>(S, w, xin, one);
            return;
        }
    }
    bpairs_body<R, M, AMOEBA>(S, w, xin, cidx);
}

// Pair-once items {segment, I | J << 16, block pair index, column group}: the
// block pair's sweep (tmv_sym.h) for the active columns of the group, into the
// segment's block-pair partials [pair][ncols][3][256] (target sides, then the
// source sides).  A group of two with one active column runs the one-column
// body, as k_bpairs.
template <typename R, int M, bool AMOEBA>
__device__ __forceinline__ void bsym_body(const BSeg &S, const int4 w, const double *__restrict__ xin,
                                          const int (&cidx)[M], char *lds) {
    const int nb = S.tblocks;
    const long np = (long)nb * (nb + 1) / 2, n = 3L * S.npol;
    R *part_i = static_cast<R *>(S.part);
    R *part_j = part_i + np * S.ncols * 3 * kSymBlock;
    const double *xcol[M];
    R *oi[M], *oj[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        xcol[m] = xin + (long)cidx[m] * n;
        const long o = ((long)w.z * S.ncols + cidx[m]) * 3 * kSymBlock;
        oi[m] = part_i + o;
        oj[m] = part_j + o;
    }
    sym_block_pair_any<R, M, AMOEBA>(S.npol, w.y & 0xffff, w.y >> 16, SegSites<R>::site(S), SegSites<R>::damp(S),
                                     xcol, oi, oj, lds);
}

template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kSymBlock) void k_bsym(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                    const int4 *__restrict__ work, int use_p,
                                                    const int *__restrict__ act, const int *__restrict__ act_n) {
    // columns per sweep: FP64 sweeps take up to four (the pair's tensor, the
    // costly part in FP64, is then built once for all of them), FP32 two
    constexpr int kPair = sizeof(R) == 8 ? (M < 4 ? M : 4) : (M < 2 ? M : 2);
    __shared__ __align__(16) char lds[sym_lds_bytes<R, kPair>()];
    const int4 w = work[blockIdx.x];
    const BSeg &S = segs[w.x];
    const int c0 = w.w * M;
    const double *xin = use_p ? S.p : S.x;
    int idx[M];
    int k = 0;
    const int na = act ? act_n[w.x] : 0;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const bool a = act ? c0 + m < na
                           : c0 + m < S.ncols && mv_active(cols[S.col0 + c0 + m], use_p);
        idx[m] = -1;
        if (a) idx[k++] = act ? act[S.col0 + c0 + m] : c0 + m;
    }
    if (k == 0) return;  // uniform across the workgroup
    if constexpr (kPair >= 4) {
        if (k == 4) {
            const int four[4] = {idx[0], idx[1], idx[2], idx[3]};
            bsym_body<R, 4, AMOEBA>(S, w, xin, four, lds);
            return;
        }
        if (k == 3) {
            const int three[3] = {idx[0], idx[1], idx[2]};
            bsym_body<R, 3, AMOEBA>(S, w, xin, three, lds);
            return;
        }
    }
    if (k == 1) {
        const int one[1] = {idx[0]};
        bsym_body<R, 1, AMOEBA>(S, w, xin, one, lds);
        return;
    }
    if constexpr (kPair >= 2 && (M == 2 || kPair >= 4)) {
        // two active columns (the wider counts were handled above)
        const int two[2] = {idx[0], idx[1]};
        bsym_body<R, 2, AMOEBA>(S, w, xin, two, lds);
    } else {
        // the active columns of the group in turn, two at a time (an odd one alone)
        int q = 0;
        for (; q + 1 < k; q += 2) {
            const int two[2] = {idx[q], idx[q + 1]};
            bsym_body<R, kPair < 2 ? 1 : 2, AMOEBA>(S, w, xin, two, lds);
        }
        if (q < k) {
            const int one[1] = {idx[q]};
            bsym_body<R, 1, AMOEBA>(S, w, xin, one, lds);
        }
    }
}

// ------------------------------------------------- bf16x6 matrix-core matvec
// Multi-RHS form of the on-the-fly FP32 matvec as a dense contraction on the
// bf16 matrix cores, to FP32 accuracy.  Every FP32 operand v is split in three
// bf16 parts, v = v0 + v1 + v2 (round to nearest at each step: 8 significant
// bits per part, 24 in all, as in FP32), and A X is summed over the six products
// A_i X_j with i + j <= 2, accumulated in FP32; the dropped ones (i + j >= 3)
// are below 2^-24 relative, the size of FP32's own rounding.
// Tiles: 32 target sites (96 rows, rows 3 t + a) x 32 source sites (K = 96,
// k = 3 s + b, three steps of v_mfma_f32_16x16x32_bf16) x 128 columns.  The
// damped dipole tensors k3 d d^T - k2 I of a tile are built in FP32 and stored
// split in LDS; the X fragments of a lane (its column, 8 consecutive k) come
// straight from the split copy of p in global memory, issued before the tile's
// tensors are built so that the loads land meanwhile.  Wave w owns columns
// 32 w .. 32 w + 31 (two 16-column blocks) and all six 16-row blocks; lane
// group g = lane / 16 supplies k = 8 g .. 8 g + 7 of each step and holds rows
// 4 g .. 4 g + 3 of each result block.  Partial fields go to
// part[split][column][3 npol] as in k_bpairs; the finalize (k_bfinal) is shared.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bx8 __attribute__((ext_vector_type(8)));
typedef const bx8 __attribute__((address_space(1))) *gbx8;  // global-memory loads, not flat
constexpr int kMfmaRows = 3 * kMfmaSites;  // 96
constexpr int kBxK = 3 * kBxSrc;           // 96
constexpr int kBxLd = kBxK + 8;            // LDS row stride (bf16): conflict-free 16-byte reads

// loads through the global address space (flat loads would also hold up the
// LDS counter that the tile's LDS accesses wait on)
__device__ __forceinline__ float4 gload(const float4 *p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const v4 __attribute__((address_space(1))) *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float gload(const float *p) { return *(const float __attribute__((address_space(1))) *)p; }

// Two FP32 values rounded to bf16 (nearest) and packed, a in the low half: one
// v_cvt_pk_bf16_f32.
__device__ __forceinline__ unsigned bx_cvt2(float a, float b) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f2){a, b}, b2));
}
// v = p0 + p1 + p2 for two values at once: three bf16 parts, each rounded to
// nearest from what the previous ones left (every subtraction is exact in FP32),
// packed in pairs as the LDS tiles hold them.
__device__ __forceinline__ void bx_split3(float v0, float v1, unsigned (&p)[kBxParts]) {
#pragma unroll
    for (int q = 0; q < kBxParts; q++) {
        p[q] = bx_cvt2(v0, v1);
        v0 -= __uint_as_float(p[q] << 16);
        v1 -= __uint_as_float(p[q] & 0xffff0000u);
    }
}

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_bpairs_bx6(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                       const int4 *__restrict__ work) {
    const int4 w = work[blockIdx.x];
    const BSeg &S = segs[w.x];
    const int npol = S.npol, ncols = S.ncols, ldx = S.ldx;
    const long n = 3L * npol;
    const int c0 = w.w * kMfmaCols;
    const float4 *__restrict__ ps32 = S.ps32;
    const float *__restrict__ pa32 = S.pa32;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, l16 = lane & 15;
    __shared__ __align__(16) __bf16 As[kBxParts][kMfmaRows][kBxLd];
    __shared__ float4 s_tp[kMfmaSites], s_sp[kBxSrc];
    __shared__ float s_ta[kMfmaSites], s_sa[kBxSrc];
    __shared__ int s_act[kMfmaCols];
    int mine = 0;
    if (tid < kMfmaCols) {
        mine = c0 + tid < ncols && cols[S.col0 + c0 + tid].s[1].phase == kPhaseInner;
        s_act[tid] = mine;
    }
    if (!__syncthreads_or(mine)) return;
    const int t0 = w.y * kMfmaSites;
    if (tid < kMfmaSites) {
        const int t = t0 + tid;
        float4 p;
        float a;
        if (t < npol) {
            p = gload(ps32 + t);
            a = gload(pa32 + t);
        } else {
            p = make_float4(0.f, 0.f, 0.f, 1.f);
            a = 0.39f;
        }
        s_tp[tid] = p;
        s_ta[tid] = a;
    }
    const int i0 = w.z * S.chunk;
    const int i1 = min(npol, i0 + S.chunk);
    // this lane's two columns and their split rows; an inactive column reads
    // row 0 (in bounds, any values): it only feeds its own result column, which
    // is not stored
    const long plane = (long)ncols * ldx;
    bool cact[2];
    gbx8 xrow[2];
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const int c = 32 * wv + 16 * cb + l16;
        cact[cb] = s_act[c];
        const long o = cact[cb] ? (long)(c0 + c) * ldx + 8 * g : 0;
        xrow[cb] = (gbx8)(reinterpret_cast<const __bf16 *>(S.xs) + o);
    }
    // source positions of the next tile go through registers
    float4 spr = make_float4(1e15f, 1e15f, 1e15f, 1.f);
    float sar = 0.39f;
    auto fetch_src = [&](int base) {
        if (tid < kBxSrc) {
            if (base + tid < i1) {
                spr = gload(ps32 + base + tid);
                sar = gload(pa32 + base + tid);
            } else {
                spr = make_float4(1e15f, 1e15f, 1e15f, 1.f);
                sar = 0.39f;
            }
        }
    };
    mfma_f4 acc[2][6];
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int rb = 0; rb < 6; rb++) acc[cb][rb] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    if (i0 < i1) {
        fetch_src(i0);
        if (tid < kBxSrc) {
            s_sp[tid] = spr;
            s_sa[tid] = sar;
        }
    }
    __syncthreads();
    for (int base = i0; base < i1; base += kBxSrc) {
        // X fragments of this tile: k = 3 base + 32 q + 8 g .. + 7 of the lane's columns
        bx8 bx[kBxParts][2][3];
        const long e = (3L * base) / 8;
#pragma unroll
        for (int pt = 0; pt < kBxParts; pt++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int q = 0; q < 3; q++) bx[pt][cb][q] = xrow[cb][pt * plane / 8 + e + 4 * q];
        // A tile: thread (wave w, lane l) builds rows 3 t + a of targets
        // t = l % 16 and l % 16 + 16 against sources 2 sp, 2 sp + 1 with
        // sp = 4 w + l / 16 (3 rows x 6 k per target), so that it reads its two
        // sources once for both targets; a wave covers 16 targets x 4 source
        // pairs (packed stores at dword 156 t + 3 sp + const: distinct banks).
        const int sp = 4 * (tid >> 6) + ((tid >> 4) & 3);
#pragma unroll
        for (int kt = 0; kt < 2; kt++) {
            const int ti = (tid & 15) + 16 * kt;
            float v[3][6];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int sj = 2 * sp + h;
                float dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom<float, AMOEBA>(s_tp[ti], s_ta[ti], s_sp[sj], s_sa[sj], dx, dy, dz, k2, k3, arg, near);
                // every lane runs both targets: the ballot is wave-uniform
                if (__ballot(near)) pair_damp<float, AMOEBA>(arg, near, k2, k3);
                const float d[3] = {dx, dy, dz};
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) v[a][3 * h + b] = k3 * d[a] * d[b] - (a == b ? k2 : 0.f);
            }
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    unsigned pk[kBxParts];
                    bx_split3(v[a][2 * u], v[a][2 * u + 1], pk);
#pragma unroll
                    for (int pt = 0; pt < kBxParts; pt++)
                        reinterpret_cast<unsigned *>(&As[pt][3 * ti + a][6 * sp])[u] = pk[pt];
                }
        }
        __syncthreads();
        // the next tile's sources, while this tile's products run
        if (base + kBxSrc < i1) {
            fetch_src(base + kBxSrc);
            if (tid < kBxSrc) {
                s_sp[tid] = spr;
                s_sa[tid] = sar;
            }
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
#pragma unroll
            for (int rb = 0; rb < 6; rb++) {
                bx8 ax[kBxParts];
#pragma unroll
                for (int pt = 0; pt < kBxParts; pt++)
                    ax[pt] = *reinterpret_cast<const bx8 *>(&As[pt][16 * rb + l16][32 * q + 8 * g]);
                // the six products A_i X_j, i + j <= 2, smallest first
#pragma unroll
                for (int cb = 0; cb < 2; cb++) {
                    mfma_f4 c = acc[cb][rb];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[2], bx[0][cb][q], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[1], bx[1][cb][q], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[0], bx[2][cb][q], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[1], bx[0][cb][q], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[0], bx[1][cb][q], c, 0, 0, 0);
                    acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[0], bx[0][cb][q], c, 0, 0, 0);
                }
            }
        }
        __syncthreads();  // this tile's A reads are done; the next sources are in LDS
    }
    // lane holds D[16 rb + 4 g + i][l16] of column blocks 2 wv, 2 wv + 1
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const int c = 32 * wv + 16 * cb + l16;
        if (!cact[cb]) continue;
        float *part = static_cast<float *>(S.part) + ((long)w.z * ncols + c0 + c) * n + 3L * t0;
#pragma unroll
        for (int rb = 0; rb < 6; rb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * rb + 4 * g + i;
                if (t0 + row / 3 < npol) part[row] = acc[cb][rb][i];
            }
    }
}

// ------------------------------------------------- FP64 matrix-core matvec
// The FP64 refinement matvec of a wide batch as a dense contraction on the FP64
// matrix cores (v_mfma_f64_16x16x4f64, whose rate on MI355X equals the FP64
// vector rate: what it saves is the per-column instruction stream of the
// on-the-fly sweeps).  Work item: 32 target sites (96 rows, rows 3 t + a) x
// kD64Cols packed refining columns (the segment's active-column list) x one
// source split, in tiles of 16 sources (K = 48, k = 3 s + b).  The damped
// tensors k3 d d^T - k2 I of a tile are built in FP64 into LDS.  Lane
// (n = lane % 16, g = lane / 16) of wave w supplies, at step q of the twelve,
// A[m][12 g + q] (m = n within the row block) and X[12 g + q][column 16 w + n]:
// the four lane groups take the tile's K in four contiguous runs, so that a
// lane's A and X fragments are 12 consecutive doubles each; the lane holds
// D[g + 4 i][n], i = 0..3, of its row blocks.  With three or four 16-column
// blocks in the item, wave w takes column block w and all six row blocks; with
// one or two, the waves share the row blocks of a column block (w of a pair
// takes three; with one block the waves take 2, 2, 1 and 1), so that an item
// with few refining columns still spreads its products over the four SIMDs.
// Partial fields go to part[split][column][3 npol] (FP64), summed by k_bfinal.
constexpr int kD64Src = 16;
constexpr int kD64K = 3 * kD64Src;  // 48
constexpr int kD64Ld = kD64K + 2;   // LDS row stride (doubles)
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_bpairs_d64(const BSeg *__restrict__ segs, const int4 *__restrict__ work,
                                                       const int *__restrict__ act, const int *__restrict__ act_n) {
    const int4 w = work[blockIdx.x];
    const BSeg &S = segs[w.x];
    const int na = act_n[w.x];
    const int c0 = w.w * kD64Cols;
    if (c0 >= na) return;  // uniform across the workgroup
    const int npol = S.npol, ncols = S.ncols;
    const long n = 3L * npol;
    const double4 *__restrict__ ps = S.ps64;
    const double *__restrict__ pa = S.pa64;
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, l16 = lane & 15;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
    __shared__ __align__(16) double As[3 * kMfmaSites][kD64Ld];
    __shared__ double4 s_tp[kMfmaSites], s_sp[kD64Src];
    __shared__ double s_ta[kMfmaSites], s_sa[kD64Src];
    const int t0 = w.y * kMfmaSites;
    if (tid < kMfmaSites) {
        const int t = t0 + tid;
        // targets past npol: far apart, their rows are not stored
        s_tp[tid] = t < npol ? ps[t] : make_double4(1e15 + 1e9 * tid, 1e15, 1e15, 1.0);
        s_ta[tid] = t < npol ? pa[t] : 0.39;
    }
    // this wave's column block and row blocks [rb0, rb1) (wave-uniform)
    const int cblocks = (min(kD64Cols, na - c0) + 15) / 16;
    int cb = wv, rb0 = 0, rb1 = 6;
    if (cblocks == 2) {
        cb = wv >> 1;
        rb0 = 3 * (wv & 1);
        rb1 = rb0 + 3;
    } else if (cblocks == 1) {
        cb = 0;
        rb0 = wv < 2 ? 2 * wv : wv + 2;
        rb1 = wv < 2 ? rb0 + 2 : rb0 + 1;
    }
    // this lane's column: position cl of the packed list; an inactive lane feeds
    // zeros to its own result column, which is not stored
    const int cl = c0 + 16 * cb + l16;
    const bool cact = cl < na;
    const int col = cact ? act[S.col0 + cl] : 0;
    const double *__restrict__ xc = S.x + (long)col * n;
    const bool wact = cb < cblocks;
    const int i0 = w.z * S.chunk, i1 = min(npol, i0 + S.chunk);
    auto fetch_src = [&](int base) {
        if (tid < kD64Src) {
            const int i = base + tid;
            s_sp[tid] = i < i1 ? ps[i] : make_double4(-1e15 - 1e9 * tid, -1e15, -1e15, 1.0);
            s_sa[tid] = i < i1 ? pa[i] : 0.39;
        }
    };
    // X fragment of the tile at base: k = 12 g .. 12 g + 11, i.e. sources
    // base + 4 g .. base + 4 g + 3, zero past the split's end
    auto fetch_x = [&](int base, double (&xf)[12]) {
        const int s0 = base + 4 * g;
        if (cact && s0 + 3 < i1) {
            // a column starts 8-byte aligned only (3 npol doubles per column)
#pragma unroll
            for (int q = 0; q < 12; q++) xf[q] = xc[3L * s0 + q];
        } else {
#pragma unroll
            for (int q = 0; q < 12; q++) xf[q] = cact && s0 + q / 3 < i1 ? xc[3L * s0 + q] : 0.0;
        }
    };
    mfma_d4 acc[6];
#pragma unroll
    for (int rb = 0; rb < 6; rb++) acc[rb] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    double xf[12];
    if (i0 < i1) {
        fetch_src(i0);
        fetch_x(i0, xf);
    }
    __syncthreads();
    for (int base = i0; base < i1; base += kD64Src) {
        // tensors of the tile: thread (t = tid % 32, s = tid / 32) builds the
        // 3 x 3 blocks of target t against sources s and s + 8
        {
            const int ti = tid & 31;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int sj = (tid >> 5) + 8 * h;
                double dx, dy, dz, k2, k3, arg;
                bool near;
                pair_geom<double, AMOEBA>(s_tp[ti], s_ta[ti], s_sp[sj], s_sa[sj], dx, dy, dz, k2, k3, arg, near);
                if (__ballot(near)) pair_damp<double, AMOEBA>(arg, near, k2, k3);
                const double d[3] = {dx, dy, dz};
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) As[3 * ti + a][3 * sj + b] = k3 * d[a] * d[b] - (a == b ? k2 : 0.0);
            }
        }
        double xn[12];
        __syncthreads();
        // the next tile's sources and X fragments, while this tile's products run
        const bool more = base + kD64Src < i1;
        if (more) {
            fetch_src(base + kD64Src);
            fetch_x(base + kD64Src, xn);
        }
        if (wact) {
#pragma unroll
            for (int rb = 0; rb < 6; rb++) {
                if (rb < rb0 || rb >= rb1) continue;  // wave-uniform
                double af[12];
                const double2 *arow = reinterpret_cast<const double2 *>(&As[16 * rb + l16][12 * g]);
#pragma unroll
                for (int u = 0; u < 6; u++) {
                    const double2 v = arow[u];
                    af[2 * u] = v.x;
                    af[2 * u + 1] = v.y;
                }
                mfma_d4 c = acc[rb];
#pragma unroll
                for (int q = 0; q < 12; q++) c = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q], xf[q], c, 0, 0, 0);
                acc[rb] = c;
            }
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < 12; q++) xf[q] = xn[q];
        }
        __syncthreads();  // this tile's A reads are done; the next sources are in LDS
    }
    if (!cact) return;
    double *part = static_cast<double *>(S.part) + ((long)w.z * ncols + col) * n + 3L * t0;
#pragma unroll
    for (int rb = 0; rb < 6; rb++) {
        if (rb < rb0 || rb >= rb1) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 16 * rb + g + 4 * i;
            if (t0 + row / 3 < npol) part[row] = acc[rb][i];
        }
    }
}

// Three-part bf16 split of the inner search directions for the matrix-core
// matvec; the first block of a column also zeroes the row's padding.
__global__ __launch_bounds__(kBlock) void k_bconvert(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                     const int4 *__restrict__ work) {
    const int4 w = work[blockIdx.x];
    if (cols[w.w].s[1].phase != kPhaseInner) return;
    const BSeg &S = segs[w.x];
    const long n = 3L * S.npol;
    const long k = (long)w.z * kBlock + threadIdx.x;
    const long plane = (long)S.ncols * S.ldx;
    __bf16 *xs = reinterpret_cast<__bf16 *>(S.xs) + (long)w.y * S.ldx;
    if (k < n) {
        unsigned pk[kBxParts];
        bx_split3((float)S.p[(long)w.y * n + k], 0.f, pk);
#pragma unroll
        for (int q = 0; q < kBxParts; q++) reinterpret_cast<unsigned short *>(xs)[q * plane + k] = (unsigned short)pk[q];
    }
    if (w.z == 0 && n + threadIdx.x < S.ldx)
#pragma unroll
        for (int q = 0; q < kBxParts; q++) xs[q * plane + n + threadIdx.x] = (__bf16)0.f;
}

// h = -(sum of split partials) + (1 - s) corrections + x / alpha for the
// active columns of one target block; FP32 passes also store h.p partials.
template <typename R, int M, bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_bfinal(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                   const int4 *__restrict__ work, int use_p) {
    const int4 w = work[blockIdx.x];
    const BSeg &S = segs[w.x];
    const int npol = S.npol, ncols = S.ncols;
    const long n = 3L * npol;
    const int c0 = w.z * M;
    bool act[M];
    bool any = false;
#pragma unroll
    for (int m = 0; m < M; m++) {
        act[m] = c0 + m < ncols && mv_active(cols[S.col0 + c0 + m], use_p);
        any = any || act[m];
    }
    if (!any) return;
    __shared__ double sh[4];
    const double *xin = use_p ? S.p : S.x;
    const R *part = static_cast<const R *>(S.part);
    const int j = w.y * kBlock + threadIdx.x;
    const bool valid = j < npol;
    double acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    if (valid && !(use_p ? S.sym32 : S.sym64)) {
        for (int s = 0; s < S.nsplit; s++)
#pragma unroll
            for (int m = 0; m < M; m++) {
                if (!act[m]) continue;
                const R *p = part + ((long)s * ncols + c0 + m) * n + 3L * j;
                acc[m][0] -= (double)p[0];
                acc[m][1] -= (double)p[1];
                acc[m][2] -= (double)p[2];
            }
    } else if (valid) {
        // block-pair partials: the row (b, J >= b) of target sides, then the
        // column (I <= b, b) of source sides; pair (I, J) has index
        // I nb - I (I - 1) / 2 + J - I
        const int nb = S.tblocks, b = w.y, li = threadIdx.x;
        const long np = (long)nb * (nb + 1) / 2, row = (long)b * nb - (long)b * (b - 1) / 2;
        const R *part_j = part + np * ncols * 3 * kSymBlock;
#pragma unroll
        for (int m = 0; m < M; m++) {
            if (!act[m]) continue;
            R sum[3] = {R(0), R(0), R(0)};
#pragma unroll 4
            for (int J = b; J < nb; J++) {
                const R *p = part + (((row + J - b) * ncols + c0 + m) * 3) * kSymBlock + li;
                sum[0] += p[0];
                sum[1] += p[kSymBlock];
                sum[2] += p[2 * kSymBlock];
            }
#pragma unroll 4
            for (int I = 0; I <= b; I++) {
                const long pidx = (long)I * nb - (long)I * (I - 1) / 2 + b - I;
                const R *p = part_j + ((pidx * ncols + c0 + m) * 3) * kSymBlock + li;
                sum[0] += p[0];
                sum[1] += p[kSymBlock];
                sum[2] += p[2 * kSymBlock];
            }
            acc[m][0] = -(double)sum[0];
            acc[m][1] = -(double)sum[1];
            acc[m][2] = -(double)sum[2];
        }
    }
    if (valid) {
        const double4 pj = S.ps64[j];
        const double aj = S.pa64[j];
        for (int k = S.pp_ri[j]; k < S.pp_ri[j + 1]; k++) {
            const int i = S.pp_ci[k];
            const double f = S.pp_f[k];
            const double4 pi = S.ps64[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3;
            dip_kernel<double, AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, S.pa64[i], aj, k2, k3);
#pragma unroll
            for (int m = 0; m < M; m++) {
                if (!act[m]) continue;
                const double *xi = xin + (long)(c0 + m) * n + 3L * i;
                const double c = k3 * (xi[0] * dx + xi[1] * dy + xi[2] * dz);
                acc[m][0] += f * (c * dx - k2 * xi[0]);
                acc[m][1] += f * (c * dy - k2 * xi[1]);
                acc[m][2] += f * (c * dz - k2 * xi[2]);
            }
        }
        const double ia = 1.0 / S.alpha[j];
#pragma unroll
        for (int m = 0; m < M; m++) {
            if (!act[m]) continue;
            const double *xj = xin + (long)(c0 + m) * n + 3L * j;
            double *yj = S.h + (long)(c0 + m) * n + 3L * j;
            for (int c = 0; c < 3; c++) {
                acc[m][c] += xj[c] * ia;
                yj[c] = acc[m][c];
            }
        }
    }
    if (use_p) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            if (!act[m]) continue;  // uniform over the block
            double v = 0.0;
            if (valid) {
                const double *xj = xin + (long)(c0 + m) * n + 3L * j;
                for (int c = 0; c < 3; c++) v += acc[m][c] * xj[c];
            }
            v = block_sum_all(v, sh);
            if (threadIdx.x == 0) S.red_hp[(long)(c0 + m) * S.tblocks + w.y] = v;
        }
    }
}

// ------------------------------------------------------------ vector kernels
// Work item {segment, column, vector block, batch column}; element k of the
// column belongs to site k / 3.  The residual and update kernels that apply
// the preconditioner take one site per thread instead (work item z < tblocks:
// sites [256 z, 256 z + 256); the other items of the column write zero
// partials), so that each neighbour's three components are read once.
//
// With the near-field preconditioner (segment lists present, the column not
// switched off): z = M^-1 r, M^-1 = 2 alpha - alpha T_near alpha (precond.hip),
// summing B_ij r_j over the neighbours j of the site, the neighbours' r formed
// by the same expression as their own sites form it.  Convergence is measured
// on r.(alpha r) as the reference does (mod_solvers.F90:188-199), beta on r.z.
// Without it z = alpha r and the two partial sums coincide.
constexpr double kBatchNbDiag = 2.0;

// acc = sum_j B_ij r_j over the neighbours of site i; r_j = rj(oj + c) with oj
// the offset of site j's x component in this column
// (neighbours taken kNbChunk at a time, their loads issued together; the sum
// keeps the list order)
constexpr int kNbChunk = 4;
template <typename RF>
__device__ __forceinline__ void nb_site(const BSeg &S, int i, long colbase, RF rj, double (&acc)[3]) {
    acc[0] = acc[1] = acc[2] = 0.0;
    const int q0 = S.nb_row_ptr[i], q1 = S.nb_row_ptr[i + 1];
    for (int qb = q0; qb < q1; qb += kNbChunk) {
        int jn[kNbChunk];
#pragma unroll
        for (int u = 0; u < kNbChunk; u++) jn[u] = qb + u < q1 ? S.nb_ci[qb + u] : -1;
        double r[kNbChunk][3];
        float b[kNbChunk][6];
#pragma unroll
        for (int u = 0; u < kNbChunk; u++) {
            if (jn[u] >= 0) {
                const long oj = colbase + 3L * jn[u];
#pragma unroll
                for (int c = 0; c < 3; c++) r[u][c] = rj(oj + c);
#pragma unroll
                for (int e = 0; e < 6; e++) b[u][e] = S.nb_b[6L * (qb + u) + e];
            }
        }
#pragma unroll
        for (int u = 0; u < kNbChunk; u++) {
            if (jn[u] < 0) break;
            const double bxx = b[u][0], bxy = b[u][1], bxz = b[u][2], byy = b[u][3], byz = b[u][4], bzz = b[u][5];
            acc[0] += bxx * r[u][0] + bxy * r[u][1] + bxz * r[u][2];
            acc[1] += bxy * r[u][0] + byy * r[u][1] + byz * r[u][2];
            acc[2] += bxz * r[u][0] + byz * r[u][1] + bzz * r[u][2];
        }
    }
}

// r.z and r.(alpha r) partials of a site-block item (zero for the items past
// the last site block)
__device__ __forceinline__ void site_partials(const BSeg &S, const int4 w, double vz, double va, double *sh) {
    vz = block_sum_all(vz, sh);
    if (threadIdx.x == 0) S.red_rz[(long)w.y * S.vblocks + w.z] = vz;
    va = block_sum_all(va, sh);
    if (threadIdx.x == 0) S.red_ra[(long)w.y * S.vblocks + w.z] = va;
}

// Outer step, part 1: r = b - T x (x = 0: r = b; known: r kept), z = M^-1 r,
// partials of r.z and r.(alpha r).
__global__ __launch_bounds__(kBlock) void k_bresid(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                   const int4 *__restrict__ work) {
    __shared__ double sh[4];
    const int4 w = work[blockIdx.x];
    const BCol &C = cols[w.w];
    if (C.s[0].phase != kPhaseOuter) return;
    const BSeg &S = segs[w.x];
    if (w.z >= S.tblocks) {
        if (threadIdx.x == 0) {
            S.red_rz[(long)w.y * S.vblocks + w.z] = 0.0;
            S.red_ra[(long)w.y * S.vblocks + w.z] = 0.0;
        }
        return;
    }
    const long n = 3L * S.npol;
    const int i = w.z * kBlock + threadIdx.x;
    const long colbase = (long)w.y * n;
    const int rinit = C.rinit;
    const bool nb = S.nb_row_ptr != nullptr && !C.nb_off;
    auto rule = [&](long q) {
        return rinit == kResXZero ? S.b[q] : (rinit == kResKnown ? S.r[q] : S.b[q] - S.h[q]);
    };
    double vz = 0.0, va = 0.0;
    if (i < S.npol) {
        const long o = colbase + 3L * i;
        const double al = S.alpha[i];
        double rr[3], zz[3], acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 3; c++) rr[c] = rule(o + c);
        if (nb) nb_site(S, i, colbase, rule, acc);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            zz[c] = nb ? kBatchNbDiag * (al * rr[c]) + acc[c] : al * rr[c];
            // a known residual is already in r (and its neighbours read it)
            if (rinit != kResKnown) S.r[o + c] = rr[c];
            S.z[o + c] = zz[c];
            vz += rr[c] * zz[c];
            va += rr[c] * (al * rr[c]);
        }
    }
    site_partials(S, w, vz, va, sh);
}

// Outer step, part 2 (state copy 0 -> 1): rms test; if not converged, start an
// inner PCG on T d = r from d = 0 (p = z, g = r.z).
__global__ __launch_bounds__(kBlock) void k_bdecide(const BSeg *__restrict__ segs, BCol *__restrict__ cols,
                                                    const int4 *__restrict__ work, double eta, double frac) {
    __shared__ double sh[4];
    const int4 w = work[blockIdx.x];
    const BCol &C = cols[w.w];
    const BColSub st = C.s[0];
    BColSub nst = st;
    if (st.phase == kPhaseOuter) {
        const BSeg &S = segs[w.x];
        const long n = 3L * S.npol;
        const double ra = reduce_parts(S.red_ra + (long)w.y * S.vblocks, S.vblocks, sh);
        const double rz = reduce_parts(S.red_rz + (long)w.y * S.vblocks, S.vblocks, sh);
        const double rms = sqrt(ra / (double)n);
        nst.rms = rms;
        if (rms < C.tol) {
            nst.phase = kPhaseDone;
            nst.status = 0;
        } else if (st.outer >= kBatchMaxOuter) {
            nst.phase = kPhaseDone;
            nst.status = 2;
        } else {
            nst.phase = kPhaseInner;
            nst.tin = fmax(eta * rms, frac * C.tol);
            nst.g = rz;
            nst.it = 0;
            nst.outer = st.outer + 1;
            const long k = (long)w.z * kBlock + threadIdx.x;
            if (k < n) {
                const long o = (long)w.y * n + k;
                S.d[o] = 0.0;
                S.p[o] = S.z[o];
            }
        }
    }
    if (w.z == 0 && threadIdx.x == 0) cols[w.w].s[1] = nst;
}

// Inner PCG, update: alpha = g / p.Tp; d += alpha p; r_new = r_old - alpha h
// (r and r2 alternate with the inner iteration); z = M^-1 r_new.
__global__ __launch_bounds__(kBlock) void k_bupdate(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                    const int4 *__restrict__ work) {
    __shared__ double sh[4];
    const int4 w = work[blockIdx.x];
    const BCol &C = cols[w.w];
    if (C.s[1].phase != kPhaseInner) return;
    const BSeg &S = segs[w.x];
    if (w.z >= S.tblocks) {
        if (threadIdx.x == 0) {
            S.red_rz[(long)w.y * S.vblocks + w.z] = 0.0;
            S.red_ra[(long)w.y * S.vblocks + w.z] = 0.0;
        }
        return;
    }
    const long n = 3L * S.npol;
    const double gama = reduce_parts(S.red_hp + (long)w.y * S.tblocks, S.tblocks, sh);
    if (gama == 0.0) return;  // the direction kernel ends this inner solve
    const double a = C.s[1].g / gama;
    const bool odd = C.s[1].it & 1;
    const double *rold = odd ? S.r2 : S.r;
    double *rnew = odd ? S.r : S.r2;
    const bool nb = S.nb_row_ptr != nullptr && !C.nb_off;
    const int i = w.z * kBlock + threadIdx.x;
    const long colbase = (long)w.y * n;
    double vz = 0.0, va = 0.0;
    if (i < S.npol) {
        const long o = colbase + 3L * i;
        const double al = S.alpha[i];
        auto rj = [&](long q) { return fma(-a, S.h[q], rold[q]); };
        double rr[3], acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            S.d[o + c] += a * S.p[o + c];
            rr[c] = rj(o + c);
        }
        if (nb) nb_site(S, i, colbase, rj, acc);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double zz = nb ? kBatchNbDiag * (al * rr[c]) + acc[c] : al * rr[c];
            rnew[o + c] = rr[c];
            S.z[o + c] = zz;
            vz += rr[c] * zz;
            va += rr[c] * (al * rr[c]);
        }
    }
    site_partials(S, w, vz, va, sh);
}

// Inner PCG, direction (state copy 1 -> 0): inner convergence ends the
// refinement step with x += d; otherwise p = z + beta p.  r.z <= 0 (the
// preconditioner is not positive definite here) also ends it, and the
// column's later inner solves use alpha alone.
__global__ __launch_bounds__(kBlock) void k_bdirection(const BSeg *__restrict__ segs, BCol *__restrict__ cols,
                                                       const int4 *__restrict__ work) {
    __shared__ double sh[4];
    const int4 w = work[blockIdx.x];
    const BCol &C = cols[w.w];
    const BColSub st = C.s[1];
    BColSub nst = st;
    bool ended = false, breakdown = false;
    if (st.phase == kPhaseInner) {
        const BSeg &S = segs[w.x];
        const long n = 3L * S.npol;
        const double gama = reduce_parts(S.red_hp + (long)w.y * S.tblocks, S.tblocks, sh);
        const double gnew = gama == 0.0 ? st.g : reduce_parts(S.red_rz + (long)w.y * S.vblocks, S.vblocks, sh);
        const double ra = gama == 0.0 ? st.g : reduce_parts(S.red_ra + (long)w.y * S.vblocks, S.vblocks, sh);
        const double rms_in = sqrt(ra / (double)n);
        breakdown = gama != 0.0 && !(gnew > 0.0) && !(rms_in < st.tin);
        ended = gama == 0.0 || rms_in < st.tin || st.it + 1 >= C.maxit || breakdown;
        const long k = (long)w.z * kBlock + threadIdx.x;
        const long o = (long)w.y * n + k;
        if (ended) {
            if (k < n) S.x[o] += S.d[o];
            nst.phase = C.oneshot ? kPhaseDone : kPhaseOuter;
            if (C.oneshot) nst.status = 0;
            nst.inner_total = st.inner_total + st.it + 1;
        } else {
            const double beta = gnew / st.g;
            if (k < n) S.p[o] = S.z[o] + beta * S.p[o];
            nst.g = gnew;
            nst.it = st.it + 1;
        }
    }
    if (w.z == 0 && threadIdx.x == 0) {
        cols[w.w].s[0] = nst;
        if (ended) cols[w.w].rinit = kResMatvec;
        if (breakdown) cols[w.w].nb_off = 1;
    }
}

// The active columns of each segment in increasing order (one workgroup per
// segment): act[col0 + k] = k-th column with mv_active, act_n[segment] = count.
__global__ __launch_bounds__(kBlock) void k_bcompact(const BSeg *__restrict__ segs, const BCol *__restrict__ cols,
                                                     int use_p, int *__restrict__ act, int *__restrict__ act_n) {
    __shared__ int s_cnt[kBlock / 64];
    const BSeg &S = segs[blockIdx.x];
    const int nc = S.ncols, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int base = 0;
    for (int c0 = 0; c0 < nc; c0 += kBlock) {
        const int c = c0 + threadIdx.x;
        const bool a = c < nc && mv_active(cols[S.col0 + c], use_p);
        const unsigned long long m = __ballot(a);
        if (lane == 0) s_cnt[wv] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int q = 0; q < wv; q++) off += s_cnt[q];
        if (a) act[S.col0 + off + __popcll(m & ((1ull << lane) - 1ull))] = c;
        for (int q = 0; q < kBlock / 64; q++) base += s_cnt[q];
        __syncthreads();
    }
    if (threadIdx.x == 0) act_n[blockIdx.x] = base;
}

// pack: column groups from the active-column lists (wide batches, FP64)
template <typename R, int M, bool AMOEBA>
static void mv_launch(hipStream_t st, const BatchTables &t, const int4 *pair, int npair, const int4 *sym, int nsym,
                      const int4 *fin, int nfin, int use_p, bool pack = false) {
    if (npair == 0 && nsym == 0) return;
    const int *act = nullptr, *act_n = nullptr;
    if (pack && t.nseg > 0) {
        hipLaunchKernelGGL(k_bcompact, dim3(t.nseg), dim3(kBlock), 0, st, t.segs, t.cols, use_p, t.act, t.act_n);
        act = t.act;
        act_n = t.act_n;
    }
    if (npair)
        hipLaunchKernelGGL((k_bpairs<R, M, AMOEBA>), dim3(npair), dim3(kBlock), 0, st, t.segs, t.cols, pair, use_p,
                           act, act_n);
    if (nsym)
        hipLaunchKernelGGL((k_bsym<R, M, AMOEBA>), dim3(nsym), dim3(kSymBlock), 0, st, t.segs, t.cols, sym, use_p,
                           act, act_n);
    hipLaunchKernelGGL((k_bfinal<R, M, AMOEBA>), dim3(nfin), dim3(kBlock), 0, st, t.segs, t.cols, fin, use_p);
}

template <bool AMOEBA>
static void tick_t(hipStream_t st, const BatchTables &t, int m32, int m64, double eta, double frac) {
    if (t.npair64m) {
        hipLaunchKernelGGL(k_bcompact, dim3(t.nseg), dim3(kBlock), 0, st, t.segs, t.cols, 0, t.act, t.act_n);
        hipLaunchKernelGGL(k_bpairs_d64<AMOEBA>, dim3(t.npair64m), dim3(kBlock), 0, st, t.segs, t.pair64m, t.act,
                           t.act_n);
        hipLaunchKernelGGL((k_bfinal<double, kBatchM64, AMOEBA>), dim3(t.nfin64), dim3(kBlock), 0, st, t.segs, t.cols,
                           t.fin64, 0);
    } else if (m64 == 2)
        mv_launch<double, 2, AMOEBA>(st, t, t.pair64, t.npair64, t.sym64, t.nsym64, t.fin64, t.nfin64, 0);
    else
        mv_launch<double, kBatchM64, AMOEBA>(st, t, t.pair64, t.npair64, t.sym64, t.nsym64, t.fin64, t.nfin64, 0,
                                             true);
    hipLaunchKernelGGL(k_bresid, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec);
    hipLaunchKernelGGL(k_bdecide, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec, eta, frac);
    if (m32 == 2) {
        mv_launch<float, 2, AMOEBA>(st, t, t.pair32, t.npair32, t.sym32, t.nsym32, t.fin32, t.nfin32, 1);
    } else if (t.npair32) {
        hipLaunchKernelGGL(k_bconvert, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec);
        hipLaunchKernelGGL(k_bpairs_bx6<AMOEBA>, dim3(t.npair32), dim3(kBlock), 0, st, t.segs, t.cols, t.pair32);
        hipLaunchKernelGGL((k_bfinal<float, kBatchM32, AMOEBA>), dim3(t.nfin32), dim3(kBlock), 0, st, t.segs,
                           t.cols, t.fin32, 1);
    }
    hipLaunchKernelGGL(k_bupdate, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec);
    hipLaunchKernelGGL(k_bdirection, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec);
    HIP_CHECK(hipGetLastError());
}

void launch_batch_tick(hipStream_t st, const BatchTables &t, bool amoeba, int m32, int m64, double eta,
                       double frac) {
    if (t.nvec == 0) return;
    if (amoeba) tick_t<true>(st, t, m32, m64, eta, frac);
    else tick_t<false>(st, t, m32, m64, eta, frac);
}

template <bool AMOEBA>
static void inner_mv_t(hipStream_t st, const BatchTables &t) {
    hipLaunchKernelGGL(k_bconvert, dim3(t.nvec), dim3(kBlock), 0, st, t.segs, t.cols, t.vec);
    hipLaunchKernelGGL(k_bpairs_bx6<AMOEBA>, dim3(t.npair32), dim3(kBlock), 0, st, t.segs, t.cols, t.pair32);
    hipLaunchKernelGGL((k_bfinal<float, kBatchM32, AMOEBA>), dim3(t.nfin32), dim3(kBlock), 0, st, t.segs, t.cols,
                       t.fin32, 1);
    HIP_CHECK(hipGetLastError());
}

void launch_batch_inner_matvec(hipStream_t st, const BatchTables &t, bool amoeba) {
    if (t.nvec == 0 || t.npair32 == 0) return;
    if (amoeba) inner_mv_t<true>(st, t);
    else inner_mv_t<false>(st, t);
}

// ------------------------------------------------------------------ helpers
__global__ void k_batch_rhs(long n, int nsets, int ncols, const double *__restrict__ e_m2d,
                            const double *__restrict__ ext, double *__restrict__ b) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= n || c >= ncols) return;
    const int rhs = c / nsets, set = c % nsets;
    b[(long)c * n + k] = e_m2d[(long)set * n + k] + ext[(long)rhs * n + k];
}

void launch_batch_rhs(hipStream_t st, long n, int nsets, int ncols, const double *e_m2d, const double *ext,
                      double *b) {
    if (n == 0 || ncols == 0) return;
    hipLaunchKernelGGL(k_batch_rhs, dim3((unsigned)((n + kBlock - 1) / kBlock), ncols), dim3(kBlock), 0, st, n,
                       nsets, ncols, e_m2d, ext, b);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_charges_field(int npol, const double4 *__restrict__ ps, int nq,
                                                          const double4 *__restrict__ qc, double *__restrict__ ext) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= npol) return;
    const double4 p = ps[j];
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int q = 0; q < nq; q++) {
        const double4 c = qc[q];
        const double dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
        const double ir = rsqrt(dx * dx + dy * dy + dz * dz);
        const double f = c.w * ir * ir * ir;
        ex += f * dx;
        ey += f * dy;
        ez += f * dz;
    }
    ext[3L * j] = ex;
    ext[3L * j + 1] = ey;
    ext[3L * j + 2] = ez;
}

void launch_charges_field(hipStream_t st, int npol, const double4 *ps64, int nq, const double4 *charges,
                          double *ext) {
    if (npol == 0) return;
    hipLaunchKernelGGL(k_charges_field, dim3((npol + kBlock - 1) / kBlock), dim3(kBlock), 0, st, npol, ps64, nq,
                       charges, ext);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_batch_start(BCol *cols, int c0, int nc, double tol, int maxit, int rinit, int oneshot) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    BCol &C = cols[c0 + c];
    BColSub s;
    s.g = 0.0;
    s.tin = 0.0;
    s.rms = 0.0;
    s.phase = kPhaseOuter;
    s.it = 0;
    s.outer = 0;
    s.inner_total = 0;
    s.status = 2;
    s.pad = 0;
    C.s[0] = s;
    C.s[1] = s;
    C.tol = tol;
    C.maxit = maxit;
    C.rinit = rinit;
    C.oneshot = oneshot;
    C.nb_off = 0;
}

void launch_batch_start(hipStream_t st, BCol *cols, int c0, int nc, double tol, int maxit, int rinit,
                        int oneshot) {
    if (nc == 0) return;
    hipLaunchKernelGGL(k_batch_start, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cols, c0, nc, tol,
                       maxit, rinit, oneshot);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_batch_shared_rhs(long n, int nsets, const double *__restrict__ e_m2d, double *__restrict__ ext,
                                   double *__restrict__ b, double *__restrict__ bs) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double e = bs[k];
    bs[k] = e - ext[k];
    ext[k] = e;
    for (int set = 0; set < nsets; set++) b[(long)set * n + k] = e_m2d[(long)set * n + k] + e;
}

void launch_batch_shared_rhs(hipStream_t st, long n, int nsets, const double *e_m2d, double *ext, double *b,
                             double *bs) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_batch_shared_rhs, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n,
                       nsets, e_m2d, ext, b, bs);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_batch_feed(long n, int nsets, double *__restrict__ x, const double *__restrict__ xs) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double v = xs[k];
    for (int set = 0; set < nsets; set++) x[(long)set * n + k] += v;
}

void launch_batch_feed(hipStream_t st, long n, int nsets, double *x, const double *xs) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_batch_feed, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, nsets,
                       x, xs);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_batch_shift_rhs(long n, int nsets, int ncols, const double *__restrict__ e_m2d,
                                  const double *__restrict__ ext, double *__restrict__ b, double *__restrict__ r) {
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= n || c >= ncols) return;
    const int rhs = c / nsets, set = c % nsets;
    const long o = (long)c * n + k;
    const double bn = e_m2d[(long)set * n + k] + ext[(long)rhs * n + k];
    r[o] += bn - b[o];
    b[o] = bn;
}

void launch_batch_shift_rhs(hipStream_t st, long n, int nsets, int ncols, const double *e_m2d, const double *ext,
                            double *b, double *r) {
    if (n == 0 || ncols == 0) return;
    hipLaunchKernelGGL(k_batch_shift_rhs, dim3((unsigned)((n + kBlock - 1) / kBlock), ncols), dim3(kBlock), 0, st,
                       n, nsets, ncols, e_m2d, ext, b, r);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_benergy_part(long n, int nsets, const double *__restrict__ x,
                                                         const double *__restrict__ e_m2d,
                                                         double *__restrict__ red) {
    __shared__ double sh[4];
    const int rhs = blockIdx.y;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    // mu_D of this right-hand side against the P-set static field
    double v = k < n ? x[(long)rhs * nsets * n + k] * e_m2d[(long)(nsets - 1) * n + k] : 0.0;
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[(long)rhs * gridDim.x + blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_benergy_final(const double *__restrict__ red, int nred,
                                                          double *__restrict__ out) {
    __shared__ double sh[4];
    const double s = reduce_parts(red + (long)blockIdx.x * nred, nred, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = -0.5 * s;
}

void launch_batch_energy(hipStream_t st, long n, int nsets, int nrhs, const double *x, const double *e_m2d,
                         double *red, double *out) {
    if (nrhs == 0) return;
    const int nb = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_benergy_part, dim3(nb, nrhs), dim3(kBlock), 0, st, n, nsets, x, e_m2d, red);
    hipLaunchKernelGGL(k_benergy_final, dim3(nrhs), dim3(kBlock), 0, st, red, nb, out);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
