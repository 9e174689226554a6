// Static field of the multipoles at the polarizable sites (E_M2D) with each
// unordered pair of polarizable sites evaluated once.
//
// The field at a polarizable site t has two parts:
//  * the polarizable sources, swept pair-once over blocks of kSymBlock sites as
//    the T matvec does (tmv_sym.h: one workgroup per block pair I <= J, the
//    sources of J in LDS, lane l meeting source (l + s) mod 64 at step s, the
//    source-side sums carried along by a DPP wave rotate, block-pair partial
//    sums added in a fixed order);
//  * the non-polarizable sources (never targets), swept one-sided by
//    k_rhs_pairs over their index list.
// For a pair (t, s) with d = r_t - r_s the damped kernels k2 = l3/r^3,
// k3 = l5/r^5 and k4 = l7/r^7 (maxder 3; damped_coulomb_kernel,
// mod_electrostatics.F90:1018-1095) are shared by both directions; the field at
// t of the multipoles of s (q_elec_prop, mu_elec_prop, quad_elec_prop,
// mod_electrostatics.F90:1097-1405, quadrupoles divided by 3) is
//   E_t += (q_s k2 + 3 (mu_s.d) k3 + 15 (d.Th_s.d) k4) d - mu_s k2 - 6 (Th_s d) k3
// and the field at s of the multipoles of t the same with d -> -d.  Wang
// (AMBER) systems carry charges only (k2 with the Wang damping).  Scale 1 for
// every pair: the finalize subtracts (1 - s_d) and (1 - s_p) times the field of
// the listed pairs, as for the per-target sweep (kernels.hip).
//
// The geometry, the reciprocal square root and the damping of a pair, a third
// of its work, are done once instead of twice; the contractions stay per
// direction.
#include <hip/hip_runtime.h>

#include "device.h"
#include "pair_math.h"
#include "tmv_sym.h"

namespace ommp {

namespace {

// Damped kernels of one pair, d = r_t - r_s.  AMOEBA: w = (1/thole)^3 of both
// sites (a u^3 = a r^3 w_t w_s), kernels up to k4; Wang: w = 1/thole, k2 only.
template <bool AMOEBA>
__device__ __forceinline__ void rhs_geom(const double4 &pt, double at, const double4 &ps, double as, double &dx,
                                         double &dy, double &dz, double &k2, double &k3, double &k4) {
    dx = pt.x - ps.x;
    dy = pt.y - ps.y;
    dz = pt.z - ps.z;
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    const double ir = fast_rsqrt(r2);
    const double ir2 = ir * ir;
    k2 = ir * ir2;
    if (AMOEBA) {
        k3 = k2 * ir2;
        k4 = k3 * ir2;
        const double r3 = (r2 * ir) * r2;
        const double arg = fmin(at, as) * (r3 * (pt.w * ps.w));
        const bool near = arg < 50.0;
        if (__ballot(near)) {
            const double e = exp(-arg);
            k2 = near ? k2 * (1.0 - e) : k2;
            k3 = near ? k3 * (1.0 - (1.0 + arg) * e) : k3;
            k4 = near ? k4 * (1.0 - (1.0 + arg + 0.6 * arg * arg) * e) : k4;
        }
    } else {
        k3 = k4 = 0.0;
        const double u = (r2 * ir) * (pt.w * ps.w);
        const bool near = u < 1.0;
        if (__ballot(near)) {
            const double u3 = u * u * u;
            k2 = near ? k2 * (4.0 * u3 - 3.0 * u3 * u) : k2;
        }
    }
}

// e += field of the multipoles q (q, mu, quadrupole xx xy yy xz yz zz) at
// displacement d from them; k3t = 3 k3, k3s = 6 k3, k4f = 15 k4.
template <bool AMOEBA>
__device__ __forceinline__ void rhs_accum(const double *q, double dx, double dy, double dz, double k2, double k3t,
                                          double k3s, double k4f, double (&e)[3]) {
    if (AMOEBA) {
        const double md = fma(q[3], dz, fma(q[2], dy, q[1] * dx));
        const double qx = fma(q[7], dz, fma(q[5], dy, q[4] * dx));
        const double qy = fma(q[8], dz, fma(q[6], dy, q[5] * dx));
        const double qz = fma(q[9], dz, fma(q[8], dy, q[7] * dx));
        const double qrr = fma(qz, dz, fma(qy, dy, qx * dx));
        const double cd = fma(qrr, k4f, fma(md, k3t, q[0] * k2));
        e[0] = fma(-k3s, qx, fma(-k2, q[1], fma(cd, dx, e[0])));
        e[1] = fma(-k3s, qy, fma(-k2, q[2], fma(cd, dy, e[1])));
        e[2] = fma(-k3s, qz, fma(-k2, q[3], fma(cd, dz, e[2])));
    } else {
        const double c = q[0] * k2;
        e[0] = fma(c, dx, e[0]);
        e[1] = fma(c, dy, e[1]);
        e[2] = fma(c, dz, e[2]);
    }
}

// One workgroup per block pair (I <= J) of polarizable sites.  Target-side sums
// of component c go to part_i[p][c][t] (t in I), source-side ones to
// part_j[p][c][t] (t in J).  Padding sites (past npol, last block only) sit far
// away at distinct points with zero multipoles, so that they add exactly 0 and
// no pair but the skipped self pair has a zero distance.
template <bool AMOEBA>
__global__ __launch_bounds__(kSymBlock) void k_rhs_sym(int npol, const int2 *__restrict__ pairs,
                                                       const double4 *__restrict__ ps, const double *__restrict__ pa,
                                                       const double *__restrict__ mq, const int *__restrict__ polar_mm,
                                                       double *__restrict__ part_i, double *__restrict__ part_j) {
    constexpr int LD = AMOEBA ? 10 : 1;
    constexpr int NQ = AMOEBA ? 6 : 1;  // double2 per source: q0..q9 and a (AMOEBA), q0 (Wang)
    constexpr int KA = AMOEBA ? 10 : 1;  // where a sits (AMOEBA)
    __shared__ double4 s_pos[kSymBlock];
    __shared__ double2 s_q[NQ][kSymBlock];
    __shared__ double s_acc[kSymWaves][3][64];
    const int2 pr = pairs[blockIdx.x];
    const int I = pr.x, J = pr.y;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double far = 1e15, gap = 1e9;
    {
        const int jj = J * kSymBlock + tid;
        double4 p;
        double q[2 * NQ];
        if (jj < npol) {
            p = ps[jj];
            if (AMOEBA) p.w = p.w * p.w * p.w;
            const double *src = mq + (long)LD * polar_mm[jj];
#pragma unroll
            for (int c = 0; c < LD; c++) q[c] = src[c];
            q[2 * NQ - 1] = 0.0;
            if (AMOEBA) q[KA] = pa[jj];
        } else {
            p.x = p.y = p.z = far + gap * (double)tid;
            p.w = 1.0;
#pragma unroll
            for (int c = 0; c < 2 * NQ; c++) q[c] = 0.0;
            if (AMOEBA) q[KA] = 0.39;
        }
        s_pos[tid] = p;
#pragma unroll
        for (int k = 0; k < NQ; k++) s_q[k][tid] = make_double2(q[2 * k], q[2 * k + 1]);
    }
    const int i = I * kSymBlock + tid;
    double4 pt;
    double at = 0.39, qt[LD];
    if (i < npol) {
        pt = ps[i];
        if (AMOEBA) pt.w = pt.w * pt.w * pt.w;
        at = pa[i];
        const double *src = mq + (long)LD * polar_mm[i];
#pragma unroll
        for (int c = 0; c < LD; c++) qt[c] = src[c];
    } else {
        pt.x = pt.y = pt.z = far + gap * (double)tid;
        pt.w = 1.0;
#pragma unroll
        for (int c = 0; c < LD; c++) qt[c] = 0.0;
    }
    auto load_src = [&](int js, double4 &psrc, double (&q)[2 * NQ]) {
        psrc = s_pos[js];
#pragma unroll
        for (int c = 0; c < NQ; c++) {
            const double2 v = s_q[c][js];
            q[2 * c] = v.x;
            q[2 * c + 1] = v.y;
        }
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
            for (int s = 1; s < 64; s++) {
                const int js = k * 64 + ((l + s) & 63);
                double4 psrc;
                double q[2 * NQ];
                load_src(js, psrc, q);
                double dx, dy, dz, k2, k3, k4;
                rhs_geom<AMOEBA>(pt, at, psrc, q[KA], dx, dy, dz, k2, k3, k4);
                rhs_accum<AMOEBA>(q, dx, dy, dz, k2, 3.0 * k3, 6.0 * k3, 15.0 * k4, acc);
            }
        } else if (!mirror) {
            for (int s = 0; s < 64; s++) {
                const int js = k * 64 + ((l + s) & 63);
                double4 psrc;
                double q[2 * NQ];
                load_src(js, psrc, q);
                double dx, dy, dz, k2, k3, k4;
                rhs_geom<AMOEBA>(pt, at, psrc, q[KA], dx, dy, dz, k2, k3, k4);
                const double k3t = 3.0 * k3, k3s = 6.0 * k3, k4f = 15.0 * k4;
                rhs_accum<AMOEBA>(q, dx, dy, dz, k2, k3t, k3s, k4f, acc);
                // the field at the source of this lane's target, d reversed
                rhs_accum<AMOEBA>(qt, -dx, -dy, -dz, k2, k3t, k3s, k4f, jacc);
#pragma unroll
                for (int c = 0; c < 3; c++) jacc[c] = lane_next(jacc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) s_acc[w][c][l] = jacc[c];
        __syncthreads();
        // source side of sub-tile k: the waves' rows added in wave order
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

// E_M2D of 64 sites per workgroup (kRhsSub workgroups per block of kSymBlock):
// kRhsQ groups of 64 threads share each site; group q adds every kRhsQ-th
// partial of the site's row (pairs (b, J >= b): target sides), column (pairs
// (I <= b, b): source sides) and non-polarizable source split, and every
// kRhsQ-th entry of its correction list (f = 1 - s for the D and P sets); group
// 0 adds the group results in a fixed order.
constexpr int kRhsQ = 8;
constexpr int kRhsSub = kSymBlock / 64;

template <bool AMOEBA>
__global__ __launch_bounds__(kRhsQ * 64) void k_rhs_sym_finalize(
    int npol, int nb, const double *__restrict__ part, const double *__restrict__ part_np, int nsplit_np,
    const double4 *__restrict__ ms, const double *__restrict__ ma, const double *__restrict__ mq,
    const double4 *__restrict__ ps, const double *__restrict__ pa, const int *__restrict__ sp_ri,
    const int *__restrict__ sp_ci, const double2 *__restrict__ sp_f, double *__restrict__ e_m2d) {
    constexpr int LD = AMOEBA ? 10 : 1;
    __shared__ double shd[kRhsQ - 1][3][64];
    __shared__ double shp[kRhsQ - 1][3][64];
    const int q = threadIdx.x >> 6, li = threadIdx.x & 63;
    const int b = blockIdx.x / kRhsSub, t = (blockIdx.x % kRhsSub) * 64 + li;  // t: site within block b
    const int j = b * kSymBlock + t;
    const bool valid = j < npol;
    const long np = (long)nb * (nb + 1) / 2, stride = 3L * kSymBlock;
    double ed[3] = {0.0, 0.0, 0.0}, ep[3] = {0.0, 0.0, 0.0};
    if (valid) {
        double sum[3] = {0.0, 0.0, 0.0};
        // pair (I, J) has index I nb - I (I - 1) / 2 + J - I
        const long row = (long)b * nb - (long)b * (b - 1) / 2;
        for (int J = b + q; J < nb; J += kRhsQ) {
            const double *p = part + (row + J - b) * stride + t;
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
        const double *part_j = part + np * stride;
        for (int I = q; I <= b; I += kRhsQ) {
            const double *p = part_j + ((long)I * nb - (long)I * (I - 1) / 2 + b - I) * stride + t;
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += p[(long)c * kSymBlock];
        }
        for (int s = q; s < nsplit_np; s += kRhsQ) {
            const double *p = part_np + (long)s * 3 * npol + 3L * j;
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += p[c];
        }
        double cd[3] = {0.0, 0.0, 0.0}, cp[3] = {0.0, 0.0, 0.0};
        const double4 pj = ps[j];
        const double aj = pa[j];
        for (int k = sp_ri[j] + q; k < sp_ri[j + 1]; k += kRhsQ) {
            const int i = sp_ci[k];
            const double2 f = sp_f[k];
            const double4 pi = ms[i];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double ir = rsqrt(dx * dx + dy * dy + dz * dz);
            double qs[LD];
#pragma unroll
            for (int c = 0; c < LD; c++) qs[c] = mq[(long)LD * i + c];
            double fx, fy, fz;
            m2d_pair<AMOEBA>(dx, dy, dz, ir, pi.w, pj.w, ma[i], aj, qs, fx, fy, fz);
            cd[0] += f.x * fx;
            cd[1] += f.x * fy;
            cd[2] += f.x * fz;
            cp[0] += f.y * fx;
            cp[1] += f.y * fy;
            cp[2] += f.y * fz;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ed[c] = sum[c] - cd[c];
            ep[c] = sum[c] - cp[c];
        }
    }
    if (q > 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            shd[q - 1][c][li] = ed[c];
            shp[q - 1][c][li] = ep[c];
        }
    }
    __syncthreads();
    if (q != 0 || !valid) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double d = ed[c], p = ep[c];
#pragma unroll
        for (int g = 0; g < kRhsQ - 1; g++) {
            d += shd[g][c][li];
            p += shp[g][c][li];
        }
        if (AMOEBA) {
            e_m2d[3L * j + c] = d;
            e_m2d[3L * npol + 3L * j + c] = p;
        } else {
            e_m2d[3L * j + c] = p;
        }
    }
}

}  // namespace

// From this size the pair-once sweep beats the per-target one, while its
// partial sums fit in the T matvec's pair-once buffer (2 GiB at most).
// Measured on MI355X (sweep kernels alone, AMOEBA / Wang): 0.96x / 1.09x at
// 4,000 sites, 1.27x / 1.29x at 5,000, 1.62x / 2.14x at 15,000, 1.76x / 2.30x
// at 30,000.
constexpr int kRhsPairOnceFrom = 5000;
bool rhs_pair_once(int npol) {
    return npol >= kRhsPairOnceFrom && tmv_sym_part_bytes(npol, 2, 64) <= (size_t(2) << 30);
}

void launch_rhs_sym(hipStream_t st, bool amoeba, int nmm, int npol, const int2 *pairs, const double4 *ms64,
                    const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                    const int *polar_mm, const int *np_src, const CsrDev &sp, void *sympart, double *part_np,
                    int nsplit_np, double *e_m2d, hipEvent_t evb, hipEvent_t eve) {
    if (npol == 0) return;
    const int nb = tmv_sym_blocks(npol);
    const long np = (long)nb * (nb + 1) / 2;
    double *part_i = (double *)sympart, *part_j = part_i + np * 3 * kSymBlock;
    const int nnp = nmm - npol;
    if (evb) HIP_CHECK(hipEventRecord(evb, st));
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_sym<true>, dim3((unsigned)np), dim3(kSymBlock), 0, st, npol, pairs, ps64, pa64, mq,
                           polar_mm, part_i, part_j);
    else
        hipLaunchKernelGGL(k_rhs_sym<false>, dim3((unsigned)np), dim3(kSymBlock), 0, st, npol, pairs, ps64, pa64,
                           mq, polar_mm, part_i, part_j);
    HIP_CHECK(hipGetLastError());
    if (nnp > 0) launch_rhs_pairs_list(st, amoeba, nnp, np_src, npol, ms64, ma64, mq, ps64, pa64, part_np, nsplit_np);
    if (eve) HIP_CHECK(hipEventRecord(eve, st));
    const int nsp = nnp > 0 ? nsplit_np : 0;
    if (amoeba)
        hipLaunchKernelGGL(k_rhs_sym_finalize<true>, dim3(nb * kRhsSub), dim3(kRhsQ * 64), 0, st, npol, nb, part_i,
                           part_np, nsp, ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    else
        hipLaunchKernelGGL(k_rhs_sym_finalize<false>, dim3(nb * kRhsSub), dim3(kRhsQ * 64), 0, st, npol, nb, part_i,
                           part_np, nsp, ms64, ma64, mq, ps64, pa64, sp.ri, sp.ci, sp.f2, e_m2d);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
