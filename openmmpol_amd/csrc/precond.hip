// Near-field preconditioner of the inner FP32 PCG of the mixed solver.
//
// The reference's CG is preconditioned by the polarizabilities alone
// (z = alpha r, mod_polarization.F90:535-557), and the plain FP64 solver keeps
// exactly that.  The inner solves of the mixed solver only have to reduce an
// FP64 residual (every outer step re-checks it with the reference's FP64
// residual and stopping rule), so they may use any symmetric positive definite
// preconditioner.  Here it is the first-order Neumann series of T^-1 around its
// diagonal, restricted to the pairs closer than kNbCut:
//
//   z_i = c alpha_i r_i - alpha_i sum_{j near i} T_ij alpha_j r_j
//       = c alpha_i r_i + sum_{j near i} B_ij r_j,
//   B_ij = alpha_i alpha_j s_ij (k3 d d^T - k2 I)    (damped dipole-field tensor),
//
// with the diagonal weighted by c = 2 so that it stays positive definite for
// strongly coupled neighbours.  On the 15,000-site sphere it cuts the inner
// iterations of a cold solve by about a quarter.  B_ij = B_ji bit for bit (the
// same operations on the same operands), so the preconditioner is exactly
// symmetric.  If it ever gives r.z <= 0 the direction kernel ends that inner
// solve and switches it off for the rest of the geometry (the flag nb_off);
// each inner solve reads the flag once, at its start, so its preconditioner
// never changes halfway.
//
// The neighbour list is rebuilt on the device for every geometry: a count
// sweep over all pairs split by source ranges, one scan, and a fill sweep that
// writes each target's neighbours in source order (deterministic).  It holds
// at most kNbCapPerSite neighbours per site on average; beyond that nb_off is
// set and the inner solves use alpha alone.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "pair_math.h"

namespace ommp {

namespace {

constexpr float kNbCut = 3.0f * 1.8897261245650f;  // 3 Angstrom in bohr
constexpr double kNbDiag = 2.0;
constexpr int kNbScanThreads = 1024;

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_nb_count(int npol, int chunk, const float4 *__restrict__ ps, float rc2,
                                                     int *__restrict__ cnt) {
    __shared__ float4 s_pos[kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const float4 pj = j < npol ? ps[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int i0 = blockIdx.y * chunk, i1 = min(npol, i0 + chunk);
    int c = 0;
    for (int base = i0; base < i1; base += kBlock) {
        const int m = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < m) s_pos[tid] = ps[base + tid];
        __syncthreads();
        for (int k = 0; k < m; k++) {
            const float dx = pj.x - s_pos[k].x, dy = pj.y - s_pos[k].y, dz = pj.z - s_pos[k].z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            c += (r2 < rc2 && base + k != j) ? 1 : 0;
        }
    }
    if (j < npol) cnt[(long)blockIdx.y * npol + j] = c;
}

// One block: row_ptr[i] = exclusive prefix of the per-target totals, off[s][i]
// = where split s of target i starts writing, *nb_off = 1 when the total
// exceeds cap (and 0 otherwise: a rebuild also clears a breakdown).
__global__ __launch_bounds__(kNbScanThreads) void k_nb_scan(int npol, int nsplit, const int *__restrict__ cnt,
                                                            int *__restrict__ row_ptr, int *__restrict__ off,
                                                            long cap, int *__restrict__ nb_off) {
    __shared__ long s_sum[kNbScanThreads];
    const int t = threadIdx.x;
    const int per = (npol + kNbScanThreads - 1) / kNbScanThreads;
    const int r0 = min(npol, t * per), r1 = min(npol, r0 + per);
    long tot = 0;
    for (int i = r0; i < r1; i++)
        for (int s = 0; s < nsplit; s++) tot += cnt[(long)s * npol + i];
    s_sum[t] = tot;
    __syncthreads();
    // Hillis-Steele inclusive scan over the thread totals
    for (int d = 1; d < kNbScanThreads; d <<= 1) {
        const long v = t >= d ? s_sum[t - d] : 0;
        __syncthreads();
        s_sum[t] += v;
        __syncthreads();
    }
    long pos = s_sum[t] - tot;
    for (int i = r0; i < r1; i++) {
        row_ptr[i] = (int)min(pos, cap);
        for (int s = 0; s < nsplit; s++) {
            off[(long)s * npol + i] = (int)min(pos, cap);
            pos += cnt[(long)s * npol + i];
        }
    }
    if (t == kNbScanThreads - 1) {
        const long total = s_sum[t];
        row_ptr[npol] = (int)min(total, cap);
        *nb_off = total > cap ? 1 : 0;
    }
}

template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_nb_fill(int npol, int chunk, const float4 *__restrict__ ps,
                                                    const float *__restrict__ pa, const double *__restrict__ alpha,
                                                    const int *__restrict__ pp_ri, const int *__restrict__ pp_ci,
                                                    const double *__restrict__ pp_f, float rc2,
                                                    const int *__restrict__ off, const int *__restrict__ nb_off,
                                                    int *__restrict__ nb_ci, float *__restrict__ nb_b) {
    if (*nb_off) return;
    __shared__ float4 s_pos[kBlock];
    __shared__ float s_a[kBlock], s_al[kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < npol;
    const float4 pj = valid ? ps[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float aj = valid ? pa[j] : 0.39f;
    const float alj = valid ? (float)alpha[j] : 0.f;
    const int i0 = blockIdx.y * chunk, i1 = min(npol, i0 + chunk);
    int o = valid ? off[(long)blockIdx.y * npol + j] : 0;
    for (int base = i0; base < i1; base += kBlock) {
        const int m = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < m) {
            s_pos[tid] = ps[base + tid];
            s_a[tid] = pa[base + tid];
            s_al[tid] = (float)alpha[base + tid];
        }
        __syncthreads();
        if (!valid) continue;
        for (int k = 0; k < m; k++) {
            const int i = base + k;
            const float4 pi = s_pos[k];
            const float dx0 = pj.x - pi.x, dy0 = pj.y - pi.y, dz0 = pj.z - pi.z;
            const float r2 = fmaf(dz0, dz0, fmaf(dy0, dy0, dx0 * dx0));
            if (!(r2 < rc2) || i == j) continue;
            float dx, dy, dz, k2, k3, arg;
            bool near;
            pair_geom<float, AMOEBA>(pj, aj, pi, s_a[k], dx, dy, dz, k2, k3, arg, near);
            pair_damp<float, AMOEBA>(arg, near, k2, k3);
            // P-P screening of the pair (list_P_P by target, f = 1 - s)
            float sij = 1.0f;
            for (int q = pp_ri[j]; q < pp_ri[j + 1]; q++)
                if (pp_ci[q] == i) sij = (float)(1.0 - pp_f[q]);
            const float c = (alj * s_al[k]) * sij;
            const float ck3 = c * k3, ck2 = c * k2;
            float *b = nb_b + 6L * o;
            b[0] = ck3 * (dx * dx) - ck2;
            b[1] = ck3 * (dx * dy);
            b[2] = ck3 * (dx * dz);
            b[3] = ck3 * (dy * dy) - ck2;
            b[4] = ck3 * (dy * dz);
            b[5] = ck3 * (dz * dz) - ck2;
            nb_ci[o] = i;
            o++;
        }
    }
}

// z_i = c alpha_i r_i + sum_j B_ij r_j with r_j = rold_j - a h_j evaluated
// exactly as the owning thread evaluates its own (fma(-a, h, rold)).
__device__ __forceinline__ void nb_apply(int i, long base, double a, const double *__restrict__ rold,
                                         const double *__restrict__ h, double ci, const double rn[3],
                                         const int *__restrict__ row_ptr, const int *__restrict__ nb_ci,
                                         const float *__restrict__ nb_b, double z[3]) {
    z[0] = ci * rn[0];
    z[1] = ci * rn[1];
    z[2] = ci * rn[2];
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        const long o = base + 3L * nb_ci[k];
        const double r0 = h ? fma(-a, h[o], rold[o]) : rold[o];
        const double r1 = h ? fma(-a, h[o + 1], rold[o + 1]) : rold[o + 1];
        const double r2 = h ? fma(-a, h[o + 2], rold[o + 2]) : rold[o + 2];
        const float *b = nb_b + 6L * k;
        const double bxx = b[0], bxy = b[1], bxz = b[2], byy = b[3], byz = b[4], bzz = b[5];
        z[0] += bxx * r0 + bxy * r1 + bxz * r2;
        z[1] += bxy * r0 + byy * r1 + byz * r2;
        z[2] += bxz * r0 + byz * r1 + bzz * r2;
    }
}

// Inner solve start from a zero guess: x = 0, r = b, z = M^-1 b, p = z,
// per-block partials of r.z (one thread per site and column).
__global__ __launch_bounds__(kBlock) void k_cg_init_nb(int npol, const double *__restrict__ b, double *__restrict__ x,
                                                       double *__restrict__ r, double *__restrict__ z,
                                                       double *__restrict__ p, const double *__restrict__ alpha,
                                                       const int *__restrict__ row_ptr, const int *__restrict__ nb_ci,
                                                       const float *__restrict__ nb_b, const int *__restrict__ nb_off,
                                                       CgState *cs, double *__restrict__ red) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const long n = 3L * npol, base = m * n;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int off = *nb_off;
    // the choice holds for the whole inner solve (k_cg_init2 leaves pad alone)
    if (blockIdx.x == 0 && threadIdx.x == 0) cs[m].pad = off;
    double v = 0.0;
    if (i < npol) {
        const long o = base + 3L * i;
        const double rn[3] = {b[o], b[o + 1], b[o + 2]};
        double zz[3];
        if (off) {
            for (int c = 0; c < 3; c++) zz[c] = alpha[i] * rn[c];
        } else {
            nb_apply(i, base, 0.0, b, nullptr, kNbDiag * alpha[i], rn, row_ptr, nb_ci, nb_b, zz);
        }
        for (int c = 0; c < 3; c++) {
            x[o + c] = 0.0;
            r[o + c] = rn[c];
            z[o + c] = zz[c];
            p[o + c] = zz[c];
            v += rn[c] * zz[c];
        }
    }
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[m * gridDim.x + blockIdx.x] = v;
}

// conjugate_gradient_solver loop body, first half, with the near-field
// preconditioner: x += a p, r_new = r_old - a h, z = M^-1 r_new; partials of
// r.z (for beta) and r.(alpha r) (the reference's convergence measure).
__global__ __launch_bounds__(kBlock) void k_cg_update_nb(int npol, int it, const double *__restrict__ h,
                                                         double *__restrict__ x, const double *__restrict__ rold,
                                                         double *__restrict__ rnew, double *__restrict__ z,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ alpha,
                                                         const double *__restrict__ red_hp, int nred_hp,
                                                         double *__restrict__ red_rz, double *__restrict__ red_ra,
                                                         CgState *cs, const int *__restrict__ row_ptr,
                                                         const int *__restrict__ nb_ci,
                                                         const float *__restrict__ nb_b) {
    __shared__ double sh[4];
    __shared__ int s_done, s_off;
    const int m = blockIdx.y;
    if (threadIdx.x == 0) {
        s_done = cs[m].done;
        s_off = cs[m].pad;
    }
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
    const double a = cs[m].g[(it + 1) & 1] / gama;
    const long n = 3L * npol, base = m * n;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double vz = 0.0, va = 0.0;
    if (i < npol) {
        const long o = base + 3L * i;
        double rn[3], zz[3];
        for (int c = 0; c < 3; c++) {
            x[o + c] += a * p[o + c];
            rn[c] = fma(-a, h[o + c], rold[o + c]);
        }
        const double al = alpha[i];
        if (s_off) {
            for (int c = 0; c < 3; c++) zz[c] = al * rn[c];
        } else {
            nb_apply(i, base, a, rold, h, kNbDiag * al, rn, row_ptr, nb_ci, nb_b, zz);
        }
        for (int c = 0; c < 3; c++) {
            rnew[o + c] = rn[c];
            z[o + c] = zz[c];
            vz += rn[c] * zz[c];
            va += rn[c] * (al * rn[c]);
        }
    }
    vz = block_sum_all(vz, sh);
    if (threadIdx.x == 0) red_rz[m * gridDim.x + blockIdx.x] = vz;
    va = block_sum_all(va, sh);
    if (threadIdx.x == 0) red_ra[m * gridDim.x + blockIdx.x] = va;
}

// Second half: convergence on sqrt(r.(alpha r) / n), as the reference measures
// it (mod_solvers.F90:188-199), beta = r.z / r.z_old; r.z <= 0 means the
// preconditioner is not positive definite here: the column stops (its progress
// so far is kept) and nb_off switches it off until the next geometry.
__global__ __launch_bounds__(kBlock) void k_cg_direction_nb(long n, int it, const double *__restrict__ z,
                                                            double *__restrict__ p, const double *__restrict__ red_rz,
                                                            const double *__restrict__ red_ra, int nred, CgState *cs,
                                                            int *__restrict__ nb_off) {
    __shared__ double sh[4];
    __shared__ int s_done;
    const int m = blockIdx.y;
    if (threadIdx.x == 0) s_done = cs[m].done;
    __syncthreads();
    if (s_done) return;
    const double gnew = reduce_parts(red_rz + m * nred, nred, sh);
    const double ra = reduce_parts(red_ra + m * nred, nred, sh);
    const double rms = sqrt(ra / (double)n);
    if (rms < cs[m].tol || !(gnew > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            cs[m].rms = rms;
            cs[m].status = rms < cs[m].tol ? 0 : 3;
            cs[m].iters = it;
            cs[m].done = 1;
            if (!(rms < cs[m].tol)) *nb_off = 1;
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

}  // namespace

int nb_split(int npol) {
    const int sb = std::max(1, site_blocks(npol));
    const int max_s = std::max(1, (npol + kBlock - 1) / kBlock);
    return std::max(1, std::min(max_s, (1024 + sb - 1) / sb));
}

void launch_nb_build(hipStream_t st, bool amoeba, int npol, const float4 *ps32, const float *pa32,
                     const double *alpha, const CsrDev &pp, int nsplit, int *cnt, int *off, int *row_ptr, long cap,
                     int *nb_ci, float *nb_b, int *nb_off) {
    if (npol == 0) return;
    const int chunk = (npol + nsplit - 1) / nsplit;
    const dim3 grid(site_blocks(npol), nsplit);
    const float rc2 = kNbCut * kNbCut;
    if (amoeba)
        hipLaunchKernelGGL(k_nb_count<true>, grid, dim3(kBlock), 0, st, npol, chunk, ps32, rc2, cnt);
    else
        hipLaunchKernelGGL(k_nb_count<false>, grid, dim3(kBlock), 0, st, npol, chunk, ps32, rc2, cnt);
    hipLaunchKernelGGL(k_nb_scan, dim3(1), dim3(kNbScanThreads), 0, st, npol, nsplit, cnt, row_ptr, off, cap, nb_off);
    if (amoeba)
        hipLaunchKernelGGL(k_nb_fill<true>, grid, dim3(kBlock), 0, st, npol, chunk, ps32, pa32, alpha, pp.ri, pp.ci,
                           pp.f, rc2, off, nb_off, nb_ci, nb_b);
    else
        hipLaunchKernelGGL(k_nb_fill<false>, grid, dim3(kBlock), 0, st, npol, chunk, ps32, pa32, alpha, pp.ri, pp.ci,
                           pp.f, rc2, off, nb_off, nb_ci, nb_b);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_init_nb(hipStream_t st, int npol, int M, const double *b, double *x, double *r, double *z, double *p,
                       const double *alpha, const NbDev &nb, double *red, CgState *cs, const double *tol,
                       const CgState *pre) {
    const int sb = site_blocks(npol);
    hipLaunchKernelGGL(k_cg_init_nb, dim3(sb, M), dim3(kBlock), 0, st, npol, b, x, r, z, p, alpha, nb.row_ptr,
                       nb.ci, nb.b, nb.off_flag, cs, red);
    HIP_CHECK(hipGetLastError());
    launch_cg_init2(st, M, red, sb, cs, tol, pre);
}

void launch_cg_update_nb(hipStream_t st, int npol, int M, int it, const double *h, double *x, const double *rold,
                         double *rnew, double *z, const double *p, const double *alpha, const double *red_hp,
                         int nred_hp, double *red_rz, double *red_ra, CgState *cs, const NbDev &nb) {
    hipLaunchKernelGGL(k_cg_update_nb, dim3(site_blocks(npol), M), dim3(kBlock), 0, st, npol, it, h, x, rold, rnew,
                       z, p, alpha, red_hp, nred_hp, red_rz, red_ra, cs, nb.row_ptr, nb.ci, nb.b);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_direction_nb(hipStream_t st, int npol, int M, int it, const double *z, double *p,
                            const double *red_rz, const double *red_ra, CgState *cs, const NbDev &nb) {
    const long n = 3L * npol;
    hipLaunchKernelGGL(k_cg_direction_nb, dim3(vec_blocks(n), M), dim3(kBlock), 0, st, n, it, z, p, red_rz, red_ra,
                       site_blocks(npol), cs, nb.off_flag);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
