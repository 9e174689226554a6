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
// The neighbour lists (pairs closer than the cut, each row in site order) are
// a function of the coordinates alone and are built on the host with a cell
// grid when the coordinates are uploaded (engine.cpp), next to the
// coincident-site check; the tensors B are computed on the device, one thread
// per list entry, before the first mixed solve of a geometry.  Four lanes share
// a site in the CG kernels below, each summing every fourth neighbour, so that
// a site's chain of dependent loads is short.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "pair_math.h"

namespace ommp {

namespace {

constexpr double kNbDiag = 2.0;

// B for every list entry e (target rowi[e], source ci[e]), FP32 from the FP32
// positions: the same operations on the same operands for (i, j) and (j, i),
// so B_ij = B_ji bit for bit.
template <bool AMOEBA>
__global__ __launch_bounds__(kBlock) void k_nb_tensors(int nnz, const int *__restrict__ rowi,
                                                       const int *__restrict__ ci, const float4 *__restrict__ ps,
                                                       const float *__restrict__ pa, const double *__restrict__ alpha,
                                                       const int *__restrict__ pp_ri, const int *__restrict__ pp_ci,
                                                       const double *__restrict__ pp_f, float *__restrict__ nb_b) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nnz) return;
    const int j = rowi[e], i = ci[e];
    float dx, dy, dz, k2, k3, arg;
    bool near;
    pair_geom<float, AMOEBA>(ps[j], pa[j], ps[i], pa[i], dx, dy, dz, k2, k3, arg, near);
    pair_damp<float, AMOEBA>(arg, near, k2, k3);
    // P-P screening of the pair (list_P_P by target, f = 1 - s)
    float sij = 1.0f;
    for (int q = pp_ri[j]; q < pp_ri[j + 1]; q++)
        if (pp_ci[q] == i) sij = (float)(1.0 - pp_f[q]);
    const float c = ((float)alpha[j] * (float)alpha[i]) * sij;
    const float ck3 = c * k3, ck2 = c * k2;
    float *b = nb_b + 6L * e;
    b[0] = ck3 * (dx * dx) - ck2;
    b[1] = ck3 * (dx * dy);
    b[2] = ck3 * (dx * dz);
    b[3] = ck3 * (dy * dy) - ck2;
    b[4] = ck3 * (dy * dz);
    b[5] = ck3 * (dz * dz) - ck2;
}

// Lane q of the four lanes of site i: sum_j B_ij r_j over every fourth
// neighbour (k = row start + q, + 4, ...), r_j = rold_j - a h_j evaluated
// exactly as the owning lane evaluates its own (fma(-a, h, rold)); the four
// partial sums are then added in a fixed order ((0 + 1) + (2 + 3)).
__device__ __forceinline__ void nb_quad_sum(int i, int q, long base, double a, const double *__restrict__ rold,
                                            const double *__restrict__ h, const int *__restrict__ row_ptr,
                                            const int *__restrict__ nb_ci, const float *__restrict__ nb_b,
                                            double z[3]) {
    z[0] = z[1] = z[2] = 0.0;
    const int k1 = row_ptr[i + 1];
    for (int k = row_ptr[i] + q; k < k1; k += 4) {
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
#pragma unroll
    for (int c = 0; c < 3; c++) {
        z[c] += __shfl_xor(z[c], 1, 64);
        z[c] += __shfl_xor(z[c], 2, 64);
    }
}

// Four lanes per site; per-block partial sums have nb_blocks(npol) entries per column.
constexpr int kNbLanes = 4;
__host__ __device__ constexpr int nb_blocks_of(int npol) { return (kNbLanes * npol + kBlock - 1) / kBlock; }

// Inner solve start from a zero guess: x = 0, r = b, z = M^-1 b, p = z,
// per-block partials of r.z.
__global__ __launch_bounds__(kBlock) void k_cg_init_nb(int npol, const double *__restrict__ b, double *__restrict__ x,
                                                       double *__restrict__ r, double *__restrict__ z,
                                                       double *__restrict__ p, const double *__restrict__ alpha,
                                                       const int *__restrict__ row_ptr, const int *__restrict__ nb_ci,
                                                       const float *__restrict__ nb_b, const int *__restrict__ nb_off,
                                                       CgState *cs, double *__restrict__ red) {
    __shared__ double sh[4];
    const int m = blockIdx.y;
    const long n = 3L * npol, base = m * n;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t >> 2, q = t & 3;
    const int off = *nb_off;
    // the choice holds for the whole inner solve (k_cg_init2 leaves pad alone)
    if (blockIdx.x == 0 && threadIdx.x == 0) cs[m].pad = off;
    double zz[3] = {0.0, 0.0, 0.0};
    if (i < npol && !off) nb_quad_sum(i, q, base, 0.0, b, nullptr, row_ptr, nb_ci, nb_b, zz);
    double v = 0.0;
    if (i < npol && q == 0) {
        const long o = base + 3L * i;
        const double al = alpha[i];
        const double ci = off ? al : kNbDiag * al;
        for (int c = 0; c < 3; c++) {
            const double rn = b[o + c];
            const double zc = ci * rn + zz[c];
            x[o + c] = 0.0;
            r[o + c] = rn;
            z[o + c] = zc;
            p[o + c] = zc;
            v += rn * zc;
        }
    }
    v = block_sum_all(v, sh);
    if (threadIdx.x == 0) red[m * gridDim.x + blockIdx.x] = v;
}

// conjugate_gradient_solver loop body, first half, with the near-field
// preconditioner: x += a p, r_new = r_old - a h, z = M^-1 r_new; partials of
// r.z (for beta) and r.(alpha r) (the reference's convergence measure).  The
// loads that do not depend on a (the column's state, the site's own vectors,
// its first neighbours' h, r_old and B) are issued before the reduction of p.Tp,
// so that their latency overlaps it; the arithmetic and its order are those of
// nb_quad_sum.
constexpr int kNbPre = 3;  // neighbour entries per lane loaded ahead of the reduction
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
    __shared__ int s_done;
    const int m = blockIdx.y;
    // done may be set by block 0 of this launch (small-step exit): one thread
    // reads it for the whole workgroup; pad is fixed for the inner solve
    if (threadIdx.x == 0) s_done = cs[m].done;
    const int off = cs[m].pad;
    const long n = 3L * npol, base = m * n;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t >> 2, q = t & 3;
    const bool site = i < npol, nbr = site && !off;
    const int k0 = nbr ? row_ptr[i] + q : 0, k1 = nbr ? row_ptr[i + 1] : 0;
    double hj[kNbPre][3], rj[kNbPre][3];
    float bj[kNbPre][6];
#pragma unroll
    for (int u = 0; u < kNbPre; u++) {
        const int k = k0 + 4 * u;
        if (k < k1) {
            const long o = base + 3L * nb_ci[k];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hj[u][c] = h[o + c];
                rj[u][c] = rold[o + c];
            }
#pragma unroll
            for (int e = 0; e < 6; e++) bj[u][e] = nb_b[6L * k + e];
        }
    }
    const long o = base + 3L * i;
    double xo[3] = {0.0, 0.0, 0.0}, po[3] = {0.0, 0.0, 0.0}, ho[3] = {0.0, 0.0, 0.0}, ro[3] = {0.0, 0.0, 0.0};
    double al = 0.0;
    if (site && q == 0) {
        al = alpha[i];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            xo[c] = x[o + c];
            po[c] = p[o + c];
            ho[c] = h[o + c];
            ro[c] = rold[o + c];
        }
    }
    __syncthreads();
    if (s_done) return;  // uniform over the workgroup
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
    double zz[3] = {0.0, 0.0, 0.0};
    if (nbr) {
        int k = k0;
#pragma unroll
        for (int u = 0; u < kNbPre; u++, k += 4) {
            if (k >= k1) break;
            const double r0 = fma(-a, hj[u][0], rj[u][0]);
            const double r1 = fma(-a, hj[u][1], rj[u][1]);
            const double r2 = fma(-a, hj[u][2], rj[u][2]);
            const double bxx = bj[u][0], bxy = bj[u][1], bxz = bj[u][2], byy = bj[u][3], byz = bj[u][4],
                         bzz = bj[u][5];
            zz[0] += bxx * r0 + bxy * r1 + bxz * r2;
            zz[1] += bxy * r0 + byy * r1 + byz * r2;
            zz[2] += bxz * r0 + byz * r1 + bzz * r2;
        }
        for (; k < k1; k += 4) {
            const long oj = base + 3L * nb_ci[k];
            const double r0 = fma(-a, h[oj], rold[oj]);
            const double r1 = fma(-a, h[oj + 1], rold[oj + 1]);
            const double r2 = fma(-a, h[oj + 2], rold[oj + 2]);
            const float *b = nb_b + 6L * k;
            const double bxx = b[0], bxy = b[1], bxz = b[2], byy = b[3], byz = b[4], bzz = b[5];
            zz[0] += bxx * r0 + bxy * r1 + bxz * r2;
            zz[1] += bxy * r0 + byy * r1 + byz * r2;
            zz[2] += bxz * r0 + byz * r1 + bzz * r2;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            zz[c] += __shfl_xor(zz[c], 1, 64);
            zz[c] += __shfl_xor(zz[c], 2, 64);
        }
    }
    double vz = 0.0, va = 0.0;
    if (site && q == 0) {
        const double ci = off ? al : kNbDiag * al;
        for (int c = 0; c < 3; c++) {
            x[o + c] = xo[c] + a * po[c];
            const double rn = fma(-a, ho[c], ro[c]);
            const double zc = ci * rn + zz[c];
            rnew[o + c] = rn;
            z[o + c] = zc;
            vz += rn * zc;
            va += rn * (al * rn);
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
// so far is kept) and nb_off switches it off until the next geometry.  z and p
// are loaded before the reductions.
__global__ __launch_bounds__(kBlock) void k_cg_direction_nb(long n, int it, const double *__restrict__ z,
                                                            double *__restrict__ p, const double *__restrict__ red_rz,
                                                            const double *__restrict__ red_ra, int nred, CgState *cs,
                                                            int *__restrict__ nb_off) {
    __shared__ double sh[4];
    __shared__ int s_done;
    const int m = blockIdx.y;
    // done may be set by block 0 of this launch: one thread reads it
    if (threadIdx.x == 0) s_done = cs[m].done;
    const long k = (long)blockIdx.x * kBlock + threadIdx.x;
    double zk = 0.0, pk = 0.0;
    if (k < n) {
        zk = z[m * n + k];
        pk = p[m * n + k];
    }
    __syncthreads();
    if (s_done) return;  // uniform over the workgroup
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
    if (k < n) p[m * n + k] = beta * pk + zk;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cs[m].g[it & 1] = gnew;
        cs[m].rms = rms;
        cs[m].last_beta = beta;
        cs[m].iters = it;
    }
}

}  // namespace

int nb_blocks(int npol) { return nb_blocks_of(npol); }

void launch_nb_tensors(hipStream_t st, bool amoeba, const NbDev &nb, const float4 *ps32, const float *pa32,
                       const double *alpha, const CsrDev &pp) {
    if (nb.nnz == 0) return;
    const int blocks = (int)((nb.nnz + kBlock - 1) / kBlock);
    if (amoeba)
        hipLaunchKernelGGL(k_nb_tensors<true>, dim3(blocks), dim3(kBlock), 0, st, nb.nnz, nb.rowi, nb.ci, ps32, pa32,
                           alpha, pp.ri, pp.ci, pp.f, nb.b);
    else
        hipLaunchKernelGGL(k_nb_tensors<false>, dim3(blocks), dim3(kBlock), 0, st, nb.nnz, nb.rowi, nb.ci, ps32,
                           pa32, alpha, pp.ri, pp.ci, pp.f, nb.b);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_init_nb(hipStream_t st, int npol, int M, const double *b, double *x, double *r, double *z, double *p,
                       const double *alpha, const NbDev &nb, double *red, CgState *cs, const double *tol,
                       const CgState *pre) {
    const int nbk = nb_blocks_of(npol);
    hipLaunchKernelGGL(k_cg_init_nb, dim3(nbk, M), dim3(kBlock), 0, st, npol, b, x, r, z, p, alpha, nb.row_ptr,
                       nb.ci, nb.b, nb.off_flag, cs, red);
    HIP_CHECK(hipGetLastError());
    launch_cg_init2(st, M, red, nbk, cs, tol, pre);
}

void launch_cg_update_nb(hipStream_t st, int npol, int M, int it, const double *h, double *x, const double *rold,
                         double *rnew, double *z, const double *p, const double *alpha, const double *red_hp,
                         int nred_hp, double *red_rz, double *red_ra, CgState *cs, const NbDev &nb) {
    hipLaunchKernelGGL(k_cg_update_nb, dim3(nb_blocks_of(npol), M), dim3(kBlock), 0, st, npol, it, h, x, rold, rnew,
                       z, p, alpha, red_hp, nred_hp, red_rz, red_ra, cs, nb.row_ptr, nb.ci, nb.b);
    HIP_CHECK(hipGetLastError());
}

void launch_cg_direction_nb(hipStream_t st, int npol, int M, int it, const double *z, double *p,
                            const double *red_rz, const double *red_ra, CgState *cs, const NbDev &nb) {
    const long n = 3L * npol;
    hipLaunchKernelGGL(k_cg_direction_nb, dim3(vec_blocks(n), M), dim3(kBlock), 0, st, n, it, z, p, red_rz, red_ra,
                       nb_blocks_of(npol), cs, nb.off_flag);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
