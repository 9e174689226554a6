// Fast multipole evaluation of the dipole field (T matvec) and of the static
// multipole field at the polarizable sites, FP64, for systems far larger than
// the direct O(N^2) sweep is worth (the role of the reference's
// src/fast_multipoles/, with its own operators: uniform octree, complex solid
// harmonics in scaled form, O(p^4) translations).
//
// Expansions (m from -n to n, index n^2 + n + m, complex as double2):
//   Phi_n^m(v)   = rho^n P_n^m(cos t) e^{i m phi} / (n + m)!       (m >= 0)
//   Psi_n^m(v)   = (n - m)! P_n^m(cos t) e^{i m phi} / rho^(n + 1)
// built from Cartesian recurrences (no trigonometry), so that the same code
// runs on dual numbers for source and target derivatives:
//   Phi_m^m = (x + iy)^m / (2^m m!),  Phi_{m+1}^m = z Phi_m^m,
//   Phi_n^m = ((2n - 1) z Phi_{n-1}^m - rho^2 Phi_{n-2}^m) / (n^2 - m^2);
//   Psi_m^m = (2m - 1)!! (x + iy)^m / rho^(2m + 1),  Psi_{m+1}^m = (2m + 1) z Psi_m^m / rho^2,
//   Psi_n^m = ((2n - 1) z Psi_{n-1}^m - (n + m - 1)(n - m - 1) Psi_{n-2}^m) / rho^2.
// Scaled harmonics (m >= 0; the -m entries from the conjugates):
//   R_n^m  = (-i)^m conj(Phi_n^m),   R_n^-m = (-i)^m Phi_n^m        (source moments)
//   S_n^m  = i^-m (-1)^n Phi_n^m,    S_n^-m = i^-m (-1)^n conj(Phi_n^m)  (local evaluation)
//   T_n^m  = i^m (-1)^n Psi_n^m,     T_n^-m = i^m (-1)^n conj(Psi_n^m)   (transfer)
// With these, the Greengard-Rokhlin translations become plain convolutions:
//   P2M  O_n^m   = sum_sources D_s R_n^m(s - c)   (D = q, mu . grad_s, Q : grad_s grad_s)
//   M2M  O'_j^k  = sum O_{j-n}^{k-m} R_n^m(c_child - c_parent)
//   M2L  L_j^k   = sum O_n^m T_{j+n}^{m-k}(c_source - c_target)
//   L2L  L'_j^k  = sum_{n >= j} L_n^m S_{n-j}^{m-k}(c_child - c_parent)
//   L2P  V(x)    = Re sum L_j^k S_j^k(x - c),  E = -grad V.
// Near field (the 27 leaves around a target's leaf, leaf side >= the near-field
// distance, so every damped or screened pair is there): the direct damped pair
// sums with scale 1, written as one partial; the far field (L2P) as a second;
// the existing finalize kernels add both and apply the sparse (1 - s) corrections.
// Every reduction has a fixed order: results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "device.h"
#include "fmm.h"

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

// ------------------------------------------------------- small arithmetic
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ void cmac(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// i^e for integer e (any sign)
__device__ __forceinline__ double2 ipow(int e, double2 a) {
    switch (e & 3) {
        case 0: return a;
        case 1: return make_double2(-a.y, a.x);
        case 2: return make_double2(-a.x, -a.y);
        default: return make_double2(a.y, -a.x);
    }
}

// Truncated Taylor numbers along one direction: value, first and second
// derivative (complex).  Only the operations the recurrences need.
struct Jet {
    double2 v, d1, d2;
};
__device__ __forceinline__ Jet jmul(const Jet &a, const Jet &b) {
    Jet r;
    r.v = cmul(a.v, b.v);
    r.d1 = cadd(cmul(a.d1, b.v), cmul(a.v, b.d1));
    r.d2 = cadd(cadd(cmul(a.d2, b.v), cmul(a.v, b.d2)), cscale(cmul(a.d1, b.d1), 2.0));
    return r;
}
__device__ __forceinline__ Jet jscale(const Jet &a, double s) {
    return {cscale(a.v, s), cscale(a.d1, s), cscale(a.d2, s)};
}
__device__ __forceinline__ Jet jsub(const Jet &a, const Jet &b) {
    return {make_double2(a.v.x - b.v.x, a.v.y - b.v.y), make_double2(a.d1.x - b.d1.x, a.d1.y - b.d1.y),
            make_double2(a.d2.x - b.d2.x, a.d2.y - b.d2.y)};
}
__device__ __forceinline__ Jet jreal(double v, double d1, double d2) {
    return {make_double2(v, 0.0), make_double2(d1, 0.0), make_double2(d2, 0.0)};
}

__device__ __forceinline__ int hidx(int n, int m) { return n * n + n + m; }

// --------------------------------------------------------------- P2M
// One thread per (leaf, n, m >= 0): the thread walks the Phi recurrence of its
// column m up to its order n for every source of the leaf (sorted order) and
// accumulates the +-m coefficients in registers; (p+1)(p+2)/2 threads per leaf,
// as many leaves per 256-thread workgroup as fit.
// KIND 0: dipoles xs[t][3] (sorted);  KIND 1: multipoles q[t][ld] (sorted MM atoms).
template <int KIND>
__global__ __launch_bounds__(256) void k_fmm_p2m(int p, int nside, double h, double3 origin, int nleaf,
                                                const int *__restrict__ leaf_start,
                                                const double4 *__restrict__ pos, const double *__restrict__ src,
                                                int ld, double2 *__restrict__ out) {
    const int npair = (p + 1) * (p + 2) / 2;
    const int per = 256 / npair;
    const int g = threadIdx.x / npair, r = threadIdx.x - g * npair;
    const int leaf = blockIdx.x * per + g;
    if (g >= per || leaf >= nleaf) return;
    // r -> (n, m), m <= n, ordered by n
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= r) n++;
    const int m = r - n * (n + 1) / 2;
    const int ix = leaf % nside, iy = (leaf / nside) % nside, iz = leaf / (nside * nside);
    const double cx = origin.x + (ix + 0.5) * h, cy = origin.y + (iy + 0.5) * h, cz = origin.z + (iz + 0.5) * h;
    const int s0 = leaf_start[leaf], s1 = leaf_start[leaf + 1];
    double2 accp = make_double2(0.0, 0.0);  // sum of D Phi_n^m over the sources
    for (int t = s0; t < s1; t++) {
        const double4 ps = pos[t];
        const double vx = ps.x - cx, vy = ps.y - cy, vz = ps.z - cz;
        const double r2v = vx * vx + vy * vy + vz * vz;
        const double2 w = make_double2(vx, vy);
        if (KIND == 0) {
            // first derivative along the dipole u: (value, derivative) pairs
            const double *d = src + 3L * t;
            const double ux = d[0], uy = d[1], uz = d[2];
            const double2 wd = make_double2(ux, uy);
            const double r2d = 2.0 * (vx * ux + vy * uy + vz * uz);
            double2 cv = make_double2(1.0, 0.0), cd = make_double2(0.0, 0.0);
            for (int k = 1; k <= m; k++) {
                const double sk = 1.0 / (2.0 * k);
                const double2 nv = cscale(cmul(cv, w), sk);
                cd = cscale(cadd(cmul(cd, w), cmul(cv, wd)), sk);
                cv = nv;
            }
            double2 pv = make_double2(0.0, 0.0), pd = make_double2(0.0, 0.0);
            for (int k = m + 1; k <= n; k++) {
                double2 nv, nd;
                if (k == m + 1) {
                    nv = cscale(cv, vz);
                    nd = cadd(cscale(cd, vz), cscale(cv, uz));
                } else {
                    const double a = 2.0 * k - 1.0, c = 1.0 / ((double)k * k - (double)m * m);
                    nv = make_double2(c * (a * vz * cv.x - r2v * pv.x), c * (a * vz * cv.y - r2v * pv.y));
                    nd = make_double2(c * (a * (uz * cv.x + vz * cd.x) - (r2d * pv.x + r2v * pd.x)),
                                      c * (a * (uz * cv.y + vz * cd.y) - (r2d * pv.y + r2v * pd.y)));
                }
                pv = cv;
                pd = cd;
                cv = nv;
                cd = nd;
            }
            accp = cadd(accp, cd);
        } else {
            const double *q = src + (long)ld * t;
            // direction 0: the charge (value) and the dipole (first derivative along
            // mu); directions 1-6: Q : H from second derivatives along x, y, z,
            // x+y, x+z, y+z
            const int nd = ld > 1 ? 7 : 1;
            for (int dd = 0; dd < nd; dd++) {
                double ux, uy, uz, wv, w1, w2;
                if (dd == 0) {
                    ux = ld > 1 ? q[1] : 0.0;
                    uy = ld > 1 ? q[2] : 0.0;
                    uz = ld > 1 ? q[3] : 0.0;
                    wv = q[0];
                    w1 = ld > 1 ? 1.0 : 0.0;
                    w2 = 0.0;
                } else {
                    const int k = dd - 1;
                    ux = (k == 0 || k == 3 || k == 4) ? 1.0 : 0.0;
                    uy = (k == 1 || k == 3 || k == 5) ? 1.0 : 0.0;
                    uz = (k == 2 || k == 4 || k == 5) ? 1.0 : 0.0;
                    wv = 0.0;
                    w1 = 0.0;
                    const double Qxx = q[4], Qxy = q[5], Qyy = q[6], Qxz = q[7], Qyz = q[8], Qzz = q[9];
                    w2 = k == 0 ? Qxx - Qxy - Qxz
                                : k == 1 ? Qyy - Qxy - Qyz
                                         : k == 2 ? Qzz - Qxz - Qyz : k == 3 ? Qxy : k == 4 ? Qxz : Qyz;
                    if (w2 == 0.0) continue;
                }
                const Jet wj = {w, make_double2(ux, uy), make_double2(0.0, 0.0)};
                const double r2d = 2.0 * (vx * ux + vy * uy + vz * uz), r2dd = 2.0 * (ux * ux + uy * uy + uz * uz);
                Jet cur = jreal(1.0, 0.0, 0.0);
                for (int k = 1; k <= m; k++) cur = jscale(jmul(cur, wj), 1.0 / (2.0 * k));
                Jet prev = jreal(0.0, 0.0, 0.0);
                for (int k = m + 1; k <= n; k++) {
                    // z = (vz, uz, 0), rho^2 = (r2v, r2d, r2dd): real jets
                    Jet zc;
                    zc.v = cscale(cur.v, vz);
                    zc.d1 = cadd(cscale(cur.d1, vz), cscale(cur.v, uz));
                    zc.d2 = cadd(cscale(cur.d2, vz), cscale(cur.d1, 2.0 * uz));
                    Jet next;
                    if (k == m + 1) {
                        next = zc;
                    } else {
                        Jet rp;
                        rp.v = cscale(prev.v, r2v);
                        rp.d1 = cadd(cscale(prev.d1, r2v), cscale(prev.v, r2d));
                        rp.d2 = cadd(cadd(cscale(prev.d2, r2v), cscale(prev.d1, 2.0 * r2d)), cscale(prev.v, r2dd));
                        next = jscale(jsub(jscale(zc, 2.0 * k - 1.0), rp), 1.0 / ((double)k * k - (double)m * m));
                    }
                    prev = cur;
                    cur = next;
                }
                accp = cadd(accp, cadd(cadd(cscale(cur.v, wv), cscale(cur.d1, w1)), cscale(cur.d2, w2)));
            }
        }
    }
    // R_n^m = (-i)^m conj(Phi), R_n^-m = (-i)^m Phi
    const int ncoef = (p + 1) * (p + 1);
    double2 *o = out + (long)leaf * ncoef;
    const int e = (-m) & 3;
    o[hidx(n, m)] = ipow(e, cconj(accp));
    if (m > 0) o[hidx(n, -m)] = ipow(e, accp);
}

// --------------------------------------------------------------- M2M
// Parent cells at level l (nside = 2^l) from their children at l + 1.  tab[oct][coef]
// = R(c_child - c_parent) for the 8 octants (oct = ox + 2 oy + 4 oz).
__global__ __launch_bounds__(256) void k_fmm_m2m(int p, int nside, const int *__restrict__ cnt_child,
                                                 const double2 *__restrict__ child, const double2 *__restrict__ tab,
                                                 int tstride, double2 *__restrict__ parent) {
    extern __shared__ double2 s_o[];
    const int cell = blockIdx.x;
    const int ncoef = (p + 1) * (p + 1);
    const int ix = cell % nside, iy = (cell / nside) % nside, iz = cell / (nside * nside);
    const int cs = 2 * nside;
    double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    for (int oct = 0; oct < 8; oct++) {
        const int cxi = 2 * ix + (oct & 1), cyi = 2 * iy + ((oct >> 1) & 1), czi = 2 * iz + (oct >> 2);
        const long ci = ((long)czi * cs + cyi) * cs + cxi;
        if (cnt_child[ci] == 0) continue;  // uniform across the block
        __syncthreads();
        for (int c = threadIdx.x; c < ncoef; c += 256) s_o[c] = child[ci * ncoef + c];
        __syncthreads();
        const double2 *R = tab + (long)oct * tstride;
        for (int r = 0; r < 2; r++) {
            const int out = threadIdx.x + 256 * r;
            if (out >= ncoef) break;
            int j = 0;
            while ((j + 1) * (j + 1) <= out) j++;
            const int k = out - j * j - j;
            double2 a = acc[r];
            for (int n = 0; n <= j; n++)
                for (int mm = -n; mm <= n; mm++) {
                    const int km = k - mm;
                    if (km < -(j - n) || km > (j - n)) continue;
                    cmac(a, s_o[hidx(j - n, km)], R[hidx(n, mm)]);
                }
            acc[r] = a;
        }
    }
    double2 *o = parent + (long)cell * ncoef;
    for (int r = 0; r < 2; r++) {
        const int out = threadIdx.x + 256 * r;
        if (out < ncoef) o[out] = acc[r];
    }
}

// --------------------------------------------------------------- M2L
// Target cells at level l: sources are the children of the parent's neighbours
// that are not neighbours of the target (offsets -3..3).  tab[(dz+3)*49+(dy+3)*7+dx+3][coef2]
// = T(c_source - c_target), coef2 over orders 0..2p.  Real sources give
// L_j^-k = (-1)^k conj(L_j^k) in this scaling, so only the (p+1)(p+2)/2
// outputs with k >= 0 are summed.  The source list is built in LDS first (fixed
// order); the workgroup runs G = 256 / nout groups (at most 8), each on every
// G-th source, and the group partials are added in group order at the end.
static __host__ __device__ inline int m2l_outputs(int p) { return (p + 1) * (p + 2) / 2; }
static __host__ __device__ inline int m2l_groups(int p) {
    const int g = 256 / m2l_outputs(p);
    return g < 1 ? 1 : (g > 8 ? 8 : g);
}

__global__ __launch_bounds__(256) void k_fmm_m2l(int p, int nside, const int *__restrict__ cnt_src,
                                                 const int *__restrict__ cnt_tgt,
                                                 const double2 *__restrict__ msrc, const double2 *__restrict__ tab,
                                                 int tstride, double2 *__restrict__ loc) {
    extern __shared__ double2 s_m2l[];
    __shared__ int s_flag[216];
    __shared__ int s_list[216];
    __shared__ int s_code[216];
    __shared__ int s_n;
    const int ncoef = (p + 1) * (p + 1);
    const int ncoef2 = (2 * p + 1) * (2 * p + 1);
    const int nout = m2l_outputs(p);
    const int G = m2l_groups(p);
    const int cell = blockIdx.x;
    const int tid = threadIdx.x;
    double2 *o = loc + (long)cell * ncoef;
    if (cnt_tgt[cell] == 0) {  // uniform across the block
        for (int c = tid; c < ncoef; c += 256) o[c] = make_double2(0.0, 0.0);
        return;
    }
    const int ix = cell % nside, iy = (cell / nside) % nside, iz = cell / (nside * nside);
    const int bx = 2 * (ix >> 1) - 2, by = 2 * (iy >> 1) - 2, bz = 2 * (iz >> 1) - 2;
    if (tid < 216) {
        const int sx = bx + tid % 6, sy = by + (tid / 6) % 6, sz = bz + tid / 36;
        const int dx = sx - ix, dy = sy - iy, dz = sz - iz;
        int ok = sx >= 0 && sy >= 0 && sz >= 0 && sx < nside && sy < nside && sz < nside;
        ok = ok && !(dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1);
        const long sc = ok ? ((long)sz * nside + sy) * nside + sx : 0;
        ok = ok && cnt_src[sc] != 0;
        s_flag[tid] = ok;
    }
    __syncthreads();
    if (tid == 0) {
        int n = 0;
        for (int c = 0; c < 216; c++)
            if (s_flag[c]) {
                const int sx = bx + c % 6, sy = by + (c / 6) % 6, sz = bz + c / 36;
                s_list[n] = (int)(((long)sz * nside + sy) * nside + sx);
                s_code[n] = (sz - iz + 3) * 49 + (sy - iy + 3) * 7 + (sx - ix + 3);
                n++;
            }
        s_n = n;
    }
    __syncthreads();
    const int nsrc = s_n;
    // thread -> (group, output r); output r -> (j, k >= 0); a thread may own
    // outputs r and r + 256 when one group covers more than 256 outputs
    const int gsize = G > 1 ? nout : 256;
    const int g = tid / gsize;
    const int r0 = tid - g * gsize;
    const bool active = g < G;
    double2 *s_o = s_m2l + (long)(active ? g : 0) * (ncoef + ncoef2);
    double2 *s_t = s_o + ncoef;
    double2 acc[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    int jj[2], kk[2];
    for (int q = 0; q < 2; q++) {
        const int rr = r0 + 256 * q;
        int j = 0;
        while ((j + 1) * (j + 2) / 2 <= rr) j++;
        jj[q] = j;
        kk[q] = rr - j * (j + 1) / 2;
    }
    const int nq = G > 1 ? 1 : 2;
    for (int it = 0; it < nsrc; it += G) {
        const int idx = it + g;
        const bool has = active && idx < nsrc;
        __syncthreads();
        if (has) {
            const double2 *O = msrc + (long)s_list[idx] * ncoef;
            const double2 *T = tab + (long)s_code[idx] * tstride;
            for (int c = r0; c < ncoef; c += gsize) s_o[c] = O[c];
            for (int c = r0; c < ncoef2; c += gsize) s_t[c] = T[c];
        }
        __syncthreads();
        if (has) {
            for (int q = 0; q < nq; q++) {
                if (r0 + 256 * q >= nout) break;
                const int j = jj[q], k = kk[q];
                double2 a = acc[q];
                for (int n = 0; n <= p; n++) {
                    const double2 *On = s_o + n * n + n;
                    const double2 *Tn = s_t + (j + n) * (j + n) + (j + n) - k;
                    for (int mm = -n; mm <= n; mm++) cmac(a, On[mm], Tn[mm]);
                }
                acc[q] = a;
            }
        }
    }
    if (G > 1) {
        __syncthreads();
        double2 *red = s_m2l;
        if (active) red[g * nout + r0] = acc[0];
        __syncthreads();
        if (g == 0) {
            double2 a = red[r0];
            for (int q = 1; q < G; q++) a = cadd(a, red[q * nout + r0]);
            acc[0] = a;
        }
    }
    for (int q = 0; q < nq; q++) {
        if (g != 0 || r0 + 256 * q >= nout) break;
        const int j = jj[q], k = kk[q];
        o[hidx(j, k)] = acc[q];
        if (k > 0) o[hidx(j, -k)] = (k & 1) ? make_double2(-acc[q].x, acc[q].y) : make_double2(acc[q].x, -acc[q].y);
    }
}

// --------------------------------------------------------------- L2L
// Child cells at level l + 1 (nside = 2^(l+1)) add their parent's local
// expansion; tab[oct][coef] = S(c_child - c_parent).
__global__ __launch_bounds__(256) void k_fmm_l2l(int p, int nside, const int *__restrict__ cnt_tgt,
                                                 const double2 *__restrict__ parent, const double2 *__restrict__ tab,
                                                 int tstride, double2 *__restrict__ child) {
    extern __shared__ double2 s_l[];
    const int cell = blockIdx.x;
    if (cnt_tgt[cell] == 0) return;  // uniform across the block
    const int ncoef = (p + 1) * (p + 1);
    const int ix = cell % nside, iy = (cell / nside) % nside, iz = cell / (nside * nside);
    const int ps = nside / 2;
    const long pc = ((long)(iz >> 1) * ps + (iy >> 1)) * ps + (ix >> 1);
    const int oct = (ix & 1) + 2 * (iy & 1) + 4 * (iz & 1);
    for (int c = threadIdx.x; c < ncoef; c += 256) s_l[c] = parent[pc * ncoef + c];
    __syncthreads();
    const double2 *S = tab + (long)oct * tstride;
    for (int out = threadIdx.x; out < ncoef; out += 256) {
        int j = 0;
        while ((j + 1) * (j + 1) <= out) j++;
        const int k = out - j * j - j;
        double2 a = child[(long)cell * ncoef + out];
        for (int n = j; n <= p; n++)
            for (int mm = -n; mm <= n; mm++) {
                const int mk = mm - k;
                if (mk < -(n - j) || mk > (n - j)) continue;
                cmac(a, s_l[hidx(n, mm)], S[hidx(n - j, mk)]);
            }
        child[(long)cell * ncoef + out] = a;
    }
}

// --------------------------------------------------------------- L2P
// One thread per polarizable site (sorted order): E = -grad Re sum L_j^k S_j^k(x - c)
// with first-order dual numbers in x, y and z; written to field[order[t]][3].
__global__ __launch_bounds__(256) void k_fmm_l2p(int p, int npol, int nside, double h, double3 origin,
                                                 const double4 *__restrict__ pos, const int *__restrict__ leaf_of,
                                                 const int *__restrict__ order,
                                                 const double2 *__restrict__ loc, double *__restrict__ field) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= npol) return;
    const double4 ps = pos[t];
    const long cell = leaf_of[t];
    const int ix = (int)(cell % nside), iy = (int)((cell / nside) % nside), iz = (int)(cell / ((long)nside * nside));
    const double vx = ps.x - (origin.x + (ix + 0.5) * h), vy = ps.y - (origin.y + (iy + 0.5) * h),
                 vz = ps.z - (origin.z + (iz + 0.5) * h);
    const int ncoef = (p + 1) * (p + 1);
    const double2 *L = loc + cell * ncoef;
    double g[3] = {0.0, 0.0, 0.0};
    // Phi with value and the three partial derivatives (complex each)
    const double r2 = vx * vx + vy * vy + vz * vz;
    double2 pmm[4] = {make_double2(1.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                      make_double2(0.0, 0.0)};
    const double2 w = make_double2(vx, vy);
    const double2 dw[3] = {make_double2(1.0, 0.0), make_double2(0.0, 1.0), make_double2(0.0, 0.0)};
    for (int m = 0; m <= p; m++) {
        if (m > 0) {
            // pmm *= w / (2m), with derivatives by the product rule
            const double s = 1.0 / (2.0 * m);
            double2 nv = cscale(cmul(pmm[0], w), s);
            double2 nd[3];
            for (int a = 0; a < 3; a++) nd[a] = cscale(cadd(cmul(pmm[1 + a], w), cmul(pmm[0], dw[a])), s);
            pmm[0] = nv;
            for (int a = 0; a < 3; a++) pmm[1 + a] = nd[a];
        }
        double2 p2[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                         make_double2(0.0, 0.0)};
        double2 p1[4] = {pmm[0], pmm[1], pmm[2], pmm[3]};
        for (int n = m; n <= p; n++) {
            double2 cur[4];
            if (n == m) {
                for (int a = 0; a < 4; a++) cur[a] = pmm[a];
            } else if (n == m + 1) {
                cur[0] = cscale(pmm[0], vz);
                cur[1] = cscale(pmm[1], vz);
                cur[2] = cscale(pmm[2], vz);
                cur[3] = cadd(cscale(pmm[3], vz), pmm[0]);
            } else {
                const double a1 = 2.0 * n - 1.0, den = 1.0 / ((double)n * n - (double)m * m);
                cur[0] = cscale(make_double2(a1 * vz * p1[0].x - r2 * p2[0].x, a1 * vz * p1[0].y - r2 * p2[0].y), den);
                const double dr2[3] = {2.0 * vx, 2.0 * vy, 2.0 * vz};
                for (int a = 0; a < 3; a++) {
                    double2 t1 = cscale(p1[1 + a], a1 * vz);
                    if (a == 2) t1 = cadd(t1, cscale(p1[0], a1));
                    double2 t2 = cadd(cscale(p2[1 + a], r2), cscale(p2[0], dr2[a]));
                    cur[1 + a] = cscale(make_double2(t1.x - t2.x, t1.y - t2.y), den);
                }
            }
            if (n > m) {
                for (int a = 0; a < 4; a++) {
                    p2[a] = p1[a];
                    p1[a] = cur[a];
                }
            }
            // S_n^m = i^-m (-1)^n Phi, S_n^-m = i^-m (-1)^n conj(Phi): Re(L S) summed over +-m
            const double sg = (n & 1) ? -1.0 : 1.0;
            const int e = (-m) & 3;
            const double2 lp = L[hidx(n, m)];
            for (int a = 0; a < 3; a++) {
                double2 sp = ipow(e, cscale(cur[1 + a], sg));
                double v = lp.x * sp.x - lp.y * sp.y;
                if (m > 0) {
                    const double2 lm = L[hidx(n, -m)];
                    double2 sm = ipow(e, cscale(cconj(cur[1 + a]), sg));
                    v += lm.x * sm.x - lm.y * sm.y;
                }
                g[a] += v;
            }
        }
    }
    double *o = field + 3L * order[t];
    o[0] = -g[0];
    o[1] = -g[1];
    o[2] = -g[2];
}

// ------------------------------------------------------------ near field
// One thread per polarizable target (sorted) and neighbour leaf (blockIdx.y =
// one of the 27 leaves around the target's leaf, kFmmNearSplits partials);
// damped kernels, scale 1 for every pair, self excluded.
// KIND 0: dipole sources (sorted polarizable sites, columns xs[col][t][3]) ->
// part[(split * ncol + col) * 3 npol + 3 order[t]];  KIND 1: MM multipole
// sources (sorted MM atoms) -> part[split * 3 npol + 3 order[t]].
template <int KIND, bool AMOEBA>
__global__ __launch_bounds__(256) void k_fmm_near(int ntgt, int nside, const double4 *__restrict__ tpos,
                                                  const double *__restrict__ ta, const int *__restrict__ tleaf,
                                                  const int *__restrict__ torder, const int *__restrict__ tself,
                                                  const int *__restrict__ leaf_start,
                                                  const double4 *__restrict__ spos, const double *__restrict__ sa,
                                                  const double *__restrict__ src, int ld, int ncol, double *__restrict__ part) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntgt) return;
    const int nb = blockIdx.y;
    const int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
    const long n3 = 3L * ntgt;
    const int ncl = KIND == 0 ? ncol : 1;
    const long cell = tleaf[t];
    const int ix = (int)(cell % nside) + dx, iy = (int)((cell / nside) % nside) + dy,
              iz = (int)(cell / ((long)nside * nside)) + dz;
    double acc[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix < nside && iy < nside && iz < nside) {
        const double4 pj = tpos[t];
        const double aj = ta[t];
        const int self = tself[t];
        const long leaf = ((long)iz * nside + iy) * nside + ix;
        for (int s = leaf_start[leaf]; s < leaf_start[leaf + 1]; s++) {
            if (s == self) continue;
            const double4 pi = spos[s];
            const double dxv = pj.x - pi.x, dyv = pj.y - pi.y, dzv = pj.z - pi.z;
            const double r2 = dxv * dxv + dyv * dyv + dzv * dzv;
            const double ir = rsqrt(r2), ir2 = ir * ir;
            double k2 = ir * ir2, k3 = k2 * ir2, k4 = k3 * ir2;
            const double u = (r2 * ir) * (pi.w * pj.w);
            if (AMOEBA) {
                const double a = fmin(sa[s], aj);
                const double au3 = a * u * u * u;
                if (au3 < 50.0) {
                    const double e = exp(-au3);
                    k2 *= 1.0 - e;
                    k3 *= 1.0 - (1.0 + au3) * e;
                    k4 *= 1.0 - (1.0 + au3 + 0.6 * au3 * au3) * e;
                }
            } else if (u < 1.0) {
                const double u3 = u * u * u, u4 = u3 * u;
                k2 *= 4.0 * u3 - 3.0 * u4;
                k3 *= u4;
            }
            if (KIND == 0) {
                for (int c = 0; c < ncl; c++) {
                    const double *x = src + c * n3 + 3L * s;
                    const double md = x[0] * dxv + x[1] * dyv + x[2] * dzv;
                    const double c3 = 3.0 * k3 * md;
                    acc[c][0] += c3 * dxv - k2 * x[0];
                    acc[c][1] += c3 * dyv - k2 * x[1];
                    acc[c][2] += c3 * dzv - k2 * x[2];
                }
            } else {
                const double *q = src + (long)ld * s;
                double cd = q[0] * k2;
                double ex = 0.0, ey = 0.0, ez = 0.0;
                if (ld > 1) {
                    const double md = q[1] * dxv + q[2] * dyv + q[3] * dzv;
                    const double qx = q[4] * dxv + q[5] * dyv + q[7] * dzv;
                    const double qy = q[5] * dxv + q[6] * dyv + q[8] * dzv;
                    const double qz = q[7] * dxv + q[8] * dyv + q[9] * dzv;
                    const double qrr = qx * dxv + qy * dyv + qz * dzv;
                    cd += 3.0 * md * k3 + 15.0 * qrr * k4;
                    ex = -q[1] * k2 - 6.0 * qx * k3;
                    ey = -q[2] * k2 - 6.0 * qy * k3;
                    ez = -q[3] * k2 - 6.0 * qz * k3;
                }
                acc[0][0] += cd * dxv + ex;
                acc[0][1] += cd * dyv + ey;
                acc[0][2] += cd * dzv + ez;
            }
        }
    }
    const long j = torder[t];
    for (int c = 0; c < ncl; c++) {
        double *o = part + ((long)nb * ncl + c) * n3 + 3L * j;
        o[0] = acc[c][0];
        o[1] = acc[c][1];
        o[2] = acc[c][2];
    }
}

// Sorted copy of the dipole columns: xs[c][t][3] = x[c][order[t]][3].
__global__ void k_fmm_gather(int n, int ncol, const int *__restrict__ order, const double *__restrict__ x,
                             long xstride, double *__restrict__ xs) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const long j = order[t];
    for (int c = 0; c < ncol; c++)
        for (int k = 0; k < 3; k++) xs[(long)c * 3 * n + 3L * t + k] = x[c * xstride + 3 * j + k];
}

// ------------------------------------------------------------- driver
static int p2m_blocks(int p, int nleaf) {
    const int per = 256 / ((p + 1) * (p + 2) / 2);
    return (nleaf + per - 1) / per;
}

static size_t lds_m2l(int p) {
    return sizeof(double2) * m2l_groups(p) * ((p + 1) * (p + 1) + (2 * p + 1) * (2 * p + 1));
}

// Far field of one source set (already expanded into the leaf multipoles at
// level L) at the polarizable sites: M2M, M2L, L2L, L2P into field[npol][3].
static void far_field(hipStream_t st, const FmmDev &f, int p, const int *cnt_src, double *field) {
    const int L = f.levels;
    const int ncoef = (p + 1) * (p + 1);
    if (L >= 2) {
        for (int l = L - 1; l >= 2; l--) {
            const int ns = 1 << l;
            hipLaunchKernelGGL(k_fmm_m2m, dim3(ns * ns * ns), dim3(256), sizeof(double2) * ncoef, st, p, ns,
                               cnt_src + f.cell_off[l + 1], f.mexp + f.cell_off[l + 1] * ncoef,
                               f.tab_m2m + (long)l * 8 * f.ncoef_tab, f.ncoef_tab, f.mexp + f.cell_off[l] * ncoef);
            HIP_CHECK(hipGetLastError());
        }
        for (int l = 2; l <= L; l++) {
            const int ns = 1 << l;
            hipLaunchKernelGGL(k_fmm_m2l, dim3(ns * ns * ns), dim3(256), lds_m2l(p), st, p, ns,
                               cnt_src + f.cell_off[l], f.cnt_pol + f.cell_off[l], f.mexp + f.cell_off[l] * ncoef,
                               f.tab_m2l + (long)l * 343 * f.ncoef2_tab, f.ncoef2_tab, f.lexp + f.cell_off[l] * ncoef);
            HIP_CHECK(hipGetLastError());
        }
        for (int l = 2; l < L; l++) {
            const int ns = 1 << (l + 1);
            hipLaunchKernelGGL(k_fmm_l2l, dim3(ns * ns * ns), dim3(256), sizeof(double2) * ncoef, st, p, ns,
                               f.cnt_pol + f.cell_off[l + 1], f.lexp + f.cell_off[l] * ncoef,
                               f.tab_l2l + (long)l * 8 * f.ncoef_tab, f.ncoef_tab, f.lexp + f.cell_off[l + 1] * ncoef);
            HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_fmm_l2p, dim3((f.npol + 255) / 256), dim3(256), 0, st, p, f.npol, 1 << L, f.leaf_h,
                           f.origin, f.pol_pos, f.pol_leaf, f.pol_order, f.lexp + f.cell_off[L] * ncoef, field);
        HIP_CHECK(hipGetLastError());
    } else {
        HIP_CHECK(hipMemsetAsync(field, 0, sizeof(double) * 3 * f.npol, st));
    }
}

void fmm_launch_matvec_fields(hipStream_t st, const FmmDev &f, bool amoeba, int M, const double *x, long xstride,
                              double *part) {
    const int L = f.levels, ns = 1 << L, nleaf = ns * ns * ns;
    const int p = f.p_pol;
    const int ncoef = (p + 1) * (p + 1);
    const long n3 = 3L * f.npol;
    hipLaunchKernelGGL(k_fmm_gather, dim3((f.npol + 255) / 256), dim3(256), 0, st, f.npol, M, f.pol_order, x,
                       xstride, f.xs);
    HIP_CHECK(hipGetLastError());
    // near field of all columns: splits 0 .. 26
    const dim3 ng((f.npol + 255) / 256, kFmmNearSplits);
    if (amoeba)
        hipLaunchKernelGGL((k_fmm_near<0, true>), ng, dim3(256), 0, st, f.npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_self, f.pol_leaf_start, f.pol_pos, f.pol_a, f.xs, 0, M, part);
    else
        hipLaunchKernelGGL((k_fmm_near<0, false>), ng, dim3(256), 0, st, f.npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_self, f.pol_leaf_start, f.pol_pos, f.pol_a, f.xs, 0, M, part);
    HIP_CHECK(hipGetLastError());
    // far field per column: split 27
    for (int c = 0; c < M; c++) {
        if (L >= 2) {
            hipLaunchKernelGGL(k_fmm_p2m<0>, dim3(p2m_blocks(p, nleaf)), dim3(256), 0, st, p, ns, f.leaf_h,
                               f.origin, nleaf, f.pol_leaf_start, f.pol_pos, f.xs + c * n3, 3,
                               f.mexp + f.cell_off[L] * ncoef);
            HIP_CHECK(hipGetLastError());
        }
        far_field(st, f, p, f.cnt_pol, part + ((long)kFmmNearSplits * M + c) * n3);
    }
}

void fmm_launch_static_fields(hipStream_t st, const FmmDev &f, bool amoeba, int ld, double *part) {
    const int L = f.levels, ns = 1 << L, nleaf = ns * ns * ns;
    const int p = f.p_static;
    const int ncoef = (p + 1) * (p + 1);
    const long n3 = 3L * f.npol;
    const dim3 ng((f.npol + 255) / 256, kFmmNearSplits);
    if (amoeba)
        hipLaunchKernelGGL((k_fmm_near<1, true>), ng, dim3(256), 0, st, f.npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_mmself, f.mm_leaf_start, f.mm_pos, f.mm_a, f.mm_q, ld, 1, part);
    else
        hipLaunchKernelGGL((k_fmm_near<1, false>), ng, dim3(256), 0, st, f.npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_mmself, f.mm_leaf_start, f.mm_pos, f.mm_a, f.mm_q, ld, 1, part);
    HIP_CHECK(hipGetLastError());
    if (L >= 2) {
        hipLaunchKernelGGL(k_fmm_p2m<1>, dim3(p2m_blocks(p, nleaf)), dim3(256), 0, st, p, ns, f.leaf_h, f.origin,
                           nleaf, f.mm_leaf_start, f.mm_pos, f.mm_q, ld, f.mexp + f.cell_off[L] * ncoef);
        HIP_CHECK(hipGetLastError());
    }
    far_field(st, f, p, f.cnt_mm, part + (long)kFmmNearSplits * n3);
}

}  // namespace ommp
