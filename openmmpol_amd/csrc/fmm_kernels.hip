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
#include "grad_math.h"

namespace ommp {


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
// add != 0 adds to the expansions already in out (two source sets in one tree).
template <int KIND>
__global__ __launch_bounds__(256) void k_fmm_p2m(int p, int nside, double h, double3 origin, int nleaf,
                                                const int *__restrict__ leaf_start,
                                                const double4 *__restrict__ pos, const double *__restrict__ src,
                                                int ld, int add, double2 *__restrict__ out) {
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
    double2 vp = ipow(e, cconj(accp)), vm = ipow(e, accp);
    if (add) {
        vp = cadd(vp, o[hidx(n, m)]);
        if (m > 0) vm = cadd(vm, o[hidx(n, -m)]);
    }
    o[hidx(n, m)] = vp;
    if (m > 0) o[hidx(n, -m)] = vm;
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

// ------------------------------------------------------------ gradient
// Derivatives of a local expansion.  With Phi^_n^-m = (-1)^m conj(Phi_n^m) the
// harmonics obey d/dz Phi^_n^m = Phi^_{n-1}^m, (d/dx + i d/dy) Phi^_n^m =
// -Phi^_{n-1}^{m+1} and (d/dx - i d/dy) Phi^_n^m = Phi^_{n-1}^{m-1}, and
// S_n^m = (-1)^n i^-m Phi^_n^m for every m.  Writing V = Re sum c_n^m Phi^_n^m
// (c_n^m = (-1)^n i^-m L_n^m), d/dx shifts the coefficients to (Y - X) / 2,
// d/dy to i (X + Y) / 2 and d/dz to c_{n+1}^m, with X c_n^m = c_{n+1}^{m-1} and
// Y c_n^m = c_{n+1}^{m+1}.  Hence for the multi-index (A, B, C), ord = A + B + C:
//   d^ord V = Re sum_{j,k} (-1)^ord i^-A [sum_s w_s (-1)^s L_{j+ord}^{k+A+B-2s}] S_j^k,
// w_s = coefficient of X^s Y^(A+B-s) in (Y - X)^A (X + Y)^B / 2^(A+B).  Real
// expansions pair the terms of k and -k into 2 Re(...) over k > 0.
__host__ __device__ constexpr double fmm_binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}
__host__ __device__ constexpr double fmm_dweight(int a, int b, int s) {
    double r = 0.0;
    for (int i = 0; i <= a; i++) {
        const int l = s - i;
        if (l < 0 || l > b) continue;
        r += fmm_binom(a, i) * ((i & 1) ? -1.0 : 1.0) * fmm_binom(b, l);
    }
    for (int t = 0; t < a + b; t++) r *= 0.5;
    return (s & 1) ? -r : r;
}
template <int A, int B, int S>
__device__ __forceinline__ void fmm_dterm(const double2 *__restrict__ Lc, double2 &c) {
    constexpr double w = fmm_dweight(A, B, S);
    if constexpr (w != 0.0) {
        const double2 l = Lc[-2 * S];
        c.x = fma(w, l.x, c.x);
        c.y = fma(w, l.y, c.y);
    }
}
// acc += f * d^(A,B,C) of the (j, k) term, sv = S_j^k(v); L = the cell's expansion.
template <int A, int B, int C>
__device__ __forceinline__ void fmm_dacc(const double2 *__restrict__ L, int p, int j, int k, double2 sv, double f,
                                         double &acc) {
    constexpr int ord = A + B + C;
    if (j + ord > p) return;
    const int n = j + ord;
    const double2 *Lc = L + n * n + n + k + A + B;  // L_n^(k + A + B - 2s) = Lc[-2s]
    double2 c = make_double2(0.0, 0.0);
    fmm_dterm<A, B, 0>(Lc, c);
    if constexpr (A + B >= 1) fmm_dterm<A, B, 1>(Lc, c);
    if constexpr (A + B >= 2) fmm_dterm<A, B, 2>(Lc, c);
    if constexpr (A + B >= 3) fmm_dterm<A, B, 3>(Lc, c);
    const double zr = c.x * sv.x - c.y * sv.y, zi = c.x * sv.y + c.y * sv.x;
    // Re((-1)^ord i^-A z)
    double v = (A & 1) ? zi : zr;
    if (A >= 2) v = -v;
    if (ord & 1) v = -v;
    acc = fma(f, v, acc);
}

// One thread per target (sorted): derivatives of the far potential at the target.
// MODE 0 (polarizable sites): Hessian H, out[c][j] = (H w)_c with w = c0 ipd[j] +
//   c1 ipd[n3 + j] (j = order[t]; the second set only when c1 != 0), i.e.
//   sum_b w_b Egrd_cb;
// MODE 1 (MM atoms, AMOEBA): E = -grad V, Egrd = H and Q : EHes = -Q_bc d_abc V
//   (q = sorted multipoles, quadrupole at 4..9), out[c][m], c < 12;
// MODE 2 (MM atoms, AMBER): E only, c < 3.
template <int MODE>
__global__ __launch_bounds__(256) void k_fmm_l2p_grad(int p, int ntgt, int nside, double h, double3 origin,
                                                      const double4 *__restrict__ pos, const int *__restrict__ leaf_of,
                                                      const int *__restrict__ order, const double2 *__restrict__ loc,
                                                      const double *__restrict__ ipd, long n3, double c0, double c1,
                                                      const double *__restrict__ q, double *__restrict__ out,
                                                      long ostride) {
    constexpr int NV = MODE == 0 ? 6 : (MODE == 1 ? 19 : 3);
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntgt) return;
    const double4 ps = pos[t];
    const long cell = leaf_of[t];
    const int ix = (int)(cell % nside), iy = (int)((cell / nside) % nside), iz = (int)(cell / ((long)nside * nside));
    const double vx = ps.x - (origin.x + (ix + 0.5) * h), vy = ps.y - (origin.y + (iy + 0.5) * h),
                 vz = ps.z - (origin.z + (iz + 0.5) * h);
    const int ncoef = (p + 1) * (p + 1);
    const double2 *L = loc + cell * ncoef;
    double acc[NV];
#pragma unroll
    for (int a = 0; a < NV; a++) acc[a] = 0.0;
    const double r2 = vx * vx + vy * vy + vz * vz;
    const double2 w = make_double2(vx, vy);
    double2 pmm = make_double2(1.0, 0.0);
    for (int m = 0; m < p; m++) {
        if (m > 0) pmm = cscale(cmul(pmm, w), 1.0 / (2.0 * m));
        double2 p2 = make_double2(0.0, 0.0), p1 = pmm;
        const double f = m > 0 ? 2.0 : 1.0;
        const int e = (-m) & 3;
        for (int n = m; n < p; n++) {
            double2 cur;
            if (n == m) {
                cur = pmm;
            } else if (n == m + 1) {
                cur = cscale(pmm, vz);
            } else {
                const double a1 = 2.0 * n - 1.0, den = 1.0 / ((double)n * n - (double)m * m);
                cur = make_double2(den * (a1 * vz * p1.x - r2 * p2.x), den * (a1 * vz * p1.y - r2 * p2.y));
            }
            if (n > m) {
                p2 = p1;
                p1 = cur;
            }
            double2 sv = ipow(e, cur);
            if (n & 1) sv = make_double2(-sv.x, -sv.y);
            if (MODE == 0) {
                fmm_dacc<2, 0, 0>(L, p, n, m, sv, f, acc[0]);
                fmm_dacc<1, 1, 0>(L, p, n, m, sv, f, acc[1]);
                fmm_dacc<0, 2, 0>(L, p, n, m, sv, f, acc[2]);
                fmm_dacc<1, 0, 1>(L, p, n, m, sv, f, acc[3]);
                fmm_dacc<0, 1, 1>(L, p, n, m, sv, f, acc[4]);
                fmm_dacc<0, 0, 2>(L, p, n, m, sv, f, acc[5]);
            } else {
                fmm_dacc<1, 0, 0>(L, p, n, m, sv, f, acc[0]);
                fmm_dacc<0, 1, 0>(L, p, n, m, sv, f, acc[1]);
                fmm_dacc<0, 0, 1>(L, p, n, m, sv, f, acc[2]);
                if (MODE == 1) {
                    fmm_dacc<2, 0, 0>(L, p, n, m, sv, f, acc[3]);
                    fmm_dacc<1, 1, 0>(L, p, n, m, sv, f, acc[4]);
                    fmm_dacc<0, 2, 0>(L, p, n, m, sv, f, acc[5]);
                    fmm_dacc<1, 0, 1>(L, p, n, m, sv, f, acc[6]);
                    fmm_dacc<0, 1, 1>(L, p, n, m, sv, f, acc[7]);
                    fmm_dacc<0, 0, 2>(L, p, n, m, sv, f, acc[8]);
                    fmm_dacc<3, 0, 0>(L, p, n, m, sv, f, acc[9]);
                    fmm_dacc<2, 1, 0>(L, p, n, m, sv, f, acc[10]);
                    fmm_dacc<1, 2, 0>(L, p, n, m, sv, f, acc[11]);
                    fmm_dacc<0, 3, 0>(L, p, n, m, sv, f, acc[12]);
                    fmm_dacc<2, 0, 1>(L, p, n, m, sv, f, acc[13]);
                    fmm_dacc<1, 1, 1>(L, p, n, m, sv, f, acc[14]);
                    fmm_dacc<0, 2, 1>(L, p, n, m, sv, f, acc[15]);
                    fmm_dacc<1, 0, 2>(L, p, n, m, sv, f, acc[16]);
                    fmm_dacc<0, 1, 2>(L, p, n, m, sv, f, acc[17]);
                    fmm_dacc<0, 0, 3>(L, p, n, m, sv, f, acc[18]);
                }
            }
        }
    }
    const long j = order[t];
    if (MODE == 0) {
        // the second set is read only when it is weighted (AMBER has one set)
        double wx = c0 * ipd[3 * j], wy = c0 * ipd[3 * j + 1], wz = c0 * ipd[3 * j + 2];
        if (c1 != 0.0) {
            wx += c1 * ipd[n3 + 3 * j];
            wy += c1 * ipd[n3 + 3 * j + 1];
            wz += c1 * ipd[n3 + 3 * j + 2];
        }
        // acc: xx xy yy xz yz zz
        out[j] = acc[0] * wx + acc[1] * wy + acc[3] * wz;
        out[ostride + j] = acc[1] * wx + acc[2] * wy + acc[4] * wz;
        out[2 * ostride + j] = acc[3] * wx + acc[4] * wy + acc[5] * wz;
    } else {
        out[j] = -acc[0];
        out[ostride + j] = -acc[1];
        out[2 * ostride + j] = -acc[2];
        if (MODE == 1) {
#pragma unroll
            for (int c = 0; c < 6; c++) out[(3 + c) * ostride + j] = acc[3 + c];
            const double *Q = q + 10L * t + 4;  // xx xy yy xz yz zz
            // third derivatives: 9 xxx, 10 xxy, 11 xyy, 12 yyy, 13 xxz, 14 xyz, 15 yyz, 16 xzz, 17 yzz, 18 zzz
            const double *T = acc + 9;
            const double hx = Q[0] * T[0] + Q[2] * T[2] + Q[5] * T[7] + 2.0 * (Q[1] * T[1] + Q[3] * T[4] + Q[4] * T[5]);
            const double hy = Q[0] * T[1] + Q[2] * T[3] + Q[5] * T[8] + 2.0 * (Q[1] * T[2] + Q[3] * T[5] + Q[4] * T[6]);
            const double hz = Q[0] * T[4] + Q[2] * T[6] + Q[5] * T[9] + 2.0 * (Q[1] * T[5] + Q[3] * T[7] + Q[4] * T[8]);
            out[9 * ostride + j] = -hx;
            out[10 * ostride + j] = -hy;
            out[11 * ostride + j] = -hz;
        }
    }
}

// Near field of the gradient at the polarizable sites (sorted target t,
// blockIdx.y = neighbour leaf): static multipoles contracted with w = 1/2 (mu_D +
// mu_P) (AMOEBA) or mu, and the dipole pairs 1/2 [mu_D,j . Egrd(mu_P) + mu_P,j .
// Egrd(mu_D)] (mu_j . Egrd(mu)); xs = sorted [D or mu][npol][3] (+ [P][npol][3]).
// Output part[nb][3][npol] at the site's own index.
template <bool AMOEBA>
__global__ __launch_bounds__(256) void k_fmm_gnear_pol(int npol, int nside, const double4 *__restrict__ ppos,
                                                       const double *__restrict__ pa, const int *__restrict__ pleaf,
                                                       const int *__restrict__ porder,
                                                       const int *__restrict__ pmmself,
                                                       const int *__restrict__ pleaf_start,
                                                       const int *__restrict__ mleaf_start,
                                                       const double4 *__restrict__ mpos,
                                                       const double *__restrict__ ma, const double *__restrict__ mq,
                                                       const double *__restrict__ xs, double *__restrict__ part) {
    constexpr int LD = AMOEBA ? 10 : 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= npol) return;
    const int nb = blockIdx.y;
    const long n3 = 3L * npol;
    const long cell = pleaf[t];
    const int ix = (int)(cell % nside) + nb % 3 - 1, iy = (int)((cell / nside) % nside) + (nb / 3) % 3 - 1,
              iz = (int)(cell / ((long)nside * nside)) + nb / 9 - 1;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix < nside && iy < nside && iz < nside) {
        const long leaf = ((long)iz * nside + iy) * nside + ix;
        const double4 pj = ppos[t];
        const double aj = pa[t];
        const double ax = xs[3L * t], ay = xs[3L * t + 1], az = xs[3L * t + 2];
        const double bx = AMOEBA ? xs[n3 + 3L * t] : ax, by = AMOEBA ? xs[n3 + 3L * t + 1] : ay,
                     bz = AMOEBA ? xs[n3 + 3L * t + 2] : az;
        const double wx = AMOEBA ? 0.5 * (ax + bx) : ax, wy = AMOEBA ? 0.5 * (ay + by) : ay,
                     wz = AMOEBA ? 0.5 * (az + bz) : az;
        const int mself = pmmself[t];
        for (int s = mleaf_start[leaf]; s < mleaf_start[leaf + 1]; s++) {
            if (s == mself) continue;
            const double4 pi = mpos[s];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, ma[s], aj, k2, k3, k4, k5);
            double q[LD];
#pragma unroll
            for (int c = 0; c < LD; c++) q[c] = mq[(long)LD * s + c];
            static_grad_pair<AMOEBA>(dx, dy, dz, k2, k3, k4, k5, q, wx, wy, wz, gx, gy, gz);
        }
        const double f = AMOEBA ? 0.5 : 1.0;
        for (int s = pleaf_start[leaf]; s < pleaf_start[leaf + 1]; s++) {
            if (s == t) continue;
            const double4 pi = ppos[s];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pj.w, pa[s], aj, k2, k3, k4, k5);
            if (AMOEBA) {
                dip_grad_pair(dx, dy, dz, k3, k4, ax, ay, az, xs[n3 + 3L * s], xs[n3 + 3L * s + 1],
                              xs[n3 + 3L * s + 2], f, gx, gy, gz);
                dip_grad_pair(dx, dy, dz, k3, k4, bx, by, bz, xs[3L * s], xs[3L * s + 1], xs[3L * s + 2], f, gx, gy,
                              gz);
            } else {
                dip_grad_pair(dx, dy, dz, k3, k4, ax, ay, az, xs[3L * s], xs[3L * s + 1], xs[3L * s + 2], f, gx, gy,
                              gz);
            }
        }
    }
    const long j = porder[t];
    double *o = part + (long)nb * n3;
    o[j] = gx;
    o[npol + j] = gy;
    o[2L * npol + j] = gz;
}

// Near field at the MM atoms (sorted target t, blockIdx.y = neighbour leaf): E,
// Egrd and Q : EHes (AMOEBA) or E of the dipoles nu (sorted, nu[npol][3]) of the
// polarizable sites around; the atom's own site is skipped.  Output
// part[nb][na][nmm] at the atom's own index.
template <bool AMOEBA>
__global__ __launch_bounds__(256) void k_fmm_gnear_mm(int nmm, int nside, const double4 *__restrict__ mpos,
                                                      const double *__restrict__ ma, const double *__restrict__ mq,
                                                      const int *__restrict__ mleaf, const int *__restrict__ morder,
                                                      const int *__restrict__ pleaf_start,
                                                      const double4 *__restrict__ ppos, const double *__restrict__ pa,
                                                      const int *__restrict__ pmmself, const double *__restrict__ nu,
                                                      double *__restrict__ part) {
    constexpr int NA = AMOEBA ? 12 : 3;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nmm) return;
    const int nb = blockIdx.y;
    const long cell = mleaf[t];
    const int ix = (int)(cell % nside) + nb % 3 - 1, iy = (int)((cell / nside) % nside) + (nb / 3) % 3 - 1,
              iz = (int)(cell / ((long)nside * nside)) + nb / 9 - 1;
    double acc[NA];
#pragma unroll
    for (int c = 0; c < NA; c++) acc[c] = 0.0;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix < nside && iy < nside && iz < nside) {
        const long leaf = ((long)iz * nside + iy) * nside + ix;
        const double4 pm = mpos[t];
        const double am = ma[t];
        double Q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (AMOEBA)
            for (int c = 0; c < 6; c++) Q[c] = mq[10L * t + 4 + c];
        const double trQ = Q[0] + Q[2] + Q[5];
        for (int s = pleaf_start[leaf]; s < pleaf_start[leaf + 1]; s++) {
            if (pmmself[s] == t) continue;
            const double4 pi = ppos[s];
            const double ex = pm.x - pi.x, ey = pm.y - pi.y, ez = pm.z - pi.z;
            const double r2 = ex * ex + ey * ey + ez * ez;
            double k2, k3, k4, k5;
            damp_k5<AMOEBA>(r2, rsqrt(r2), pi.w, pm.w, pa[s], am, k2, k3, k4, k5);
            dip_at_mm<AMOEBA>(ex, ey, ez, k2, k3, k4, k5, nu[3L * s], nu[3L * s + 1], nu[3L * s + 2], Q, trQ, acc);
        }
    }
    const long m = morder[t];
    double *o = part + (long)nb * NA * nmm;
#pragma unroll
    for (int c = 0; c < NA; c++) o[(long)c * nmm + m] = acc[c];
}

// Sorted dipole columns of the gradient: xs[0] = mu_D (or mu), xs[1] = mu_P,
// xs[2] = 1/2 (mu_D + mu_P) (AMOEBA); ipd = [set][npol][3].
__global__ void k_fmm_gather_ipd(int n, bool amoeba, const int *__restrict__ order, const double *__restrict__ ipd,
                                 double *__restrict__ xs) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const long j = order[t], n3 = 3L * n;
    for (int k = 0; k < 3; k++) {
        const double d = ipd[3 * j + k];
        xs[3L * t + k] = d;
        if (amoeba) {
            const double pv = ipd[n3 + 3 * j + k];
            xs[n3 + 3L * t + k] = pv;
            xs[2 * n3 + 3L * t + k] = 0.5 * (d + pv);
        }
    }
}

// dst[leaf][c] = src[leaf][c] for the first n coefficients (orders up to that of
// dst); src has sstride coefficients per leaf.
__global__ void k_fmm_leafcopy(long total, int n, int sstride, const double2 *__restrict__ src,
                               double2 *__restrict__ dst) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const long leaf = i / n;
    const int c = (int)(i - leaf * n);
    dst[i] = src[leaf * sstride + c];
}

// ------------------------------------------------------------- driver
static int p2m_blocks(int p, int nleaf) {
    const int per = 256 / ((p + 1) * (p + 2) / 2);
    return (nleaf + per - 1) / per;
}

static size_t lds_m2l(int p) {
    return sizeof(double2) * m2l_groups(p) * ((p + 1) * (p + 1) + (2 * p + 1) * (2 * p + 1));
}

// Leaf moments of the static multipoles at order p (<= p_tab) from the cache.
static void leaf_static(hipStream_t st, const FmmDev &f, int p, double2 *leafm) {
    const int ns = 1 << f.levels;
    const long n = (long)(p + 1) * (p + 1), total = (long)ns * ns * ns * n;
    hipLaunchKernelGGL(k_fmm_leafcopy, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total, (int)n,
                       f.ncoef_tab, f.mstatic, leafm);
    HIP_CHECK(hipGetLastError());
}

// Local expansions at the leaves of one source set (already expanded into the
// leaf multipoles at level L): M2M, M2L, L2L; cells with no target (cnt_tgt)
// are skipped.  False when the tree has no far field (L < 2).
static bool far_locals(hipStream_t st, const FmmDev &f, int p, const int *cnt_src, const int *cnt_tgt) {
    const int L = f.levels;
    const int ncoef = (p + 1) * (p + 1);
    if (L < 2) return false;
    for (int l = L - 1; l >= 2; l--) {
        const int ns = 1 << l;
        hipLaunchKernelGGL(k_fmm_m2m, dim3(ns * ns * ns), dim3(256), sizeof(double2) * ncoef, st, p, ns,
                           cnt_src + f.cell_off[l + 1], f.mexp + f.cell_off[l + 1] * ncoef,
                           f.tab_m2m + (long)l * 8 * f.ncoef_tab, f.ncoef_tab, f.mexp + f.cell_off[l] * ncoef);
        HIP_CHECK(hipGetLastError());
    }
    for (int l = 2; l <= L; l++) {
        const int ns = 1 << l;
        hipLaunchKernelGGL(k_fmm_m2l, dim3(ns * ns * ns), dim3(256), lds_m2l(p), st, p, ns, cnt_src + f.cell_off[l],
                           cnt_tgt + f.cell_off[l], f.mexp + f.cell_off[l] * ncoef,
                           f.tab_m2l + (long)l * 343 * f.ncoef2_tab, f.ncoef2_tab, f.lexp + f.cell_off[l] * ncoef);
        HIP_CHECK(hipGetLastError());
    }
    for (int l = 2; l < L; l++) {
        const int ns = 1 << (l + 1);
        hipLaunchKernelGGL(k_fmm_l2l, dim3(ns * ns * ns), dim3(256), sizeof(double2) * ncoef, st, p, ns,
                           cnt_tgt + f.cell_off[l + 1], f.lexp + f.cell_off[l] * ncoef,
                           f.tab_l2l + (long)l * 8 * f.ncoef_tab, f.ncoef_tab, f.lexp + f.cell_off[l + 1] * ncoef);
        HIP_CHECK(hipGetLastError());
    }
    return true;
}

// Far field of one source set at the polarizable sites: local expansions, then
// L2P into field[npol][3].
static void far_field(hipStream_t st, const FmmDev &f, int p, const int *cnt_src, double *field) {
    const int L = f.levels;
    const int ncoef = (p + 1) * (p + 1);
    if (far_locals(st, f, p, cnt_src, f.cnt_pol)) {
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
                               f.origin, nleaf, f.pol_leaf_start, f.pol_pos, f.xs + c * n3, 3, 0,
                               f.mexp + f.cell_off[L] * ncoef);
            HIP_CHECK(hipGetLastError());
        }
        far_field(st, f, p, f.cnt_pol, part + ((long)kFmmNearSplits * M + c) * n3);
    }
}

void fmm_launch_static_fields(hipStream_t st, const FmmDev &f, bool amoeba, int ld, double *part) {
    const int L = f.levels, ns = 1 << L;
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
    if (L >= 2) leaf_static(st, f, p, f.mexp + f.cell_off[L] * ncoef);
    far_field(st, f, p, f.cnt_mm, part + (long)kFmmNearSplits * n3);
}

void fmm_launch_static_moments(hipStream_t st, const FmmDev &f, int ld) {
    const int L = f.levels, ns = 1 << L, nleaf = ns * ns * ns;
    if (L < 2) return;
    hipLaunchKernelGGL(k_fmm_p2m<1>, dim3(p2m_blocks(f.p_tab, nleaf)), dim3(256), 0, st, f.p_tab, ns, f.leaf_h,
                       f.origin, nleaf, f.mm_leaf_start, f.mm_pos, f.mm_q, ld, 0, f.mstatic);
    HIP_CHECK(hipGetLastError());
}

void fmm_launch_grad_fields(hipStream_t st, const FmmDev &f, bool amoeba, int ld, const double *ipd,
                            double *part_pol, double *part_mm) {
    const int L = f.levels, ns = 1 << L, nleaf = ns * ns * ns;
    const int npol = f.npol, nmm = f.nmm;
    const long n3 = 3L * npol;
    const int na = amoeba ? 12 : 3;
    if (ld != (amoeba ? 10 : 1)) {
        fatal("fast multipole gradient: unexpected multipole width " + std::to_string(ld));
    }
    hipLaunchKernelGGL(k_fmm_gather_ipd, dim3((npol + 255) / 256), dim3(256), 0, st, npol, amoeba, f.pol_order, ipd,
                       f.xs);
    HIP_CHECK(hipGetLastError());
    const double *nu = amoeba ? f.xs + 2 * n3 : f.xs;
    const dim3 gp((npol + 255) / 256, kFmmNearSplits), gm((nmm + 255) / 256, kFmmNearSplits);
    if (amoeba) {
        hipLaunchKernelGGL(k_fmm_gnear_pol<true>, gp, dim3(256), 0, st, npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_mmself, f.pol_leaf_start, f.mm_leaf_start, f.mm_pos, f.mm_a, f.mm_q,
                           f.xs, part_pol);
        hipLaunchKernelGGL(k_fmm_gnear_mm<true>, gm, dim3(256), 0, st, nmm, ns, f.mm_pos, f.mm_a, f.mm_q, f.mm_leaf,
                           f.mm_order, f.pol_leaf_start, f.pol_pos, f.pol_a, f.pol_mmself, nu, part_mm);
    } else {
        hipLaunchKernelGGL(k_fmm_gnear_pol<false>, gp, dim3(256), 0, st, npol, ns, f.pol_pos, f.pol_a, f.pol_leaf,
                           f.pol_order, f.pol_mmself, f.pol_leaf_start, f.mm_leaf_start, f.mm_pos, f.mm_a, f.mm_q,
                           f.xs, part_pol);
        hipLaunchKernelGGL(k_fmm_gnear_mm<false>, gm, dim3(256), 0, st, nmm, ns, f.mm_pos, f.mm_a, f.mm_q, f.mm_leaf,
                           f.mm_order, f.pol_leaf_start, f.pol_pos, f.pol_a, f.pol_mmself, nu, part_mm);
    }
    HIP_CHECK(hipGetLastError());
    const int npass = amoeba ? 2 : 1;
    double *far_pol = part_pol + (long)kFmmNearSplits * n3;
    double *far_mm = part_mm + (long)kFmmNearSplits * na * nmm;
    if (L < 2) {
        HIP_CHECK(hipMemsetAsync(far_pol, 0, sizeof(double) * npass * n3, st));
        HIP_CHECK(hipMemsetAsync(far_mm, 0, sizeof(double) * na * nmm, st));
        return;
    }
    const int p = f.p_tab;
    const int ncoef = (p + 1) * (p + 1);
    double2 *leafm = f.mexp + f.cell_off[L] * ncoef;
    const double2 *leafl = f.lexp + f.cell_off[L] * ncoef;
    const dim3 lp((npol + 255) / 256), lm((nmm + 255) / 256);
    // polarizable targets: static multipoles + one dipole set per pass, contracted
    // with half of the other set (AMOEBA: P sources against D / 2, D against P / 2)
    for (int pass = 0; pass < npass; pass++) {
        leaf_static(st, f, p, leafm);
        const double *dip = amoeba ? f.xs + (pass == 0 ? n3 : 0) : f.xs;
        hipLaunchKernelGGL(k_fmm_p2m<0>, dim3(p2m_blocks(p, nleaf)), dim3(256), 0, st, p, ns, f.leaf_h, f.origin,
                           nleaf, f.pol_leaf_start, f.pol_pos, dip, 3, 1, leafm);
        HIP_CHECK(hipGetLastError());
        far_locals(st, f, p, f.cnt_mm, f.cnt_pol);
        const double c0 = amoeba ? (pass == 0 ? 0.5 : 0.0) : 1.0, c1 = amoeba ? (pass == 0 ? 0.0 : 0.5) : 0.0;
        hipLaunchKernelGGL(k_fmm_l2p_grad<0>, lp, dim3(256), 0, st, p, npol, ns, f.leaf_h, f.origin, f.pol_pos,
                           f.pol_leaf, f.pol_order, leafl, ipd, n3, c0, c1, f.mm_q, far_pol + pass * n3, (long)npol);
        HIP_CHECK(hipGetLastError());
    }
    // MM targets: the dipoles nu
    hipLaunchKernelGGL(k_fmm_p2m<0>, dim3(p2m_blocks(p, nleaf)), dim3(256), 0, st, p, ns, f.leaf_h, f.origin, nleaf,
                       f.pol_leaf_start, f.pol_pos, nu, 3, 0, leafm);
    HIP_CHECK(hipGetLastError());
    far_locals(st, f, p, f.cnt_pol, f.cnt_mm);
    if (amoeba)
        hipLaunchKernelGGL(k_fmm_l2p_grad<1>, lm, dim3(256), 0, st, p, nmm, ns, f.leaf_h, f.origin, f.mm_pos,
                           f.mm_leaf, f.mm_order, leafl, ipd, n3, 0.0, 0.0, f.mm_q, far_mm, (long)nmm);
    else
        hipLaunchKernelGGL(k_fmm_l2p_grad<2>, lm, dim3(256), 0, st, p, nmm, ns, f.leaf_h, f.origin, f.mm_pos,
                           f.mm_leaf, f.mm_order, leafl, ipd, n3, 0.0, 0.0, f.mm_q, far_mm, (long)nmm);
    HIP_CHECK(hipGetLastError());
}

}  // namespace ommp
