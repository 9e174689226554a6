// Device-side arithmetic shared by the single-system and the batched kernels:
// block reductions with a fixed summation order, and the per-pair geometry,
// Thole damping and dipole-field accumulation of the T matvec
// (damped_coulomb_kernel + mu_elec_prop, mod_electrostatics.F90:1018-1284).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace ommp {

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in every thread.
__device__ __forceinline__ double block_sum_all(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return t;
}

// Fixed-order sum of np per-block partials, valid in every thread.
__device__ __forceinline__ double reduce_parts(const double *p, int np, double *sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += kBlock) v += p[i];
    return block_sum_all(v, sh);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}


template <typename R> struct V4;
template <> struct V4<double> { using t = double4; };
template <> struct V4<float> { using t = float4; };

__device__ __forceinline__ float fast_exp(float x) { return __expf(x); }
__device__ __forceinline__ double fast_exp(double x) { return exp(x); }
// r2 is either 0 (self pair, masked by the caller) or >= ~1 bohr^2, never
// denormal, so the raw v_rsq_f32 (1 ulp) needs no range fix-up.
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64: the hardware estimate y of x^-1/2 refined by one second-order step of
// the binomial series, x^-1/2 = y (1 - e)^-1/2 = y (1 + e/2 + 3 e^2/8 + ...)
// with e = 1 - x y^2 (the estimate's error cubed remains).  For a normal,
// positive, finite x only: the squared distance of a pair of distinct sites
// (coincident sites are rejected when the coordinates are set; a self pair's
// result is discarded by its caller), so the special-value handling of a
// general rsqrt is left out.
__device__ __forceinline__ double fast_rsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}
__device__ __forceinline__ float fmaa(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmaa(double a, double b, double c) { return fma(a, b, c); }

// Damped dipole-field kernel (maxder 2): returns k2 = lambda3 / r^3 and
// k3 = 3 lambda5 / r^5 so that E = k3 (x.d) d - k2 x and T x = -E.
// Damping only acts inside a few Angstrom; the exponential is evaluated under
// a wave-uniform branch that waves without a near pair skip entirely.
template <typename R, bool AMOEBA>
__device__ __forceinline__ void dip_kernel(R r2, R ir, R ti, R tj, R ai, R aj, R &k2, R &k3) {
    const R ir2 = ir * ir;
    k2 = ir * ir2;
    k3 = (R(3) * k2) * ir2;
    const R u = (r2 * ir) * (ti * tj);
    if (AMOEBA) {
        const R a = ai < aj ? ai : aj;
        const R au3 = a * (u * u * u);
        const bool near = au3 < R(50);
        if (__ballot(near)) {
            const R e = fast_exp(-au3);
            const R l3 = R(1) - e;
            const R l5 = R(1) - (R(1) + au3) * e;
            k2 = near ? k2 * l3 : k2;
            k3 = near ? k3 * l5 : k3;
        }
    } else {
        const bool near = u < R(1);
        if (__ballot(near)) {
            const R u3 = u * u * u, u4 = u3 * u;
            k2 = near ? k2 * (R(4) * u3 - R(3) * u4) : k2;
            k3 = near ? k3 * u4 : k3;
        }
    }
}

// Geometry of one pair: d = r_j - r_i, undamped k2 = 1/r^3 and k3 = 3/r^5,
// the damping argument (a u^3 for AMOEBA, u = r/s for Wang) and whether the
// pair needs damping at all.
// min of two positive finite values.  FP32: their bit patterns order as
// integers, so one integer min (no NaN quieting); FP64 keeps the float min
// (there is no 64-bit integer min instruction).
__device__ __forceinline__ float pos_min(float a, float b) {
    return __int_as_float(min(__float_as_int(a), __float_as_int(b)));
}
__device__ __forceinline__ double pos_min(double a, double b) { return fmin(a, b); }

template <typename R, bool AMOEBA, bool GUARD = true>
__device__ __forceinline__ void pair_geom(const typename V4<R>::t &pj, R aj, const typename V4<R>::t &pi, R ai,
                                          R &dx, R &dy, R &dz, R &k2, R &k3, R &arg, bool &near) {
    dx = pj.x - pi.x;
    dy = pj.y - pi.y;
    dz = pj.z - pi.z;
    const R r2 = fmaa(dz, dz, fmaa(dy, dy, dx * dx));
    // GUARD: a zero distance (the self pair) gives zero kernels; without it the
    // caller guarantees distinct positions
    const R ir = !GUARD || r2 > R(0) ? fast_rsqrt(r2) : R(0);
    const R ir2 = ir * ir;
    k2 = ir * ir2;
    k3 = (R(3) * k2) * ir2;
    const R u = (r2 * ir) * (pi.w * pj.w);
    if (AMOEBA) {
        const R a = pos_min(ai, aj);
        arg = a * (u * u * u);
        near = arg < R(50);
    } else {
        arg = u;
        near = u < R(1);
    }
}

template <typename R, bool AMOEBA>
__device__ __forceinline__ void pair_damp(R arg, bool near, R &k2, R &k3) {
    if (AMOEBA) {
        const R e = fast_exp(-arg);
        const R l3 = R(1) - e;
        const R l5 = R(1) - (R(1) + arg) * e;
        k2 = near ? k2 * l3 : k2;
        k3 = near ? k3 * l5 : k3;
    } else {
        const R u3 = arg * arg * arg, u4 = u3 * arg;
        k2 = near ? k2 * (R(4) * u3 - R(3) * u4) : k2;
        k3 = near ? k3 * u4 : k3;
    }
}

template <typename R, int M>
__device__ __forceinline__ void pair_accum(R (&acc)[M][3], const typename V4<R>::t *xv, R dx, R dy, R dz, R k2,
                                           R k3) {
#pragma unroll
    for (int m = 0; m < M; m++) {
        const R c = k3 * fmaa(xv[m].z, dz, fmaa(xv[m].y, dy, xv[m].x * dx));
        const R nk2 = -k2;
        acc[m][0] = fmaa(c, dx, fmaa(nk2, xv[m].x, acc[m][0]));
        acc[m][1] = fmaa(c, dy, fmaa(nk2, xv[m].y, acc[m][1]));
        acc[m][2] = fmaa(c, dz, fmaa(nk2, xv[m].z, acc[m][2]));
    }
}

// ----------------------------------------------------------- static field
// Field at polarizable target j of MM source i (charge, dipole, quadrupole),
// damped kernel with maxder 3 (AMOEBA) or 1 (Wang), as damped_coulomb_kernel +
// q/mu/quad_elec_prop (mod_electrostatics.F90:1018-1405).  d = r_j - r_i.
template <bool AMOEBA>
__device__ __forceinline__ void m2d_pair(double dx, double dy, double dz, double ir, double ti,
                                         double tj, double ai, double aj, const double *q,
                                         double &ex, double &ey, double &ez) {
    const double ir2 = ir * ir;
    double k2 = ir * ir2;
    const double r = (dx * dx + dy * dy + dz * dz) * ir;
    const double u = r * ti * tj;
    if (AMOEBA) {
        double k3 = k2 * ir2, k4 = k3 * ir2;
        const double a = fmin(ai, aj);
        const double au3 = a * u * u * u;
        const bool near = au3 < 50.0;
        if (__ballot(near)) {
            // au3 = +inf for a non-polarizable source: masked by the select
            const double e = exp(-au3);
            k2 = near ? k2 * (1.0 - e) : k2;
            k3 = near ? k3 * (1.0 - (1.0 + au3) * e) : k3;
            k4 = near ? k4 * (1.0 - (1.0 + au3 + 0.6 * au3 * au3) * e) : k4;
        }
        const double md = q[1] * dx + q[2] * dy + q[3] * dz;
        // quadrupole storage xx, xy, yy, xz, yz, zz
        const double qx = q[4] * dx + q[5] * dy + q[7] * dz;
        const double qy = q[5] * dx + q[6] * dy + q[8] * dz;
        const double qz = q[7] * dx + q[8] * dy + q[9] * dz;
        const double qrr = qx * dx + qy * dy + qz * dz;
        const double cd = q[0] * k2 + 3.0 * md * k3 + 15.0 * qrr * k4;
        ex = cd * dx - q[1] * k2 - 6.0 * qx * k3;
        ey = cd * dy - q[2] * k2 - 6.0 * qy * k3;
        ez = cd * dz - q[3] * k2 - 6.0 * qz * k3;
    } else {
        const bool near = u < 1.0;
        if (__ballot(near)) {
            const double u3 = u * u * u;
            k2 = near ? k2 * (4.0 * u3 - 3.0 * u3 * u) : k2;
        }
        ex = q[0] * k2 * dx;
        ey = q[0] * k2 * dy;
        ez = q[0] * k2 * dz;
    }
}


}  // namespace ommp
