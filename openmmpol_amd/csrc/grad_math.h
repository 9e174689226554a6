// Per-pair arithmetic of the polarization-energy gradient, shared by the direct
// sweeps (grad_kernels.hip) and the near field of the fast multipole gradient
// (fmm_kernels.hip).  Conventions: d = r_target - r_source, Egrd_ab = -dE_b/dr_a,
// EHes_abc = d2E_c/dr_a dr_b (mod_electrostatics.F90:1097-1405).
#pragma once

#include <hip/hip_runtime.h>

namespace ommp {

// Damped k2..k5 of one pair.  ir = 1/r, or 0 for the self pair (every k is then
// 0).  ti, tj = 1/thole (+inf when not polarizable: undamped), ai, aj = Thole
// a (AMOEBA, +inf when not polarizable so that min picks the polarizable end).
// Wang damping defines k2..k4 only (k5 is left undamped and never used).
template <bool AMOEBA>
__device__ __forceinline__ void damp_k5(double r2, double ir, double ti, double tj, double ai, double aj,
                                        double &k2, double &k3, double &k4, double &k5) {
    const double ir2 = ir * ir;
    k2 = ir * ir2;
    k3 = k2 * ir2;
    k4 = k3 * ir2;
    k5 = k4 * ir2;
    const double u = (r2 * ir) * (ti * tj);
    if (AMOEBA) {
        const double a = fmin(ai, aj);
        const double au3 = a * (u * u * u);
        const bool near = au3 < 50.0;
        if (__ballot(near)) {
            const double e = exp(-au3);
            const double x2 = au3 * au3;
            k2 = near ? k2 * (1.0 - e) : k2;
            k3 = near ? k3 * (1.0 - (1.0 + au3) * e) : k3;
            k4 = near ? k4 * (1.0 - (1.0 + au3 + 0.6 * x2) * e) : k4;
            k5 = near ? k5 * (1.0 - (1.0 + au3 + (18.0 / 35.0) * x2 + (9.0 / 35.0) * x2 * au3) * e) : k5;
        }
    } else {
        const bool near = u < 1.0;
        if (__ballot(near)) {
            const double u3 = u * u * u, u4 = u3 * u;
            k2 = near ? k2 * (4.0 * u3 - 3.0 * u4) : k2;
            k3 = near ? k3 * u4 : k3;
            k4 = near ? k4 * (0.2 * u4) : k4;
        }
    }
}

// g += sum_b w_b Egrd_ab of a static multipole q (charge; AMOEBA: + dipole q[1..3],
// quadrupole / 3 in q[4..9] as xx, xy, yy, xz, yz, zz).
// static_grad_pair with w.d given (wd).
template <bool AMOEBA>
__device__ __forceinline__ void static_grad_pair_wd(double dx, double dy, double dz, double k2, double k3, double k4,
                                                    double k5, const double *q, double wx, double wy, double wz,
                                                    double wd, double &gx, double &gy, double &gz) {
    if (AMOEBA) {
        const double mx = q[1], my = q[2], mz = q[3];
        const double md = mx * dx + my * dy + mz * dz;
        const double mw = mx * wx + my * wy + mz * wz;
        const double qx = q[4] * dx + q[5] * dy + q[7] * dz;
        const double qy = q[5] * dx + q[6] * dy + q[8] * dz;
        const double qz = q[7] * dx + q[8] * dy + q[9] * dz;
        const double qrr = qx * dx + qy * dy + qz * dz;
        const double wqd = wx * qx + wy * qy + wz * qz;
        const double Qwx = q[4] * wx + q[5] * wy + q[7] * wz;
        const double Qwy = q[5] * wx + q[6] * wy + q[8] * wz;
        const double Qwz = q[7] * wx + q[8] * wy + q[9] * wz;
        const double cd = 3.0 * q[0] * k3 * wd + 15.0 * k4 * md * wd - 3.0 * k3 * mw + 105.0 * k5 * qrr * wd -
                          30.0 * k4 * wqd;
        const double cw = -q[0] * k2 - 3.0 * k3 * md - 15.0 * k4 * qrr;
        const double cm = -3.0 * k3 * wd;
        const double cq = -30.0 * k4 * wd;
        const double c6 = 6.0 * k3;
        gx += cd * dx + cw * wx + cm * mx + cq * qx + c6 * Qwx;
        gy += cd * dy + cw * wy + cm * my + cq * qy + c6 * Qwy;
        gz += cd * dz + cw * wz + cm * mz + cq * qz + c6 * Qwz;
    } else {
        const double c = 3.0 * q[0] * k3 * wd, cw = -q[0] * k2;
        gx += c * dx + cw * wx;
        gy += c * dy + cw * wy;
        gz += c * dz + cw * wz;
    }
}

template <bool AMOEBA>
__device__ __forceinline__ void static_grad_pair(double dx, double dy, double dz, double k2, double k3, double k4,
                                                 double k5, const double *q, double wx, double wy, double wz,
                                                 double &gx, double &gy, double &gz) {
    static_grad_pair_wd<AMOEBA>(dx, dy, dz, k2, k3, k4, k5, q, wx, wy, wz, wx * dx + wy * dy + wz * dz, gx, gy, gz);
}

// g += f * sum_b w_b Egrd_ab(mu) = f * [15 k4 (mu.d)(w.d) d - 3 k3 (mu (w.d) + w (mu.d) + (mu.w) d)]
__device__ __forceinline__ void dip_grad_pair(double dx, double dy, double dz, double k3, double k4, double wx,
                                              double wy, double wz, double mx, double my, double mz, double f,
                                              double &gx, double &gy, double &gz) {
    const double wd = wx * dx + wy * dy + wz * dz;
    const double md = mx * dx + my * dy + mz * dz;
    const double mw = mx * wx + my * wy + mz * wz;
    const double cd = f * (15.0 * k4 * md * wd - 3.0 * k3 * mw);
    const double cw = -3.0 * f * k3 * md, cm = -3.0 * f * k3 * wd;
    gx += cd * dx + cw * wx + cm * mx;
    gy += cd * dy + cw * wy + cm * my;
    gz += cd * dz + cw * wz + cm * mz;
}

// Field and (AMOEBA) field gradient at the target of a dipole nu,
// e = r_target - r_source:
//   E    = 3 k3 (nu.e) e - k2 nu
//   Egrd = 15 k4 (nu.e) e e - 3 k3 (nu e + e nu + (nu.e) I)
// acc: E (0-2), Egrd xx xy yy xz yz zz (3-8).  Returns nu.e.
template <bool AMOEBA>
__device__ __forceinline__ double dip_field_at(double ex, double ey, double ez, double k2, double k3, double k4,
                                               double nx, double ny, double nz, double *acc) {
    const double ne = nx * ex + ny * ey + nz * ez;
    const double c3 = 3.0 * k3 * ne;
    acc[0] += c3 * ex - k2 * nx;
    acc[1] += c3 * ey - k2 * ny;
    acc[2] += c3 * ez - k2 * nz;
    if (AMOEBA) {
        const double c4 = 15.0 * k4 * ne, t3 = 3.0 * k3;
        acc[3] += c4 * ex * ex - t3 * (2.0 * nx * ex + ne);
        acc[4] += c4 * ex * ey - t3 * (nx * ey + ny * ex);
        acc[5] += c4 * ey * ey - t3 * (2.0 * ny * ey + ne);
        acc[6] += c4 * ex * ez - t3 * (nx * ez + nz * ex);
        acc[7] += c4 * ey * ez - t3 * (ny * ez + nz * ey);
        acc[8] += c4 * ez * ez - t3 * (2.0 * nz * ez + ne);
    }
    return ne;
}

// Field, field gradient and Q : EHes at the target of a dipole nu, e = r_target - r_source:
//   E, Egrd as dip_field_at
//   Q:EHes = 105 k5 (nu.e)(eQe) e - 15 k4 [nu (eQe) + 2 e (nu.Qe) + 2 (nu.e) Qe + (nu.e) tr(Q) e]
//            + 3 k3 [nu tr(Q) + 2 Q nu]
// acc: E (0-2), Egrd xx xy yy xz yz zz (3-8), Q:EHes (9-11).
template <bool AMOEBA>
__device__ __forceinline__ void dip_at_mm(double ex, double ey, double ez, double k2, double k3, double k4, double k5,
                                          double nx, double ny, double nz, const double *Q, double trQ,
                                          double *acc) {
    const double ne = dip_field_at<AMOEBA>(ex, ey, ez, k2, k3, k4, nx, ny, nz, acc);
    if (AMOEBA) {
        // Q = xx, xy, yy, xz, yz, zz
        const double qx = Q[0] * ex + Q[1] * ey + Q[3] * ez;
        const double qy = Q[1] * ex + Q[2] * ey + Q[4] * ez;
        const double qz = Q[3] * ex + Q[4] * ey + Q[5] * ez;
        const double eqe = qx * ex + qy * ey + qz * ez;
        const double nqe = nx * qx + ny * qy + nz * qz;
        const double qnx = Q[0] * nx + Q[1] * ny + Q[3] * nz;
        const double qny = Q[1] * nx + Q[2] * ny + Q[4] * nz;
        const double qnz = Q[3] * nx + Q[4] * ny + Q[5] * nz;
        const double ce = 105.0 * k5 * ne * eqe - 15.0 * k4 * (2.0 * nqe + ne * trQ);
        const double cn = -15.0 * k4 * eqe + 3.0 * k3 * trQ;
        const double cq = -30.0 * k4 * ne;
        const double cqn = 6.0 * k3;
        acc[9] += ce * ex + cn * nx + cq * qx + cqn * qnx;
        acc[10] += ce * ey + cn * ny + cq * qy + cqn * qny;
        acc[11] += ce * ez + cn * nz + cq * qz + cqn * qnz;
    }
}

}  // namespace ommp
