// Potential and field of the static multipoles and of the induced dipoles at
// arbitrary external points (potential_M2E / potential_D2E / field_M2E /
// field_D2E, mod_electrostatics.F90:3067-3336): undamped, unscreened, FP64.
// One thread per external point, sources staged through LDS in tiles of 256,
// the source range split over gridDim.y; the per-split partial results are
// added in a fixed order by the finalize kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

#include "device.h"

namespace ommp {

constexpr double kEps2 = kEpsRpDev * kEpsRpDev;

// Multipole sources: ms[i] = (x, y, z, .), mq[i][LD] = q, mu, quadrupole / 3
// (xx, xy, yy, xz, yz, zz).  V = q k1 + (mu.d) k2 + 3 (d.Q d) k3;
// E = q k2 d + 3 (mu.d) k3 d - k2 mu + 15 (d.Q d) k4 d - 6 k3 Q d, d = r_ext - r_i.
template <int LD>
__global__ __launch_bounds__(kBlock) void k_m2e(int n, const double *__restrict__ cext, int nsrc, int chunk,
                                                const double4 *__restrict__ ms, const double *__restrict__ mq,
                                                double4 *__restrict__ part, int *__restrict__ bad) {
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_q[LD][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < n;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (valid) {
        px = cext[3L * j];
        py = cext[3L * j + 1];
        pz = cext[3L * j + 2];
    }
    double v = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    int hit = 0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(nsrc, i0 + chunk);
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            s_pos[tid] = ms[base + tid];
#pragma unroll
            for (int c = 0; c < LD; c++) s_q[c][tid] = mq[(long)LD * (base + tid) + c];
        }
        __syncthreads();
        if (!valid) continue;
        for (int k = 0; k < cnt; k++) {
            const double dx = px - s_pos[k].x, dy = py - s_pos[k].y, dz = pz - s_pos[k].z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < kEps2) {
                hit = 1;
                continue;
            }
            const double k1 = rsqrt(r2), ir2 = k1 * k1, k2 = k1 * ir2;
            const double q = s_q[0][k];
            if (LD == 1) {
                v += q * k1;
                ex += q * k2 * dx;
                ey += q * k2 * dy;
                ez += q * k2 * dz;
            } else {
                const double k3 = k2 * ir2, k4 = k3 * ir2;
                const double mx = s_q[1][k], my = s_q[2][k], mz = s_q[3][k];
                const double md = mx * dx + my * dy + mz * dz;
                const double qx = s_q[4][k] * dx + s_q[5][k] * dy + s_q[7][k] * dz;
                const double qy = s_q[5][k] * dx + s_q[6][k] * dy + s_q[8][k] * dz;
                const double qz = s_q[7][k] * dx + s_q[8][k] * dy + s_q[9][k] * dz;
                const double qrr = qx * dx + qy * dy + qz * dz;
                v += q * k1 + md * k2 + 3.0 * qrr * k3;
                const double cd = q * k2 + 3.0 * md * k3 + 15.0 * qrr * k4;
                ex += cd * dx - mx * k2 - 6.0 * qx * k3;
                ey += cd * dy - my * k2 - 6.0 * qy * k3;
                ez += cd * dz - mz * k2 - 6.0 * qz * k3;
            }
        }
    }
    if (valid) part[(long)blockIdx.y * n + j] = make_double4(v, ex, ey, ez);
    if (hit) *bad = 1;
}

// Induced-dipole sources: ps[i] = position, potential from muV (the D set for
// AMOEBA), field from muE (1/2 (mu_D + mu_P) for AMOEBA, mod_electrostatics.F90:3250).
__global__ __launch_bounds__(kBlock) void k_d2e(int n, const double *__restrict__ cext, int nsrc, int chunk,
                                                const double4 *__restrict__ ps, const double *__restrict__ ipd,
                                                int amoeba, double4 *__restrict__ part, int *__restrict__ bad) {
    __shared__ double4 s_pos[kBlock];
    __shared__ double s_mv[3][kBlock], s_me[3][kBlock];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kBlock + tid;
    const bool valid = j < n;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (valid) {
        px = cext[3L * j];
        py = cext[3L * j + 1];
        pz = cext[3L * j + 2];
    }
    double v = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    int hit = 0;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(nsrc, i0 + chunk);
    const long np3 = 3L * nsrc;
    for (int base = i0; base < i1; base += kBlock) {
        const int cnt = min(kBlock, i1 - base);
        __syncthreads();
        if (tid < cnt) {
            const long i = base + tid;
            s_pos[tid] = ps[i];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double d = ipd[3 * i + c];
                s_mv[c][tid] = d;
                s_me[c][tid] = amoeba ? 0.5 * (ipd[np3 + 3 * i + c] + d) : d;
            }
        }
        __syncthreads();
        if (!valid) continue;
        for (int k = 0; k < cnt; k++) {
            const double dx = px - s_pos[k].x, dy = py - s_pos[k].y, dz = pz - s_pos[k].z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < kEps2) {
                hit = 1;
                continue;
            }
            const double k1 = rsqrt(r2), ir2 = k1 * k1, k2 = k1 * ir2, k3 = k2 * ir2;
            v += (s_mv[0][k] * dx + s_mv[1][k] * dy + s_mv[2][k] * dz) * k2;
            const double mx = s_me[0][k], my = s_me[1][k], mz = s_me[2][k];
            const double c3 = 3.0 * (mx * dx + my * dy + mz * dz) * k3;
            ex += c3 * dx - mx * k2;
            ey += c3 * dy - my * k2;
            ez += c3 * dz - mz * k2;
        }
    }
    if (valid) part[(long)blockIdx.y * n + j] = make_double4(v, ex, ey, ez);
    if (hit) *bad = 1;
}

__global__ __launch_bounds__(kBlock) void k_x2e_finalize(int n, int nparts, const double4 *__restrict__ part,
                                                         double4 *__restrict__ out) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    double4 a = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int s = 0; s < nparts; s++) {
        const double4 p = part[(long)s * n + j];
        a.x += p.x;
        a.y += p.y;
        a.z += p.z;
        a.w += p.w;
    }
    out[j] = a;
}

int x2e_splits(int n, int nsrc) {
    const int pb = (n + kBlock - 1) / kBlock;
    const int max_s = std::max(1, (nsrc + kBlock - 1) / kBlock);
    const int want = (2048 + pb - 1) / pb;
    return std::max(1, std::min(want, max_s));
}

void launch_x2e(hipStream_t st, int n, const double *cext, bool mm, bool pol, bool amoeba, int ld, int nmm,
                const double4 *ms, const double *mq, int npol, const double4 *ps, const double *ipd, double4 *part,
                int smm, int spol, double4 *out, int *bad) {
    if (n == 0) return;
    const int pb = (n + kBlock - 1) / kBlock;
    int off = 0;
    if (mm && nmm > 0) {
        const int chunk = (nmm + smm - 1) / smm;
        if (ld == 10)
            hipLaunchKernelGGL(k_m2e<10>, dim3(pb, smm), dim3(kBlock), 0, st, n, cext, nmm, chunk, ms, mq, part,
                               bad);
        else
            hipLaunchKernelGGL(k_m2e<1>, dim3(pb, smm), dim3(kBlock), 0, st, n, cext, nmm, chunk, ms, mq, part,
                               bad);
        HIP_CHECK(hipGetLastError());
        off = smm;
    }
    if (pol && npol > 0) {
        const int chunk = (npol + spol - 1) / spol;
        hipLaunchKernelGGL(k_d2e, dim3(pb, spol), dim3(kBlock), 0, st, n, cext, npol, chunk, ps, ipd,
                           amoeba ? 1 : 0, part + (long)off * n, bad);
        HIP_CHECK(hipGetLastError());
        off += spol;
    }
    hipLaunchKernelGGL(k_x2e_finalize, dim3(pb), dim3(kBlock), 0, st, n, off, part, out);
    HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------ no-allocation variant
// ommp_amd_field_mmpol2ext_noalloc keeps everything beyond its inputs and
// outputs in one file-scope device array per GPU, `scratch`: the external
// points and the caller's field (copied in), the per-split partial sums, the
// summed result and the coincident-point flag.  Split counts are those of the
// allocating entry (x2e_splits), so both give the same bits.  For n points
// each source kind takes at most ceil(2048 / ceil(n / 256)) splits, so the
// partial sums need at most 2 (2048 + ceil(n / 256) - 1) 256 double4.
// Every call on a GPU uses that GPU's array: calls that overlap on one GPU
// (other host threads, other streams) race; each call waits for its own work.
constexpr int kX2eNoAllocMaxPoints = 8192;
constexpr long kX2eNoAllocMaxParts = 2L * (2048 + (kX2eNoAllocMaxPoints + kBlock - 1) / kBlock - 1) * kBlock;
constexpr size_t kScrPart = 0;
constexpr size_t kScrOut = kScrPart + sizeof(double4) * kX2eNoAllocMaxParts;
constexpr size_t kScrC = kScrOut + sizeof(double4) * kX2eNoAllocMaxPoints;
constexpr size_t kScrE = kScrC + sizeof(double) * 3 * kX2eNoAllocMaxPoints;
constexpr size_t kScrBad = kScrE + sizeof(double) * 3 * kX2eNoAllocMaxPoints;
constexpr size_t kScrBytes = kScrBad + 256;
static_assert(kScrOut % 256 == 0 && kScrC % 256 == 0 && kScrE % 256 == 0 && kScrBad % 256 == 0,
              "scratch sub-buffers stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScrBytes];

int x2e_noalloc_max_points() { return kX2eNoAllocMaxPoints; }
size_t x2e_noalloc_scratch_bytes() { return kScrBytes; }

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (each GPU holds its own copy),
// looked up once per device with that device current.
unsigned char *scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fatal("Device index out of range for the no-allocation entry point");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) HIP_CHECK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) HIP_CHECK(hipSetDevice(cur));
        HIP_CHECK(e);
        if (!p) fatal("No device address for the scratch array of the no-allocation entry point");
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    return g_scratch_ptr[device];
}

__global__ void k_zero_flag(int *__restrict__ flag) {
    if (threadIdx.x == 0) *flag = 0;
}

// e[j] += field part of out[j] (out = V, Ex, Ey, Ez)
__global__ __launch_bounds__(kBlock) void k_add_x2e_field(int n, const double4 *__restrict__ out,
                                                          double *__restrict__ e) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double4 o = out[j];
    e[3L * j] += o.y;
    e[3L * j + 1] += o.z;
    e[3L * j + 2] += o.w;
}

}  // namespace

int x2e_field_noalloc(hipStream_t st, int device, int n, const double *cext, bool mm, bool pol, bool amoeba, int ld,
                      int nmm, const double4 *ms, const double *mq, int npol, const double4 *ps, const double *ipd,
                      double *e) {
    if (n <= 0) return 0;
    if (n > kX2eNoAllocMaxPoints)
        fatal("ommp_amd_field_mmpol2ext_noalloc takes at most " + std::to_string(kX2eNoAllocMaxPoints) +
              " external points per call (" + std::to_string(n) + " requested)");
    const int smm = mm ? x2e_splits(n, nmm) : 0, spol = pol ? x2e_splits(n, npol) : 0;
    if ((long)(smm + spol) * n > kX2eNoAllocMaxParts) fatal("Scratch of the no-allocation entry point too small");
    unsigned char *base = scratch_on(device);
    double4 *part = reinterpret_cast<double4 *>(base + kScrPart);
    double4 *out = reinterpret_cast<double4 *>(base + kScrOut);
    double *c = reinterpret_cast<double *>(base + kScrC);
    double *es = reinterpret_cast<double *>(base + kScrE);
    int *bad = reinterpret_cast<int *>(base + kScrBad);
    HIP_CHECK(hipMemcpyAsync(c, cext, sizeof(double) * 3L * n, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(es, e, sizeof(double) * 3L * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(64), 0, st, bad);
    HIP_CHECK(hipGetLastError());
    launch_x2e(st, n, c, mm, pol, amoeba, ld, nmm, ms, mq, npol, ps, ipd, part, smm, spol, out, bad);
    hipLaunchKernelGGL(k_add_x2e_field, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, out, es);
    HIP_CHECK(hipGetLastError());
    int bad_h = 0;
    HIP_CHECK(hipMemcpyAsync(e, es, sizeof(double) * 3L * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return bad_h;
}

}  // namespace ommp
