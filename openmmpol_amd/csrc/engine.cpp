// Host-side drivers of the device path: HBM residency of a system, the static
// field, the polarization() driver (mod_polarization.F90:49-313) and its three
// solvers, the polarization energy.  Every arithmetic step runs on the GPU;
// the host only sequences launches and reads back convergence flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "batch.h"
#include "batch_engine.h"
#include "device.h"
#include "fmm.h"
#include "fmm_host.h"
#include "ommp_internal.h"

namespace ommp {

void hip_fatal(hipError_t e, const char *what, const char *file, int line) {
    const char *base = std::strrchr(file, '/');
    fatal(std::string("HIP error: ") + hipGetErrorString(e) + " (" + what + ", " + (base ? base + 1 : file) + ":" +
          std::to_string(line) + ")");
}

#define HCHECK(x) HIP_CHECK(x)

// kOtf64 / kOtf32 pick the pair-once sweep (kSym*) or the per-target one
// (kPlain*) by size (tmv_pair_once); the explicit kinds force one of them.
enum MvKind { kOtf64 = 0, kOtf32 = 1, kPack32 = 2, kPack64 = 3, kFmm = 4, kSym64 = 5, kSym32 = 6, kPlain64 = 7,
              kPlain32 = 8 };

struct DeviceState {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int npol = 0, nmm = 0, ld = 1;
    long n = 0;
    // sites
    double4 *ps64 = nullptr;
    double *pa64 = nullptr, *alpha = nullptr;
    float4 *ps32 = nullptr;
    float *pa32 = nullptr;
    int *polar_mm = nullptr;
    double4 *ms64 = nullptr;
    double *ma64 = nullptr, *mq = nullptr;
    CsrDev sp, pp, pd;
    CsrDev ss;  // S-S pairs by target, f = 1 - s (fixed-multipole energy)
    // fields, dipoles, work vectors ([2][n] each)
    double *e_m2d = nullptr, *ipd = nullptr, *ext = nullptr;
    double *vb = nullptr, *vx = nullptr, *vh = nullptr, *vr = nullptr, *vz = nullptr, *vp = nullptr;
    double *vd = nullptr, *vr2 = nullptr;
    double *red_a = nullptr, *red_b = nullptr, *red_c = nullptr;
    double *scal = nullptr;
    double *tol_dev = nullptr;
    CgState *cs = nullptr;
    CgState *cs_host = nullptr;
    double *scal_host = nullptr;
    // mixed solve enqueued ahead: per-column refinement state, then the inner
    // CG state of each refinement step; pinned copy (allocated on first use)
    CgState *cs_spec = nullptr;
    CgState *cs_spec_host = nullptr;
    void *mvpart = nullptr;
    int2 *sym_pairs = nullptr;  // block pairs of the pair-once matvec, on first use
    void *sympart = nullptr;
    size_t mvpart_bytes = 0;
    int nsplit = 1;
    double *rhs_part = nullptr;
    int rhs_split = 1;
    // static field with the pair-once sweep: the non-polarizable MM sources and
    // the source split of their one-sided sweep
    int *np_src = nullptr;
    int rhs_split_np = 1;
    // packed T
    void *tpack = nullptr;
    int tpack_prec = 0;
    bool tpack_valid = false;
    void *tpart = nullptr;
    // dense T / LU (inversion)
    double *tdense = nullptr;
    int *piv = nullptr;
    bool lu_valid = false;
    double *lu_work = nullptr;
    // DIIS
    double *diis_x = nullptr, *diis_e = nullptr, *diis_coef = nullptr;
    // near-field preconditioner of the inner FP32 solves (on first use), the
    // second residual buffer and the r.(alpha r) partials its CG steps need
    NbDev nb;
    double *vr3 = nullptr, *red_d = nullptr;
    // fixed-multipole energy (on first use)
    double *mm_part = nullptr, *mm_red = nullptr, *mm_out = nullptr;
    int *mm_bad = nullptr;
    int mm_split = 1;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<void *> allocs;
    std::vector<size_t> alloc_bytes;
    // batched multi-RHS solver state (cached between calls)
    BatchEngine *batch = nullptr;
    // external-point potentials and fields (grown on demand)
    double *x2e_c = nullptr;
    double4 *x2e_part = nullptr, *x2e_out = nullptr;
    int *x2e_bad = nullptr;
    long x2e_cap = 0, x2e_cap_part = 0;
    // live kernel timing: event pairs in flight and accumulated totals per class
    bool prof = false;
    std::vector<hipEvent_t> prof_pool;
    std::vector<int> prof_class;
    int prof_used = 0;
    int prof_every = 1;        // bracket every prof_every-th launch of a class
    long prof_seq[6] = {0, 0, 0, 0, 0, 0};
    double prof_ms[6] = {0, 0, 0, 0, 0, 0};
    long prof_count[6] = {0, 0, 0, 0, 0, 0};
    // geometric gradient (built on first use): MM-target correction list, frames,
    // per-atom rotation index, partial sums and results
    CsrDev mp;
    int *mm_polar = nullptr;
    double *q0 = nullptr;
    int4 *frames = nullptr;
    int *rev_ri = nullptr, *rev_ci = nullptr;
    double *g_part_pol = nullptr, *g_part_mm = nullptr, *g_pol = nullptr, *g_tr = nullptr;
    double *g_emm = nullptr, *g_egrd = nullptr, *g_rotc = nullptr, *g_out = nullptr;
    int g_split_s = 1, g_split_p = 1, g_split_m = 1;
    bool grad_static_ready = false;
    // fast multipoles (built on first use, rebuilt after a geometry change)
    FmmDev fmm;
    bool fmm_valid = false;
    std::vector<void *> fmm_ptrs;
    double *fmm_part = nullptr, *fmm_rhs_part = nullptr;
    double *fmm_gpart_pol = nullptr, *fmm_gpart_mm = nullptr;  // gradient partials, on first use
};

namespace {

void *dmalloc(DeviceState &d, size_t bytes) {
    if (bytes == 0) bytes = 8;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        size_t fr = 0, tot = 0;
        (void)hipGetLastError();
        (void)hipMemGetInfo(&fr, &tot);
        char buf[256];
        std::snprintf(buf, sizeof(buf), "Cannot allocate %.3f GB of device memory (%.3f GB free of %.3f GB)",
                      bytes / 1e9, fr / 1e9, tot / 1e9);
        fatal(buf);
    }
    d.allocs.push_back(p);
    d.alloc_bytes.push_back(bytes);
    return p;
}

template <typename T> void dfree(DeviceState &d, T *&p) {
    if (!p) return;
    void *v = (void *)p;
    (void)hipFree(v);
    for (size_t i = 0; i < d.allocs.size(); i++)
        if (d.allocs[i] == v) {
            d.allocs.erase(d.allocs.begin() + i);
            d.alloc_bytes.erase(d.alloc_bytes.begin() + i);
            break;
        }
    p = nullptr;
}

template <typename T> T *dalloc(DeviceState &d, size_t count) { return (T *)dmalloc(d, count * sizeof(T)); }

constexpr int kProfPool = 512;

void prof_flush(DeviceState &d) {
    if (d.prof_used == 0) return;
    HCHECK(hipEventSynchronize(d.prof_pool[2 * (d.prof_used - 1) + 1]));
    for (int i = 0; i < d.prof_used; i++) {
        float ms = 0.0f;
        HCHECK(hipEventElapsedTime(&ms, d.prof_pool[2 * i], d.prof_pool[2 * i + 1]));
        d.prof_ms[d.prof_class[i]] += ms;
        d.prof_count[d.prof_class[i]] += 1;
    }
    d.prof_used = 0;
}

// Returns the index of an event pair to bracket one launch, or -1 when off.
int prof_begin(DeviceState &d, int cls) {
    if (!d.prof) return -1;
    if (d.prof_seq[cls]++ % d.prof_every != 0) return -1;
    if (d.prof_pool.empty()) {
        d.prof_pool.resize(2 * kProfPool);
        d.prof_class.resize(kProfPool);
        for (auto &e : d.prof_pool) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    if (d.prof_used == kProfPool) prof_flush(d);
    const int i = d.prof_used++;
    d.prof_class[i] = cls;
    HCHECK(hipEventRecord(d.prof_pool[2 * i], d.stream));
    return i;
}

void upload(const void *h, void *d, size_t bytes) {
    if (bytes) HCHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
}

// Per-target CSR with one or two factors per entry from reference-layout lists.
void upload_csr(DeviceState &d, int nrow, const std::vector<std::vector<std::pair<int, double2>>> &rows,
                CsrDev &out, bool two) {
    std::vector<int> ri(nrow + 1, 0), ci;
    std::vector<double> f;
    std::vector<double2> f2;
    for (int j = 0; j < nrow; j++) {
        ri[j + 1] = ri[j] + (int)rows[j].size();
        for (auto &e : rows[j]) {
            ci.push_back(e.first);
            if (two) f2.push_back(e.second);
            else f.push_back(e.second.x);
        }
    }
    out.nnz = ri[nrow];
    out.ri = dalloc<int>(d, nrow + 1);
    out.ci = dalloc<int>(d, std::max(out.nnz, 1));
    upload(ri.data(), out.ri, sizeof(int) * (nrow + 1));
    upload(ci.data(), out.ci, sizeof(int) * out.nnz);
    if (two) {
        out.f2 = dalloc<double2>(d, std::max(out.nnz, 1));
        upload(f2.data(), out.f2, sizeof(double2) * out.nnz);
    } else {
        out.f = dalloc<double>(d, std::max(out.nnz, 1));
        upload(f.data(), out.f, sizeof(double) * out.nnz);
    }
}

void free_csr(DeviceState &d, CsrDev &c) {
    dfree(d, c.ri);
    dfree(d, c.ci);
    dfree(d, c.f);
    dfree(d, c.f2);
    c = CsrDev();
}

// S->P corrections per polarizable target: MM source, f = 1 - s for the D and
// P sets (AMBER: one set, stored in both slots); the self pair is never summed.
std::vector<std::vector<std::pair<int, double2>>> sp_rows(const System &s) {
    const int npol = s.npol;
    std::vector<std::vector<std::pair<int, double2>>> rows(npol);
    auto add = [&](int i, int jp, double sd, double spv, bool setd, bool setp) {
        if (s.polar_mm[jp] == i) return;
        for (auto &e : rows[jp])
            if (e.first == i) {
                if (setd) e.second.x = 1.0 - sd;
                if (setp) e.second.y = 1.0 - spv;
                return;
            }
        double2 f;
        f.x = setd ? 1.0 - sd : 0.0;
        f.y = setp ? 1.0 - spv : 0.0;
        rows[jp].push_back({i, f});
    };
    for (int i = 0; i < s.nmm; i++)
        for (int k = s.list_S_P_P.ri[i]; k < s.list_S_P_P.ri[i + 1]; k++) {
            const double sc = s.list_S_P_P.val[k];
            add(i, s.list_S_P_P.ci[k], sc, sc, !s.amoeba, true);
        }
    if (s.amoeba)
        for (int i = 0; i < s.nmm; i++)
            for (int k = s.list_S_P_D.ri[i]; k < s.list_S_P_D.ri[i + 1]; k++)
                add(i, s.list_S_P_D.ci[k], s.list_S_P_D.val[k], 1.0, true, false);
    for (auto &r : rows) std::sort(r.begin(), r.end(), [](auto &a, auto &b) { return a.first < b.first; });
    return rows;
}

void upload_lists(System &s, DeviceState &d) {
    batch_release(d.batch);  // its segment points at the old lists
    free_csr(d, d.sp);
    free_csr(d, d.pp);
    free_csr(d, d.pd);
    free_csr(d, d.mp);
    free_csr(d, d.ss);
    const int npol = s.npol;
    auto rows = sp_rows(s);
    upload_csr(d, npol, rows, d.sp, true);
    // S-S corrections of the fixed-multipole energy (list_S_S, rows = source; transpose)
    rows.assign(s.nmm, {});
    for (int i = 0; i < s.nmm; i++)
        for (int k = s.list_S_S.ri[i]; k < s.list_S_S.ri[i + 1]; k++) {
            double2 f;
            f.x = 1.0 - s.list_S_S.val[k];
            f.y = 0.0;
            rows[s.list_S_S.ci[k]].push_back({i, f});
        }
    for (auto &r : rows) std::sort(r.begin(), r.end(), [](auto &a, auto &b) { return a.first < b.first; });
    upload_csr(d, s.nmm, rows, d.ss, false);
    // P->P corrections of the on-the-fly matvec (list_P_P, rows = source; transpose)
    rows.assign(npol, {});
    for (int ip = 0; ip < npol; ip++)
        for (int k = s.list_P_P.ri[ip]; k < s.list_P_P.ri[ip + 1]; k++) {
            double2 f;
            f.x = 1.0 - s.list_P_P.val[k];
            f.y = 0.0;
            rows[s.list_P_P.ci[k]].push_back({ip, f});
        }
    for (auto &r : rows) std::sort(r.begin(), r.end(), [](auto &a, auto &b) { return a.first < b.first; });
    upload_csr(d, npol, rows, d.pp, false);
    // dense-rule scale factors (values s, not 1 - s), rows = i
    rows.assign(npol, {});
    for (int ip = 0; ip < npol; ip++)
        for (int k = s.list_P_P_dense.ri[ip]; k < s.list_P_P_dense.ri[ip + 1]; k++) {
            double2 f;
            f.x = s.list_P_P_dense.val[k];
            f.y = 0.0;
            rows[ip].push_back({s.list_P_P_dense.ci[k], f});
        }
    upload_csr(d, npol, rows, d.pd, false);
}

void upload_geometry(System &s, DeviceState &d) {
    check_coincident_sites(s);  // once per geometry, not per solve
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<double4> ps(s.npol), ms(s.nmm);
    std::vector<float4> ps32(s.npol);
    std::vector<double> pa(s.npol), ma(s.nmm);
    std::vector<float> pa32(s.npol);
    for (int i = 0; i < s.nmm; i++) {
        const double t = s.thole[i];
        ms[i] = make_double4(s.cmm[3 * i], s.cmm[3 * i + 1], s.cmm[3 * i + 2], t > kEpsRp ? 1.0 / t : inf);
        ma[i] = s.mm_polar[i] >= 0 ? s.thole_damping_mm[i] : inf;
    }
    for (int p = 0; p < s.npol; p++) {
        const int i = s.polar_mm[p];
        ps[p] = ms[i];
        pa[p] = s.thole_damping_mm[i];
        ps32[p] = make_float4((float)ps[p].x, (float)ps[p].y, (float)ps[p].z, (float)ps[p].w);
        pa32[p] = (float)pa[p];
    }
    upload(ps.data(), d.ps64, sizeof(double4) * s.npol);
    upload(ps32.data(), d.ps32, sizeof(float4) * s.npol);
    upload(pa.data(), d.pa64, sizeof(double) * s.npol);
    upload(pa32.data(), d.pa32, sizeof(float) * s.npol);
    upload(s.pol.data(), d.alpha, sizeof(double) * s.npol);
    upload(s.polar_mm.data(), d.polar_mm, sizeof(int) * s.npol);
    upload(ms.data(), d.ms64, sizeof(double4) * s.nmm);
    upload(ma.data(), d.ma64, sizeof(double) * s.nmm);
    upload(s.q.data(), d.mq, sizeof(double) * s.ld_cart * s.nmm);
    d.tpack_valid = false;
    d.lu_valid = false;
    d.nb.list_valid = false;
    d.nb.valid = false;
    s.dev_geometry_dirty = false;
}

DeviceState &ensure_device(System &s) {
    if (!s.dev) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            fatal("No GPU available: this library runs the polarization solver on an AMD Instinct GPU only");
        DeviceState *dp = new DeviceState();
        DeviceState &d = *dp;
        HCHECK(hipGetDevice(&d.device));
        if (s.user_stream) {
            d.stream = (hipStream_t)s.user_stream;
        } else {
            HCHECK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
            d.own_stream = true;
        }
        // ev[0..2] only time phases; the host waits on ev[3] before reading results.
        // Timing-only events skip the cache writeback and invalidation a default
        // event performs when it completes (it costs ~5-10 us of idle GPU each).
        for (int k = 0; k < 3; k++) HCHECK(hipEventCreateWithFlags(&d.ev[k], hipEventDisableSystemFence));
        HCHECK(hipEventCreate(&d.ev[3]));
        d.npol = s.npol;
        d.nmm = s.nmm;
        d.ld = s.ld_cart;
        d.n = 3L * s.npol;
        const long n = d.n;
        d.ps64 = dalloc<double4>(d, s.npol);
        d.ps32 = dalloc<float4>(d, s.npol);
        d.pa64 = dalloc<double>(d, s.npol);
        d.pa32 = dalloc<float>(d, s.npol);
        d.alpha = dalloc<double>(d, s.npol);
        d.polar_mm = dalloc<int>(d, s.npol);
        d.ms64 = dalloc<double4>(d, s.nmm);
        d.ma64 = dalloc<double>(d, s.nmm);
        d.mq = dalloc<double>(d, (size_t)s.ld_cart * s.nmm);
        d.e_m2d = dalloc<double>(d, 2 * n);
        d.ipd = dalloc<double>(d, 2 * n);
        d.ext = dalloc<double>(d, n);
        double **vecs[] = {&d.vb, &d.vx, &d.vh, &d.vr, &d.vz, &d.vp, &d.vd, &d.vr2};
        for (auto v : vecs) *v = dalloc<double>(d, 2 * n);
        const int nb = std::max(vec_blocks(n), site_blocks(s.npol)) + 8;
        d.red_a = dalloc<double>(d, 24 * (size_t)nb);
        d.red_b = dalloc<double>(d, 24 * (size_t)nb);
        d.red_c = dalloc<double>(d, 24 * (size_t)nb);
        d.scal = dalloc<double>(d, 64);
        d.tol_dev = dalloc<double>(d, 4);
        d.cs = dalloc<CgState>(d, 2);
        HCHECK(hipHostMalloc((void **)&d.cs_host, 2 * sizeof(CgState), hipHostMallocDefault));
        HCHECK(hipHostMalloc((void **)&d.scal_host, 64 * sizeof(double), hipHostMallocDefault));
        d.nsplit = tmv_split(s.npol);
        d.mvpart_bytes = tmv_part_bytes(s.npol, 2, 64, d.nsplit);
        d.mvpart = dmalloc(d, d.mvpart_bytes);
        {
            hipDeviceProp_t prop;
            HCHECK(hipGetDeviceProperties(&prop, d.device));
            // the FP64 static-field sweep hides its latency better with about six
            // workgroups per CU than with the ~3 of the balanced split (-8% at 15,000 sites)
            const int tb = std::max(1, site_blocks(s.npol)), ncu = prop.multiProcessorCount;
            const int max_s = std::max(1, (s.nmm + kBlock - 1) / kBlock);
            d.rhs_split = std::min(max_s, std::max(balanced_split(tb, s.nmm, ncu), (6 * ncu + tb - 1) / tb));
            const int nnp = s.nmm - s.npol, max_np = std::max(1, (nnp + kBlock - 1) / kBlock);
            d.rhs_split_np = std::min(max_np, std::max(balanced_split(tb, std::max(nnp, 1), ncu), (6 * ncu + tb - 1) / tb));
        }
        d.rhs_part = dalloc<double>(d, (size_t)std::max(d.rhs_split, d.rhs_split_np) * n + 1);
        if (s.nmm > s.npol) {
            std::vector<int> np_src;
            np_src.reserve(s.nmm - s.npol);
            for (int i = 0; i < s.nmm; i++)
                if (s.mm_polar[i] < 0) np_src.push_back(i);
            d.np_src = dalloc<int>(d, np_src.size());
            upload(np_src.data(), d.np_src, sizeof(int) * np_src.size());
        }
        HCHECK(hipMemset(d.ipd, 0, sizeof(double) * 2 * n));
        HCHECK(hipMemset(d.e_m2d, 0, sizeof(double) * 2 * n));
        s.dev = dp;
        upload_lists(s, d);
        s.dev_geometry_dirty = true;
    }
    DeviceState &d = *s.dev;
    if (s.dev_geometry_dirty) upload_geometry(s, d);
    return d;
}

template <typename T> T *fmm_alloc(DeviceState &d, size_t count) {
    T *p = dalloc<T>(d, count);
    d.fmm_ptrs.push_back((void *)p);
    return p;
}

void fmm_release(DeviceState &d) {
    for (void *p : d.fmm_ptrs) dfree(d, p);
    d.fmm_ptrs.clear();
    d.fmm = FmmDev();
    d.fmm_part = d.fmm_rhs_part = d.fmm_gpart_pol = d.fmm_gpart_mm = nullptr;
    d.fmm_valid = false;
}

// Octree, sorted copies of the sites and atoms, tables and expansion storage
// for the current coordinates.
void ensure_fmm(System &s, DeviceState &d) {
    if (d.fmm_valid) return;
    fmm_release(d);
    const FmmHost h = fmm_build(s, s.fmm_lmax_pol, s.fmm_lmax_static, s.fmm_min_cell, s.fmm_distance);
    const int npol = s.npol, nmm = s.nmm, ld = s.ld_cart;
    FmmDev &f = d.fmm;
    f.npol = npol;
    f.nmm = nmm;
    f.levels = h.levels;
    f.p_pol = h.p_pol;
    f.p_static = h.p_static;
    f.p_tab = h.p_tab;
    f.ncoef_tab = h.ncoef_tab;
    f.ncoef2_tab = h.ncoef2_tab;
    f.leaf_h = h.leaf_h;
    f.origin = make_double3(h.origin[0], h.origin[1], h.origin[2]);
    for (int l = 0; l < (int)h.cell_off.size(); l++) f.cell_off[l] = h.cell_off[l];
    const double inf = std::numeric_limits<double>::infinity();
    auto site = [&](int i) {
        const double t = s.thole[i];
        return make_double4(s.cmm[3 * i], s.cmm[3 * i + 1], s.cmm[3 * i + 2], t > kEpsRp ? 1.0 / t : inf);
    };
    std::vector<double4> ppos(npol), mpos(nmm);
    std::vector<double> pa(npol), ma(nmm), mq((size_t)nmm * ld);
    std::vector<int> pself(npol);
    for (int t = 0; t < npol; t++) {
        const int i = s.polar_mm[h.pol_order[t]];
        ppos[t] = site(i);
        pa[t] = s.thole_damping_mm[i];
        pself[t] = t;
    }
    for (int t = 0; t < nmm; t++) {
        const int i = h.mm_order[t];
        mpos[t] = site(i);
        ma[t] = s.mm_polar[i] >= 0 ? s.thole_damping_mm[i] : inf;
        for (int c = 0; c < ld; c++) mq[(size_t)t * ld + c] = s.q[(size_t)i * ld + c];
    }
    auto up_i = [&](const std::vector<int> &v) {
        int *p = fmm_alloc<int>(d, std::max<size_t>(v.size(), 1));
        upload(v.data(), p, sizeof(int) * v.size());
        return p;
    };
    f.pol_order = up_i(h.pol_order);
    f.pol_leaf = up_i(h.pol_leaf);
    f.pol_self = up_i(pself);
    f.pol_mmself = up_i(h.pol_mmself);
    f.pol_leaf_start = up_i(h.pol_leaf_start);
    f.mm_order = up_i(h.mm_order);
    f.mm_leaf = up_i(h.mm_leaf);
    f.mm_leaf_start = up_i(h.mm_leaf_start);
    f.cnt_pol = up_i(h.cnt_pol);
    f.cnt_mm = up_i(h.cnt_mm);
    f.pol_pos = fmm_alloc<double4>(d, npol);
    upload(ppos.data(), f.pol_pos, sizeof(double4) * npol);
    f.pol_a = fmm_alloc<double>(d, npol);
    upload(pa.data(), f.pol_a, sizeof(double) * npol);
    f.mm_pos = fmm_alloc<double4>(d, nmm);
    upload(mpos.data(), f.mm_pos, sizeof(double4) * nmm);
    f.mm_a = fmm_alloc<double>(d, nmm);
    upload(ma.data(), f.mm_a, sizeof(double) * nmm);
    f.mm_q = fmm_alloc<double>(d, mq.size());
    upload(mq.data(), f.mm_q, sizeof(double) * mq.size());
    auto up_c = [&](const std::vector<double2> &v) {
        double2 *p = fmm_alloc<double2>(d, std::max<size_t>(v.size(), 1));
        upload(v.data(), p, sizeof(double2) * v.size());
        return p;
    };
    f.tab_m2m = up_c(h.tab_m2m);
    f.tab_l2l = up_c(h.tab_l2l);
    f.tab_m2l = up_c(h.tab_m2l);
    const size_t ncell = (size_t)h.cell_off.back();
    f.mexp = fmm_alloc<double2>(d, ncell * f.ncoef_tab);
    f.lexp = fmm_alloc<double2>(d, ncell * f.ncoef_tab);
    f.xs = fmm_alloc<double>(d, 9L * npol);
    f.mstatic = fmm_alloc<double2>(d, std::max<size_t>((size_t)(h.cell_off.back() - h.cell_off[h.levels]), 1) *
                                          f.ncoef_tab);
    d.fmm_part = fmm_alloc<double>(d, (size_t)(kFmmNearSplits + 1) * 2 * 3 * npol);
    d.fmm_rhs_part = fmm_alloc<double>(d, (size_t)(kFmmNearSplits + 1) * 3 * npol);
    fmm_launch_static_moments(d.stream, f, ld);
    d.fmm_valid = true;
}

// Block pairs and block-pair partials of the pair-once sweeps (T matvec and
// static field), on first use.
void ensure_sym(const System &s, DeviceState &d) {
    if (d.sym_pairs) return;
    const int nb = tmv_sym_blocks(s.npol);
    std::vector<int2> pr;
    pr.reserve((size_t)nb * (nb + 1) / 2);
    for (int I = 0; I < nb; I++)
        for (int J = I; J < nb; J++) pr.push_back(make_int2(I, J));
    d.sym_pairs = dalloc<int2>(d, pr.size());
    upload(pr.data(), d.sym_pairs, sizeof(int2) * pr.size());
    d.sympart = dmalloc(d, tmv_sym_part_bytes(s.npol, 2, 64));
}

// Number of y.x partials per column that matvec() of this kind writes to dotpart.
static int matvec_dot_parts(const System &s, int kind, int M) {
    if ((kind == kOtf64 || kind == kOtf32) && tmv_pair_once(s.npol, kind == kOtf32 ? 32 : 64, M))
        return tmv_sym_dot_parts(s.npol);
    return site_blocks(s.npol);
}

void matvec(System &s, DeviceState &d, int kind, int M, const double *x, double *y, bool dodiag,
            double *dotpart, const CgState *skip) {
    const long n = d.n;
    if (kind == kFmm) {
        ensure_fmm(s, d);
        fmm_launch_matvec_fields(d.stream, d.fmm, s.amoeba, M, x, n, d.fmm_part);
        launch_tmv_finalize64(d.stream, s.amoeba, s.npol, M, d.fmm_part, kFmmNearSplits + 1, d.ps64, d.pa64, d.alpha,
                              d.pp, x, n, y, n, dodiag, dotpart, reinterpret_cast<const int *>(skip));
        return;
    }
    if (kind == kOtf64 || kind == kOtf32) {
        const bool sym = tmv_pair_once(s.npol, kind == kOtf32 ? 32 : 64, M);
        kind = kind == kOtf32 ? (sym ? kSym32 : kPlain32) : (sym ? kSym64 : kPlain64);
    }
    if (kind == kSym64 || kind == kSym32) {
        ensure_sym(s, d);
        const int pi = prof_begin(d, kind == kSym32 ? kOtf32 : kOtf64);
        launch_tmv_sym(d.stream, kind == kSym32 ? 32 : 64, s.amoeba, s.npol, M, d.sym_pairs, d.ps64, d.pa64, d.ps32,
                       d.pa32, d.alpha, d.pp, x, n, y, n, dodiag, dotpart, d.sympart,
                       reinterpret_cast<const int *>(skip), pi >= 0 ? d.prof_pool[2 * pi] : nullptr,
                       pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr);
        return;
    }
    if (kind == kPlain64 || kind == kPlain32) {
        const int pi = prof_begin(d, kind == kPlain32 ? kOtf32 : kOtf64);
        launch_tmv_otf(d.stream, kind == kPlain32 ? 32 : 64, s.amoeba, s.npol, M, d.ps64, d.pa64, d.ps32, d.pa32,
                       d.alpha, d.pp, x, n, y, n, dodiag, dotpart, d.mvpart, d.nsplit,
                       reinterpret_cast<const int *>(skip), pi >= 0 ? d.prof_pool[2 * pi] : nullptr,
                       pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr);
    } else {
        const int prec = kind == kPack32 ? 32 : 64;
        if (!d.tpack_valid || d.tpack_prec != prec) {
            dfree(d, d.tpack);
            dfree(d, d.tpart);
            size_t fr = 0, tot = 0;
            HCHECK(hipMemGetInfo(&fr, &tot));
            const size_t need = tpack_bytes(s.npol, prec) + tpack_part_bytes(s.npol, 2);
            if (need > fr) {
                char buf[256];
                std::snprintf(buf, sizeof(buf),
                              "In-core T matrix needs %.2f GB of device memory, %.2f GB are free", need / 1e9,
                              fr / 1e9);
                fatal(buf);
            }
            d.tpack = dmalloc(d, tpack_bytes(s.npol, prec));
            d.tpart = dmalloc(d, tpack_part_bytes(s.npol, 2));
            launch_tpack_build(d.stream, prec, s.amoeba, s.npol, d.ps64, d.pa64, d.tpack, d.pd.ri, d.pd.ci, d.pd.f);
            d.tpack_prec = prec;
            d.tpack_valid = true;
        }
        const int pi = prof_begin(d, kind);
        launch_tpack_mv(d.stream, prec, s.npol, M, d.tpack, d.alpha, x, n, y, n, dodiag, dotpart, d.tpart,
                        reinterpret_cast<const int *>(skip), pi >= 0 ? d.prof_pool[2 * pi] : nullptr,
                        pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr);
    }
}

// Neighbour lists of the near-field preconditioner: for every polarizable site
// the sites closer than kNbCut, in site order.  A function of the coordinates
// alone, built on the host with a grid of cells at least kNbCut wide (only the
// 27 cells around a site can hold its neighbours) when the coordinates reach
// the device; distances in FP64, the same for (i, j) and (j, i), so the lists
// are symmetric.
void build_nb_lists(const System &s, std::vector<int> &row_ptr, std::vector<int> &ci) {
    const int n = s.npol;
    row_ptr.assign(n + 1, 0);
    ci.clear();
    if (n == 0) return;
    double lo[3], hi[3];
    for (int c = 0; c < 3; c++) lo[c] = hi[c] = s.cpol[c];
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], s.cpol[3 * i + c]);
            hi[c] = std::max(hi[c], s.cpol[3 * i + c]);
        }
    // cells at least kNbCut wide, at most ~8 per site
    double h = kNbCut;
    long dim[3];
    for (;;) {
        for (int c = 0; c < 3; c++) dim[c] = (long)std::floor((hi[c] - lo[c]) / h) + 1;
        if (dim[0] * dim[1] * dim[2] <= 8L * n + 64) break;
        h *= 1.5;
    }
    auto cell_of = [&](int i, long *g) {
        for (int c = 0; c < 3; c++)
            g[c] = std::min(dim[c] - 1, (long)std::floor((s.cpol[3 * i + c] - lo[c]) / h));
    };
    const long ncell = dim[0] * dim[1] * dim[2];
    std::vector<int> start(ncell + 1, 0), order(n), cell(n);
    for (int i = 0; i < n; i++) {
        long g[3];
        cell_of(i, g);
        cell[i] = (int)((g[2] * dim[1] + g[1]) * dim[0] + g[0]);
        start[cell[i] + 1]++;
    }
    for (long k = 0; k < ncell; k++) start[k + 1] += start[k];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; i++) order[fill[cell[i]]++] = i;
    }
    const double rc2 = kNbCut * kNbCut;
    std::vector<int> row;
    ci.reserve((size_t)n * 12);
    for (int i = 0; i < n; i++) {
        long g[3];
        cell_of(i, g);
        row.clear();
        for (long z = std::max(0L, g[2] - 1); z <= std::min(dim[2] - 1, g[2] + 1); z++)
            for (long y = std::max(0L, g[1] - 1); y <= std::min(dim[1] - 1, g[1] + 1); y++)
                for (long x = std::max(0L, g[0] - 1); x <= std::min(dim[0] - 1, g[0] + 1); x++) {
                    const long k = (z * dim[1] + y) * dim[0] + x;
                    for (int q = start[k]; q < start[k + 1]; q++) {
                        const int j = order[q];
                        if (j == i) continue;
                        const double dx = s.cpol[3 * i] - s.cpol[3 * j], dy = s.cpol[3 * i + 1] - s.cpol[3 * j + 1],
                                     dz = s.cpol[3 * i + 2] - s.cpol[3 * j + 2];
                        if (dx * dx + dy * dy + dz * dz < rc2) row.push_back(j);
                    }
                }
        std::sort(row.begin(), row.end());
        ci.insert(ci.end(), row.begin(), row.end());
        row_ptr[i + 1] = (int)ci.size();
    }
}

// The preconditioner of the current geometry on the device: lists uploaded
// once per coordinates, tensors B recomputed before the first mixed solve of a
// geometry (and after a reset of the static field), which also clears a
// breakdown switch-off.
void ensure_nb(System &s, DeviceState &d) {
    NbDev &nb = d.nb;
    if (!nb.list_valid) {
        std::vector<int> rp, ci;
        build_nb_lists(s, rp, ci);
        const long nnz = (long)ci.size();
        if (!nb.row_ptr) {
            nb.row_ptr = dalloc<int>(d, s.npol + 1);
            nb.off_flag = dalloc<int>(d, 1);
            d.vr3 = dalloc<double>(d, 2 * d.n);
            d.red_d = dalloc<double>(d, 2 * (size_t)nb_blocks(s.npol) + 8);
        }
        if (nnz > nb.cap) {
            dfree(d, nb.ci);
            dfree(d, nb.rowi);
            dfree(d, nb.b);
            nb.cap = std::max(1L, nnz + nnz / 4);
            nb.ci = dalloc<int>(d, nb.cap);
            nb.rowi = dalloc<int>(d, nb.cap);
            nb.b = dalloc<float>(d, 6 * (size_t)nb.cap);
        }
        std::vector<int> rowi(nnz);
        for (int i = 0; i < s.npol; i++)
            for (int k = rp[i]; k < rp[i + 1]; k++) rowi[k] = i;
        upload(rp.data(), nb.row_ptr, sizeof(int) * (s.npol + 1));
        upload(ci.data(), nb.ci, sizeof(int) * nnz);
        upload(rowi.data(), nb.rowi, sizeof(int) * nnz);
        nb.nnz = nnz;
        nb.list_valid = true;
        nb.valid = false;
    }
    if (!nb.valid) {
        HCHECK(hipMemsetAsync(nb.off_flag, 0, sizeof(int), d.stream));
        launch_nb_tensors(d.stream, s.amoeba, nb, d.ps32, d.pa32, d.alpha, d.pp);
        nb.valid = true;
    }
}

// conjugate_gradient_solver (mod_solvers.F90:86-216) on M columns that share
// the matvec.  Vectors are column-strided by n.  Returns per-column status.
struct CgResult {
    int iters[2] = {0, 0};
    int status[2] = {2, 2};
    double rms[2] = {0, 0};
    double beta[2] = {0, 0};
};

// With nb (inner solves of the mixed solver, zero guess only) the iterations
// use the near-field preconditioner, r alternating between vr and vr3.
CgResult cg_solve(System &s, DeviceState &d, int kind, int M, const double *b, double *x, const double *tol,
                  int maxit, bool zero_guess, double small_step_eps, int slot, bool nb = false) {
    const long n = d.n;
    CgResult res;
    double te[4] = {tol[0], M > 1 ? tol[1] : tol[0], small_step_eps, small_step_eps};
    HCHECK(hipMemcpyAsync(d.tol_dev, te, sizeof(te), hipMemcpyHostToDevice, d.stream));
    if (nb && !zero_guess) fatal("Internal error: the near-field preconditioned solve starts from zero");
    if (!zero_guess) {
        launch_norm2_guess(d.stream, n, M, x, d.alpha, d.red_a);
        matvec(s, d, kind, M, x, d.vh, true, nullptr, nullptr);
    }
    double *rb[2] = {d.vr, nb ? d.vr3 : d.vr};
    if (nb)
        launch_cg_init_nb(d.stream, s.npol, M, b, x, rb[0], d.vz, d.vp, d.alpha, d.nb, d.red_a, d.cs, d.tol_dev,
                          nullptr);
    else
        launch_cg_init(d.stream, n, M, b, d.vh, x, d.vr, d.vz, d.vp, d.alpha, d.red_a, d.cs, zero_guess, d.tol_dev);
    const int nred_hp = matvec_dot_parts(s, kind, M);
    int it = 1;
    // first batch: as many iterations as this call site needed last time
    const int pred = (slot >= 0 && slot < 16) ? s.pred_iters[slot] : 0;
    int chunk = pred > 0 ? pred : 4;
    bool all_done = false;
    while (it <= maxit && !all_done) {
        const int end = std::min(maxit, it + chunk - 1);
        for (; it <= end; it++) {
            matvec(s, d, kind, M, d.vp, d.vh, true, d.red_b, d.cs);
            if (nb) {
                launch_cg_update_nb(d.stream, s.npol, M, it, d.vh, x, rb[(it - 1) & 1], rb[it & 1], d.vz, d.vp,
                                    d.alpha, d.red_b, nred_hp, d.red_c, d.red_d, d.cs, d.nb);
                launch_cg_direction_nb(d.stream, s.npol, M, it, d.vz, d.vp, d.red_c, d.red_d, d.cs, d.nb);
            } else {
                launch_cg_update(d.stream, n, M, it, d.vh, x, d.vr, d.vz, d.vp, d.alpha, d.red_b, nred_hp, d.red_c,
                                 d.cs);
                launch_cg_direction(d.stream, n, M, it, d.vz, d.vp, d.red_c, d.cs, nullptr);
            }
        }
        HCHECK(hipMemcpyAsync(d.cs_host, d.cs, sizeof(CgState) * M, hipMemcpyDeviceToHost, d.stream));
        HCHECK(hipStreamSynchronize(d.stream));
        all_done = true;
        for (int m = 0; m < M; m++) all_done = all_done && d.cs_host[m].done;
        chunk = 2;
    }
    int most = 0;
    for (int m = 0; m < M; m++) most = std::max(most, d.cs_host[m].done ? d.cs_host[m].iters : maxit);
    if (slot >= 0 && slot < 16) s.pred_iters[slot] = most;
    for (int m = 0; m < M; m++) {
        res.iters[m] = d.cs_host[m].done ? d.cs_host[m].iters : maxit;
        res.status[m] = d.cs_host[m].done ? d.cs_host[m].status : 2;
        res.rms[m] = d.cs_host[m].rms;
        res.beta[m] = d.cs_host[m].last_beta;
    }
    return res;
}

// Pulay DIIS on Jacobi iterations (jacobi_diis_solver, diis, makeb;
// mod_solvers.F90:218-469).  One column; the B matrix and its (m+1)^2 solve
// stay on the host, vectors and dot products on the device.
bool host_lu_solve(int n, std::vector<double> A, std::vector<double> &b) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[i * n + k]) > mx) {
                mx = std::fabs(A[i * n + k]);
                p = i;
            }
        piv[k] = p;
        if (A[p * n + k] == 0.0) return false;
        if (p != k)
            for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[p * n + j]);
        for (int i = k + 1; i < n; i++) {
            A[i * n + k] /= A[k * n + k];
            for (int j = k + 1; j < n; j++) A[i * n + j] -= A[i * n + k] * A[k * n + j];
        }
    }
    for (int k = 0; k < n; k++)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) b[i] -= A[i * n + j] * b[j];
    for (int i = n - 1; i >= 0; i--) {
        for (int j = i + 1; j < n; j++) b[i] -= A[i * n + j] * b[j];
        b[i] /= A[i * n + i];
    }
    return true;
}

int diis_solve(System &s, DeviceState &d, int kind, const double *rhs, double *x, double tol, int maxit,
               int *iters_out) {
    const long n = d.n;
    const int ndiis = 20;
    if (!d.diis_x) {
        d.diis_x = dalloc<double>(d, (size_t)ndiis * n);
        d.diis_e = dalloc<double>(d, (size_t)ndiis * n);
        d.diis_coef = dalloc<double>(d, 32);
    }
    // zero-norm guess -> inv_diag * rhs (mod_solvers.F90:300-308)
    launch_dot(d.stream, n, x, x, d.red_a, d.scal);
    HCHECK(hipMemcpyAsync(d.scal_host, d.scal, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    if (d.scal_host[0] < kEpsRp) {
        HCHECK(hipMemsetAsync(d.vz, 0, sizeof(double) * n, d.stream));
        launch_jacobi(d.stream, n, rhs, d.vz, d.vz, d.alpha, x, d.vr2);  // x = alpha * rhs
    }
    std::vector<int> slot(ndiis);
    for (int k = 0; k < ndiis; k++) slot[k] = k;
    std::vector<double> B((ndiis + 1) * (ndiis + 1), 0.0);
    const int W = ndiis + 1;
    int nmat = 1;
    double max_norm = 0.0;
    int it;
    double *xnew = d.vd, *y = d.vh;
    for (it = 1; it <= maxit; it++) {
        matvec(s, d, kind, 1, x, y, false, nullptr, nullptr);
        double *xs = d.diis_x + (size_t)slot[nmat - 1] * n;
        double *es = d.diis_e + (size_t)slot[nmat - 1] * n;
        launch_jacobi(d.stream, n, rhs, y, x, d.alpha, xs, es);
        // diis(): shift when full (keep the last ndiis-10 points)
        if (nmat >= ndiis) {
            for (int j = 2; j <= nmat - 10; j++)
                for (int k = 2; k <= nmat - 10; k++) B[(j - 1) * W + (k - 1)] = B[(j + 9) * W + (k + 9)];
            std::vector<int> ns(slot);
            for (int j = 1; j <= nmat - 10; j++) ns[j - 1] = slot[j + 9];
            for (int j = nmat - 10 + 1; j <= ndiis; j++) ns[j - 1] = slot[j - 1 - (nmat - 10)];
            slot = ns;
            nmat -= 10;
        }
        // makeb: new row against all stored error vectors
        std::vector<const double *> ev(nmat);
        for (int q = 0; q < nmat; q++) ev[q] = d.diis_e + (size_t)slot[q] * n;
        launch_multi_dot(d.stream, n, nmat, ev.data(), ev[nmat - 1], d.red_a, d.scal);
        HCHECK(hipMemcpyAsync(d.scal_host, d.scal, sizeof(double) * nmat, hipMemcpyDeviceToHost, d.stream));
        HCHECK(hipStreamSynchronize(d.stream));
        if (nmat == 1) {
            B[0] = 0.0;
            B[1] = 1.0;
            B[W] = 1.0;
            B[W + 1] = d.scal_host[0];
        } else {
            B[nmat * W] = 1.0;
            B[nmat] = 1.0;
            for (int i = 1; i <= nmat - 1; i++) B[nmat * W + i] = B[i * W + nmat] = d.scal_host[i - 1];
            B[nmat * W + nmat] = d.scal_host[nmat - 1];
        }
        const int nm1 = nmat + 1;
        std::vector<double> bloc(nm1 * nm1), cex(nm1, 0.0);
        for (int i = 0; i < nm1; i++)
            for (int j = 0; j < nm1; j++) bloc[i * nm1 + j] = B[i * W + j];
        cex[0] = 1.0;
        const double *jac = xs;
        if (host_lu_solve(nm1, bloc, cex)) {
            std::vector<const double *> xv(nmat);
            for (int q = 0; q < nmat; q++) xv[q] = d.diis_x + (size_t)slot[q] * n;
            HCHECK(hipMemcpyAsync(d.diis_coef, cex.data() + 1, sizeof(double) * nmat, hipMemcpyHostToDevice,
                                  d.stream));
            launch_lincomb(d.stream, n, nmat, xv.data(), d.diis_coef, xnew);
            nmat += 1;
        } else {
            launch_copy(d.stream, n, jac, xnew);
            nmat = 1;
        }
        launch_maxabs_diff(d.stream, n, xnew, x, d.red_a, d.scal);
        launch_copy(d.stream, n, xnew, x);
        HCHECK(hipMemcpyAsync(d.scal_host, d.scal, sizeof(double), hipMemcpyDeviceToHost, d.stream));
        HCHECK(hipStreamSynchronize(d.stream));
        max_norm = d.scal_host[0];
        if (max_norm < tol) break;
    }
    *iters_out = std::min(it, maxit);
    return max_norm > tol ? 1 : 0;
}

// FP32 matvec inside FP64 iterative refinement, columns solved together.
struct MixedResult {
    int outer = 0, inner = 0;
    bool ok = false;
};

// Refinement steps enqueued without host round trips, for a solve shaped like
// the previous one: K FP64 residuals with K - 1 inner solves of the predicted
// lengths between them.  Every decision is taken on the device
// (k_refine_decide): a converged column's later inner solves start done and its
// FP64 residual matvec is skipped once every column has converged.  An inner
// solve cut short by its predicted length only leaves a larger residual for the
// next step.  Returns -1 when there is no prediction (nothing enqueued), 0 when
// some column has not converged within the K steps (res.inner counts the inner
// iterations run; the caller goes on from x), K when every column converged.
constexpr int kSpecMaxOuter = 6;

int mixed_solve_ahead(System &s, DeviceState &d, int M, const double *b, double *x, bool zero_guess, double tol,
                      int kind32, bool nb, MixedResult &res) {
    const int K = s.pred_outer;
    if (K < 2 || K > kSpecMaxOuter) return -1;
    for (int o = 0; o < K - 1; o++)
        if (s.pred_iters[2 + o] <= 0) return -1;
    const long n = d.n;
    const int W = 2 * (kSpecMaxOuter + 1);
    if (!d.cs_spec) {
        d.cs_spec = dalloc<CgState>(d, W);
        HCHECK(hipHostMalloc((void **)&d.cs_spec_host, W * sizeof(CgState), hipHostMallocDefault));
    }
    CgState *out = d.cs_spec;
    const int nred_hp = matvec_dot_parts(s, kind32, M);
    bool xzero = zero_guess;
    for (int o = 0; o < K; o++) {
        if (xzero) {
            HCHECK(hipMemsetAsync(d.vh, 0, sizeof(double) * M * n, d.stream));
        } else {
            // out still holds the previous solve's state until step 0 is decided
            matvec(s, d, kOtf64, M, x, d.vh, true, nullptr, o == 0 ? nullptr : out);
        }
        launch_residual_rms(d.stream, n, M, b, d.vh, d.vr2, d.alpha, d.red_a, d.scal);
        launch_refine_decide(d.stream, M, d.scal, out, tol, refine_eta(), refine_frac(), d.tol_dev, o);
        if (o == K - 1) break;
        CgState *cs = d.cs_spec + 2 * (o + 1);
        double *rb[2] = {d.vr, nb ? d.vr3 : d.vr};
        if (nb)
            launch_cg_init_nb(d.stream, s.npol, M, d.vr2, d.vd, rb[0], d.vz, d.vp, d.alpha, d.nb, d.red_a, cs,
                              d.tol_dev, out);
        else
            launch_cg_init(d.stream, n, M, d.vr2, d.vh, d.vd, d.vr, d.vz, d.vp, d.alpha, d.red_a, cs, true,
                           d.tol_dev, out);
        for (int it = 1; it <= s.pred_iters[2 + o]; it++) {
            matvec(s, d, kind32, M, d.vp, d.vh, true, d.red_b, cs);
            if (nb) {
                launch_cg_update_nb(d.stream, s.npol, M, it, d.vh, d.vd, rb[(it - 1) & 1], rb[it & 1], d.vz, d.vp,
                                    d.alpha, d.red_b, nred_hp, d.red_c, d.red_d, cs, d.nb);
                launch_cg_direction_nb(d.stream, s.npol, M, it, d.vz, d.vp, d.red_c, d.red_d, cs, d.nb);
            } else {
                launch_cg_update(d.stream, n, M, it, d.vh, d.vd, d.vr, d.vz, d.vp, d.alpha, d.red_b, nred_hp, d.red_c,
                                 cs);
                launch_cg_direction(d.stream, n, M, it, d.vz, d.vp, d.red_c, cs, nullptr);
            }
        }
        launch_axpy(d.stream, n, M, 1.0, d.vd, x);
        xzero = false;
    }
    HCHECK(hipMemcpyAsync(d.cs_spec_host, d.cs_spec, sizeof(CgState) * 2 * K, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    const CgState *h = d.cs_spec_host;
    // next predictions: an inner solve that ran out of launches gets two more
    for (int o = 0; o < K - 1; o++) {
        const int ran = s.pred_iters[2 + o];
        int most = 0;
        for (int m = 0; m < M; m++)
            most = std::max(most, h[2 * (o + 1) + m].done ? h[2 * (o + 1) + m].iters : ran + 2);
        s.pred_iters[2 + o] = most;
        res.inner += std::min(most, ran);
    }
    bool conv = true;
    res.outer = 0;
    for (int m = 0; m < M; m++) {
        conv = conv && h[m].done;
        res.outer = std::max(res.outer, h[m].iters);
    }
    if (!conv) return 0;
    res.ok = true;
    return K;
}

MixedResult mixed_solve(System &s, DeviceState &d, int M, const double *b, double *x, bool zero_guess, double tol,
                        int maxit, int kind32) {
    const long n = d.n;
    MixedResult res;
    const int max_outer = 30;
    // Inner target per refinement step: max(refine_eta() * rms, refine_frac() * tol).
    const double eta_min = refine_eta(), frac = refine_frac();
    const int K = s.pred_outer;
    // near-field preconditioner for the inner solves (any SPD one will do: the
    // FP64 residual decides convergence)
    const bool nb = s.inner_precond == 1 && s.npol > 0;
    if (nb) ensure_nb(s, d);
    const int ahead = mixed_solve_ahead(s, d, M, b, x, zero_guess, tol, kind32, nb, res);
    if (ahead > 0) {
        s.pred_outer = res.outer;
        return res;
    }
    // one refinement step per host round trip, from x (after the steps enqueued
    // ahead, if any: the last of them is redone with its inner solve)
    bool xzero = zero_guess && ahead < 0;
    for (int outer = ahead < 0 ? 0 : K - 1; outer < max_outer; outer++) {
        // FP64 residual r = b - T x and its preconditioned rms per column
        if (xzero) {
            HCHECK(hipMemsetAsync(d.vh, 0, sizeof(double) * M * n, d.stream));
            launch_residual_rms(d.stream, n, M, b, d.vh, d.vr2, d.alpha, d.red_a, d.scal);
        } else {
            matvec(s, d, kOtf64, M, x, d.vh, true, nullptr, nullptr);
            launch_residual_rms(d.stream, n, M, b, d.vh, d.vr2, d.alpha, d.red_a, d.scal);
        }
        HCHECK(hipMemcpyAsync(d.scal_host, d.scal, sizeof(double) * M, hipMemcpyDeviceToHost, d.stream));
        HCHECK(hipStreamSynchronize(d.stream));
        res.outer = outer + 1;
        bool conv = true;
        double tin[2];
        for (int m = 0; m < M; m++) {
            const double rms = d.scal_host[m];
            conv = conv && rms < tol;
            tin[m] = rms < tol ? 1e300 : std::max(eta_min * rms, frac * tol);
        }
        if (conv) {
            res.ok = true;
            s.pred_outer = res.outer;
            return res;
        }
        // inner FP32 PCG on T delta = r from delta = 0
        CgResult cr = cg_solve(s, d, kind32, M, d.vr2, d.vd, tin, maxit, true, 0.0, 2 + std::min(outer, 13), nb);
        res.inner += std::max(cr.iters[0], M > 1 ? cr.iters[1] : 0);
        launch_axpy(d.stream, n, M, 1.0, d.vd, x);
        xzero = false;
    }
    return res;
}

}  // namespace

// The FP32 T matvec is accurate to ~1e-7 relative, so an inner solve can
// reduce the FP64 residual by ~1e-6 before its recursive residual drifts from
// the true one; asking for that much (instead of stopping at the tolerance)
// makes one refinement step enough for most warm-started solves.
double refine_eta() { return 1e-6; }
double refine_frac() { return 0.8; }

// ===================================================================== API
void device_release(System &s) {
    if (!s.dev) return;
    DeviceState &d = *s.dev;
    (void)hipStreamSynchronize(d.stream);
    batch_release(d.batch);
    d.batch = nullptr;
    for (void *p : d.allocs) (void)hipFree(p);
    d.allocs.clear();
    d.alloc_bytes.clear();
    if (d.cs_host) (void)hipHostFree(d.cs_host);
    if (d.cs_spec_host) (void)hipHostFree(d.cs_spec_host);
    if (d.scal_host) (void)hipHostFree(d.scal_host);
    for (auto &e : d.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : d.prof_pool) (void)hipEventDestroy(e);
    if (d.own_stream) (void)hipStreamDestroy(d.stream);
    delete s.dev;
    s.dev = nullptr;
}

void invalidate_geometry(System &s) {
    s.dev_geometry_dirty = true;
    if (s.dev) {
        s.dev->tpack_valid = false;
        s.dev->lu_valid = false;
        s.dev->fmm_valid = false;
    }
}

void invalidate_precond(System &s) {
    if (s.dev) s.dev->nb.valid = false;
}

void fmm_settings_changed(System &s) {
    if (s.dev) s.dev->fmm_valid = false;
    s.m2d_done = false;
    s.ipd_done = false;
}

// prepare_polelec (mod_electrostatics.F90:1487-1596), cached until the geometry changes.
void prepare_polelec(System &s) {
    if (s.m2d_done) return;
    DeviceState &d = ensure_device(s);
    if (s.npol > 0 && s.fmm_requested) {
        ensure_fmm(s, d);
        fmm_launch_static_fields(d.stream, d.fmm, s.amoeba, s.ld_cart, d.fmm_rhs_part);
        launch_rhs_finalize(d.stream, s.amoeba, s.npol, kFmmNearSplits + 1, d.fmm_rhs_part, d.ms64, d.ma64, d.mq,
                            d.ps64, d.pa64, d.sp, d.e_m2d);
    } else if (s.npol > 0 && rhs_pair_once(s.npol)) {
        ensure_sym(s, d);
        const int pi = prof_begin(d, 4);
        launch_rhs_sym(d.stream, s.amoeba, s.nmm, s.npol, d.sym_pairs, d.ms64, d.ma64, d.mq, d.ps64, d.pa64,
                       d.polar_mm, d.np_src, d.sp, d.sympart, d.rhs_part, d.rhs_split_np, d.e_m2d,
                       pi >= 0 ? d.prof_pool[2 * pi] : nullptr, pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr);
    } else if (s.npol > 0) {
        const int pi = prof_begin(d, 4);
        launch_rhs_m2d(d.stream, s.amoeba, s.nmm, s.npol, d.ms64, d.ma64, d.mq, d.ps64, d.pa64, d.polar_mm, d.sp,
                       d.rhs_part, d.rhs_split, d.e_m2d, pi >= 0 ? d.prof_pool[2 * pi] : nullptr,
                       pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr);
    }
    s.m2d_done = true;
    s.host_m2d_stale = true;
}

void refresh_host_m2d(System &s) {
    prepare_polelec(s);
    if (!s.host_m2d_stale) return;
    DeviceState &d = *s.dev;
    if (s.npol > 0)
        HCHECK(hipMemcpyAsync(s.e_m2d.data(), d.e_m2d, sizeof(double) * s.n_ipd * d.n, hipMemcpyDeviceToHost,
                              d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    s.host_m2d_stale = false;
}

// polarization() (mod_polarization.F90:49-313).  d_field: [n_ipd][3 npol] on the device.
void polarization(System &s, const double *d_field, int solver, int matv, int mask) {
    if (s.npol == 0) return;
    DeviceState &d = ensure_device(s);
    const long n = d.n;
    if (solver == kSolverNone) solver = s.def_solver;
    if (matv == kMatvNone) matv = s.def_matv;
    if (s.n_ipd == 1) mask = 1;
    const bool mD = mask & 1, mP = s.n_ipd > 1 && (mask & 2);
    if (solver != kSolverCG && solver != kSolverDIIS && solver != kSolverInversion)
        fatal("Unknown solver for calculation of the induced point dipoles");
    if (solver != kSolverInversion && matv != kMatvIncore && matv != kMatvDirect)
        fatal("Unknown matrix-vector method requested");
    // with fast multipoles every matvec is the FP64 multipole one (the mixed
    // refinement is for the direct pair sweep)
    const bool use_fmm = s.fmm_requested && solver != kSolverInversion;
    const bool mixed = s.precision == 1 && !use_fmm;
    const int kind = use_fmm ? kFmm
                             : (matv == kMatvIncore ? (mixed ? kPack32 : kPack64) : (mixed ? kOtf32 : kOtf64));
    // ipd0 = 0, then the guess for the requested sets
    HCHECK(hipMemsetAsync(d.vx, 0, sizeof(double) * 2 * n, d.stream));
    const bool guess = s.ipd_use_guess && solver != kSolverInversion;
    if (guess) {
        if (mD) launch_copy(d.stream, n, d.ipd, d.vx);
        if (mP) launch_copy(d.stream, n, d.ipd + n, d.vx + n);
    }
    launch_copy(d.stream, s.n_ipd * n, d_field, d.vb);
    HCHECK(hipEventRecord(d.ev[1], d.stream));
    for (int k = 0; k < 4; k++) s.last_iters[k] = 0;
    s.last_status = 0;
    if (solver == kSolverCG && mixed) {
        // FP32 T matvec + FP64 refinement; both sets at once when both are requested
        const int M = (mD && mP) ? 2 : 1;
        const int c0 = mD ? 0 : 1;
        MixedResult mr = mixed_solve(s, d, M, d.vb + c0 * n, d.vx + c0 * n, !guess, s.tol, s.max_iter,
                                     matv == kMatvIncore ? kPack32 : kOtf32);
        s.last_iters[2] = mr.outer;
        s.last_iters[3] = mr.inner;
        s.last_iters[0] = s.last_iters[1] = mr.inner;
        if (!mr.ok) fatal("Iterative solver did not converged");
    } else if (solver == kSolverCG) {
        for (int k = 0; k < s.n_ipd; k++) {
            if (!(k == 0 ? mD : mP)) continue;
            // P guess = converged D when there is no previous guess (mod_polarization.F90:241-243)
            if (k == 1 && mD && !s.ipd_use_guess) launch_copy(d.stream, n, d.vx, d.vx + n);
            CgState *cs_save = d.cs;
            d.cs = cs_save + k;
            double tol = s.tol;
            CgResult r = cg_solve(s, d, kind, 1, d.vb + k * n, d.vx + k * n, &tol, s.max_iter, false, kEpsRp,
                                  s.ipd_use_guess ? -1 : k);
            d.cs = cs_save;
            s.last_iters[k] = r.iters[0];
            if (r.status[0] == 1) s.last_status = 1;
            // mod_solvers.F90:212-214
            if (r.status[0] == 2 && std::fabs(r.beta[0]) > kEpsRp) fatal("Iterative solver did not converged");
        }
    } else if (solver == kSolverDIIS) {
        for (int k = 0; k < s.n_ipd; k++) {
            if (!(k == 0 ? mD : mP)) continue;
            if (k == 1 && mD && !s.ipd_use_guess) launch_copy(d.stream, n, d.vx, d.vx + n);
            int it = 0;
            // DIIS runs the FP64 matvec of the requested kind
            const int dk = use_fmm ? kFmm : (matv == kMatvIncore ? kPack64 : kOtf64);
            int bad = diis_solve(s, d, dk, d.vb + k * n, d.vx + k * n, s.tol, s.max_iter, &it);
            s.last_iters[k] = it;
            if (bad) fatal("Iterative solver did not converged");
        }
    } else {
        // inversion_solver: dense T, LU with partial pivoting, triangular solves
        if (n > 40000) fatal("The inversion solver is limited to 13333 polarizable sites; use CG or DIIS");
        if (!d.lu_valid) {
            dfree(d, d.tdense);
            dfree(d, d.piv);
            dfree(d, d.lu_work);
            d.tdense = dalloc<double>(d, (size_t)n * n);
            d.piv = dalloc<int>(d, n);
            d.lu_work = dalloc<double>(d, 8);
            launch_tdense_build(d.stream, s.amoeba, s.npol, d.ps64, d.pa64, d.alpha, d.pd.ri, d.pd.ci, d.pd.f,
                                d.tdense);
            int info = launch_lu_factor(d.stream, n, d.tdense, d.piv, d.lu_work);
            if (info != 0) fatal("Polarization matrix is singular (LU factorization failed)");
            d.lu_valid = true;
        }
        for (int k = 0; k < s.n_ipd; k++) {
            if (!(k == 0 ? mD : mP)) continue;
            launch_copy(d.stream, n, d.vb + k * n, d.vx + k * n);
            launch_lu_apply(d.stream, n, d.tdense, d.piv, d.vx + k * n);
        }
    }
    HCHECK(hipEventRecord(d.ev[2], d.stream));
    // eel%ipd = ipd0: unrequested sets become zero
    launch_copy(d.stream, s.n_ipd * n, d.vx, d.ipd);
    s.ipd_done = true;
    s.ipd_use_guess = true;
    s.host_ipd_stale = true;
}

double polelec_energy(System &s) {
    if (s.npol == 0) return 0.0;
    if (!s.ipd_done) {
        prepare_polelec(s);
        polarization(s, s.dev->e_m2d, kSolverNone, kMatvNone, 3);
    }
    DeviceState &d = *s.dev;
    const long n = d.n;
    // -1/2 mu_D . E_M2D(P) for AMOEBA, -1/2 mu . E_M2D for AMBER
    launch_energy(d.stream, n, d.ipd, d.e_m2d + (s.amoeba ? n : 0), d.red_a, d.scal);
    HCHECK(hipMemcpyAsync(d.scal_host, d.scal, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    return d.scal_host[0];
}

// energy_MM_MM (mod_electrostatics.F90:899-941) through prepare_fixedelec /
// elec_prop_M2M (:1407-1485, 1678-1952): one FP64 sweep over all ordered MM
// pairs, the S-S screened pairs corrected in the finalize, cached until the
// geometry or the scale factors change.  Always the exact direct sums (also
// when fast multipoles are enabled for the solver).
double fixedelec_energy(System &s) {
    if (s.nmm == 0) return 0.0;
    if (s.m2m_done) return s.e_mm;
    DeviceState &d = ensure_device(s);
    if (!d.mm_part) {
        hipDeviceProp_t prop;
        HCHECK(hipGetDeviceProperties(&prop, d.device));
        const int tb = site_blocks(s.nmm), ncu = prop.multiProcessorCount;
        const int max_s = std::max(1, (s.nmm + kBlock - 1) / kBlock);
        d.mm_split = std::min(max_s, std::max(balanced_split(tb, s.nmm, ncu), (6 * ncu + tb - 1) / tb));
        d.mm_part = dalloc<double>(d, (size_t)d.mm_split * s.nmm);
        d.mm_red = dalloc<double>(d, tb);
        d.mm_out = dalloc<double>(d, 1);
        d.mm_bad = dalloc<int>(d, 1);
    }
    HCHECK(hipMemsetAsync(d.mm_bad, 0, sizeof(int), d.stream));
    launch_mm_energy(d.stream, s.amoeba, s.nmm, d.ms64, d.mq, d.ss, d.mm_part, d.mm_split, d.mm_red, d.mm_bad,
                     d.mm_out);
    HCHECK(hipMemcpyAsync(d.scal_host, d.mm_out, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipMemcpyAsync(d.scal_host + 1, d.mm_bad, sizeof(int), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    int bad = 0;
    std::memcpy(&bad, d.scal_host + 1, sizeof(int));
    if (bad)
        fatal("Requesting Coulomb kernel for two atoms placed in the same point, this could be an internal bug "
              "or a problem in your input file, please check.");
    s.e_mm = d.scal_host[0];
    s.m2m_done = true;
    return s.e_mm;
}

void refresh_host_ipd(System &s) {
    if (!s.host_ipd_stale || !s.dev) return;
    DeviceState &d = *s.dev;
    HCHECK(hipMemcpyAsync(s.ipd.data(), d.ipd, sizeof(double) * s.n_ipd * d.n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    s.host_ipd_stale = false;
}

// Device bytes ensure_device() allocates for a system (packed T, dense T and
// the batched multi-RHS state come on top, on demand).
size_t device_footprint_estimate(const System &s) {
    const size_t np = s.npol, nm = s.nmm, n = 3 * np;
    size_t b = np * (sizeof(double4) + sizeof(float4) + 3 * sizeof(double) + sizeof(float) + sizeof(int));
    b += nm * (sizeof(double4) + sizeof(double) + sizeof(double) * s.ld_cart + sizeof(int));
    b += sizeof(double) * (2 * n * 2 + n + 8 * 2 * n);     // e_m2d, ipd, ext, work vectors
    b += 3 * 24 * sizeof(double) * (n / kBlock + 16);      // reduction scratch
    const int nsplit = tmv_split(s.npol);
    b += tmv_part_bytes(s.npol, 2, 64, nsplit) + sizeof(double) * n * 16;  // pair and static-field partials
    if (tmv_pair_once(s.npol, 64, 1) || rhs_pair_once(s.npol)) {  // block-pair partials and list of the pair-once sweeps
        const size_t nb = tmv_sym_blocks(s.npol);
        b += tmv_sym_part_bytes(s.npol, 2, 64) + sizeof(int2) * nb * (nb + 1) / 2;
    }
    b += (size_t)(s.list_S_P_P.nnz() + s.list_S_P_D.nnz() + 2 * s.list_P_P.nnz() + s.list_P_P_dense.nnz()) * 24;
    return b + (1 << 20);
}

DeviceView device_view(System &s) {
    DeviceState &d = ensure_device(s);
    prepare_polelec(s);
    HCHECK(hipStreamSynchronize(d.stream));
    DeviceView v;
    v.device = d.device;
    v.stream = d.stream;
    v.npol = s.npol;
    v.amoeba = s.amoeba ? 1 : 0;
    v.n_ipd = s.n_ipd;
    v.ps64 = d.ps64;
    v.pa64 = d.pa64;
    v.ps32 = d.ps32;
    v.pa32 = d.pa32;
    v.alpha = d.alpha;
    v.pp = d.pp;
    v.e_m2d = d.e_m2d;
    v.ipd = d.ipd;
    if (s.inner_precond == 1 && s.npol > 0) {
        ensure_nb(s, d);
        HCHECK(hipStreamSynchronize(d.stream));
        v.nb_row_ptr = d.nb.row_ptr;
        v.nb_ci = d.nb.ci;
        v.nb_b = d.nb.b;
    }
    v.bytes = 0;
    for (size_t b : d.alloc_bytes) v.bytes += b;
    return v;
}

// Several external fields of one system solved together (no reference
// counterpart: the reference solves one field per polarization() call).  Each
// right-hand side is the system's E_M2D plus its own field, D and P sets as two
// columns; all columns share every launch of the batched mixed solver.  The
// guess of every column is the system's current dipoles when a previous solve
// left them (ipd_use_guess), else zero.  The system's own dipoles and energy
// state are left as they were.
void solve_multi(System &s, int nrhs, const double *ext, bool ext_dev, double *ipd_out, bool out_dev,
                 double *epol) {
    if (nrhs < 1) fatal("The number of right-hand sides must be positive");
    const long n = 3L * s.npol;
    if (s.npol == 0) {
        if (epol)
            for (int k = 0; k < nrhs; k++) epol[k] = 0.0;
        return;
    }
    DeviceView v = device_view(s);
    DeviceState &d = *s.dev;
    const int nsets = s.n_ipd, ncols = nrhs * nsets;
    if (d.batch && (int)d.batch->col_owner.size() != ncols) batch_release(d.batch);
    if (!d.batch) {
        d.batch = new BatchEngine(d.stream, s.amoeba, ncols > 2, 1, ncols);
        if (d.batch->add(v, nrhs) < 0) {
            batch_release(d.batch);
            fatal("Cannot allocate device memory for the multi-RHS solve");
        }
    }
    BatchEngine &E = *d.batch;
    SegInfo &S = E.segs[0];
    HCHECK(hipEventRecord(d.ev[0], d.stream));
    HCHECK(hipMemcpyAsync(S.ext, ext, sizeof(double) * nrhs * n,
                          ext_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, d.stream));
    E.set_rhs(0);
    const bool guess = s.ipd_use_guess;
    if (guess) {
        for (int r = 0; r < nrhs; r++)
            HCHECK(hipMemcpyAsync(S.dev.x + (long)r * nsets * n, d.ipd, sizeof(double) * nsets * n,
                                  hipMemcpyDeviceToDevice, d.stream));
        // every column starts from the same dipoles x0, so the first residuals
        // share T x0: r = (E_M2D - T x0) + ext, one FP64 matvec per dipole set
        // instead of one per column
        matvec(s, d, kOtf64, nsets, d.ipd, d.vh, true, nullptr, nullptr);
        launch_residual_rms(d.stream, n, nsets, d.e_m2d, d.vh, d.vr2, d.alpha, d.red_a, d.scal);
        launch_batch_rhs(d.stream, n, nsets, ncols, d.vr2, S.ext, S.dev.r);
    } else {
        HCHECK(hipMemsetAsync(S.dev.x, 0, sizeof(double) * ncols * n, d.stream));
    }
    E.start(0, s.tol, s.max_iter, guess ? kResKnown : kResXZero);
    const long max_ticks = (long)(kBatchMaxOuter + 1) * (s.max_iter + 2);
    long prev = -1;
    const BCol *snap = nullptr;
    for (long t = 0;; t++) {
        E.tick(refine_eta(), refine_frac());
        const long seq = E.enqueue_readback();
        if (prev >= 0) {
            snap = E.wait_readback(prev);
            bool all = true;
            for (int c = 0; c < ncols && all; c++) all = snap[S.col0 + c].s[0].phase == kPhaseDone;
            if (all) break;
        }
        prev = seq;
        if (t > max_ticks) fatal("Iterative solver did not converged");
    }
    int outer = 0, inner = 0;
    for (int c = 0; c < ncols; c++) {
        const BColSub &cs = snap[S.col0 + c].s[0];
        if (cs.status != 0) fatal("Iterative solver did not converged");
        outer = std::max(outer, cs.outer + 1);
        inner = std::max(inner, cs.inner_total);
    }
    s.last_iters[0] = s.last_iters[1] = inner;
    s.last_iters[2] = outer;
    s.last_iters[3] = inner;
    s.last_status = 0;
    HCHECK(hipEventRecord(d.ev[1], d.stream));
    E.energies(0);
    if (ipd_out)
        HCHECK(hipMemcpyAsync(ipd_out, S.dev.x, sizeof(double) * ncols * n,
                              out_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, d.stream));
    if (epol) HCHECK(hipMemcpyAsync(epol, S.eout, sizeof(double) * nrhs, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipEventRecord(d.ev[3], d.stream));
    HCHECK(hipEventSynchronize(d.ev[3]));
    float ms = 0.0f;
    s.last_ms[0] = 0.0;
    (void)hipEventElapsedTime(&ms, d.ev[0], d.ev[1]);
    s.last_ms[1] = ms;
    (void)hipEventElapsedTime(&ms, d.ev[1], d.ev[3]);
    s.last_ms[2] = ms;
}

// E_n2p (mod_qm_helper.F90:376-391) of nq charges [nq][4] at the polarizable sites.
void charges_field_host(System &s, int nq, const double *charges, double *out) {
    if (s.npol == 0) return;
    if (nq < 0) fatal("The number of charges must not be negative");
    DeviceState &d = ensure_device(s);
    double4 *dq = nullptr;
    double *dout = d.vr2;
    if (nq > 0) HCHECK(hipMalloc((void **)&dq, sizeof(double4) * nq));
    if (nq > 0) HCHECK(hipMemcpyAsync(dq, charges, sizeof(double4) * nq, hipMemcpyHostToDevice, d.stream));
    launch_charges_field(d.stream, s.npol, d.ps64, nq, dq, dout);
    HCHECK(hipMemcpyAsync(out, dout, sizeof(double) * d.n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    if (dq) HCHECK(hipFree(dq));
}

// potential_M2E / potential_D2E / field_M2E / field_D2E at n external points
// (mod_electrostatics.F90:3067-3336), added to V[n] and / or E[n][3] on the host
// as the reference adds to the caller's arrays.
void x2e_host(System &s, int n, const double *cext, bool mm, bool pol, double *V, double *E) {
    if (n <= 0) return;
    if (n > (1 << 28)) fatal("Too many external points");
    pol = pol && s.npol > 0;
    if (pol && !s.ipd_done)
        fatal(V ? "IPD should be computed before D2E potential." : "IPD should be computed before computing D2E field.");
    if (!mm && !pol) return;
    DeviceState &d = ensure_device(s);
    const int smm = mm ? x2e_splits(n, s.nmm) : 0, spol = pol ? x2e_splits(n, s.npol) : 0;
    const long nparts = (long)(smm + spol) * n;
    if (n > d.x2e_cap) {
        dfree(d, d.x2e_c);
        dfree(d, d.x2e_out);
        d.x2e_c = dalloc<double>(d, 3L * n);
        d.x2e_out = dalloc<double4>(d, n);
        d.x2e_cap = n;
    }
    if (nparts > d.x2e_cap_part) {
        dfree(d, d.x2e_part);
        d.x2e_part = dalloc<double4>(d, nparts);
        d.x2e_cap_part = nparts;
    }
    if (!d.x2e_bad) d.x2e_bad = dalloc<int>(d, 1);
    std::vector<double> out(4L * n);
    int bad = 0;
    HCHECK(hipMemsetAsync(d.x2e_bad, 0, sizeof(int), d.stream));
    HCHECK(hipMemcpyAsync(d.x2e_c, cext, sizeof(double) * 3L * n, hipMemcpyHostToDevice, d.stream));
    launch_x2e(d.stream, n, d.x2e_c, mm, pol, s.amoeba, s.ld_cart, s.nmm, d.ms64, d.mq, s.npol, d.ps64, d.ipd,
               d.x2e_part, smm, spol, d.x2e_out, d.x2e_bad);
    HCHECK(hipMemcpyAsync(out.data(), d.x2e_out, sizeof(double) * 4L * n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipMemcpyAsync(&bad, d.x2e_bad, sizeof(int), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    if (bad)
        fatal("Requesting Coulomb kernel for two atoms placed in the same point, this could be an internal bug "
              "or a problem in your input file, please check.");
    for (long j = 0; j < n; j++) {
        if (V) V[j] += out[4 * j];
        if (E)
            for (int c = 0; c < 3; c++) E[3 * j + c] += out[4 * j + 1 + c];
    }
}

// ommp_field_mmpol2ext without any device allocation: the system's device
// state must already exist (a solve or another device call made it); the
// points, the partial sums and the result live in a static per-device scratch
// array (x2e_kernels.hip), so at most x2e_noalloc_max_points() points per call.
void x2e_field_noalloc_host(System &s, int n, const double *cext, double *E) {
    if (n <= 0) return;
    if (n > x2e_noalloc_max_points())
        fatal("ommp_amd_field_mmpol2ext_noalloc takes at most " + std::to_string(x2e_noalloc_max_points()) +
              " external points per call (" + std::to_string(n) + " requested)");
    const bool pol = s.npol > 0;
    if (pol && !s.ipd_done) fatal("IPD should be computed before computing D2E field.");
    if (!s.dev)
        fatal("ommp_amd_field_mmpol2ext_noalloc needs the system's device state: call it after a solve (or "
              "another device call) on this system");
    DeviceState &d = ensure_device(s);  // exists: at most re-uploads coordinates, no allocation
    const int bad = x2e_field_noalloc(d.stream, d.device, n, cext, true, pol, s.amoeba, s.ld_cart, s.nmm, d.ms64,
                                      d.mq, s.npol, d.ps64, d.ipd, E);
    if (bad)
        fatal("Requesting Coulomb kernel for two atoms placed in the same point, this could be an internal bug "
              "or a problem in your input file, please check.");
}

namespace {

// Index lists and buffers of the gradient, built on first use.  The MM-target
// correction list is the transpose of the S->P one (dropped with it when the
// lists change); frames, unrotated multipoles and the per-atom rotation index
// do not depend on the coordinates.
void ensure_grad(System &s, DeviceState &d) {
    if (!d.mp.ri) {
        const auto rows = sp_rows(s);
        std::vector<std::vector<std::pair<int, double2>>> mrows(s.nmm);
        for (int jp = 0; jp < s.npol; jp++)
            for (const auto &e : rows[jp]) mrows[e.first].push_back({jp, e.second});
        upload_csr(d, s.nmm, mrows, d.mp, true);
    }
    if (d.grad_static_ready) return;
    const int nmm = s.nmm, npol = s.npol;
    d.mm_polar = dalloc<int>(d, nmm);
    upload(s.mm_polar.data(), d.mm_polar, sizeof(int) * nmm);
    std::vector<int> rri(nmm + 1, 0), rci;
    if (s.amoeba) {
        d.q0 = dalloc<double>(d, 10L * nmm);
        upload(s.q0.data(), d.q0, sizeof(double) * 10 * nmm);
        std::vector<int4> fr(nmm);
        std::vector<std::vector<int>> users(nmm);
        for (int j = 0; j < nmm; j++) {
            fr[j] = make_int4(s.mol_frame[j], s.iz[j], s.ix[j], s.iy[j]);
            if (s.mol_frame[j] == 0) continue;
            if (s.iz[j] >= nmm || s.ix[j] >= nmm || s.iy[j] >= nmm || s.mol_frame[j] < 0 || s.mol_frame[j] > 5)
                fatal("Invalid multipole frame definition");
            const int at[4] = {j, s.ix[j] >= 0 ? s.ix[j] : j, s.iy[j] >= 0 ? s.iy[j] : j,
                               s.iz[j] >= 0 ? s.iz[j] : j};
            for (int r = 0; r < 4; r++) users[at[r]].push_back(4 * j + r);
        }
        for (int a = 0; a < nmm; a++) {
            rri[a + 1] = rri[a] + (int)users[a].size();
            rci.insert(rci.end(), users[a].begin(), users[a].end());
        }
        d.frames = dalloc<int4>(d, nmm);
        upload(fr.data(), d.frames, sizeof(int4) * nmm);
        d.g_rotc = dalloc<double>(d, 12L * nmm);
        d.g_egrd = dalloc<double>(d, 6L * nmm);
    }
    d.rev_ri = dalloc<int>(d, nmm + 1);
    d.rev_ci = dalloc<int>(d, std::max<size_t>(rci.size(), 1));
    upload(rri.data(), d.rev_ri, sizeof(int) * (nmm + 1));
    upload(rci.data(), d.rev_ci, sizeof(int) * rci.size());
    int cu = 0;
    HCHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d.device));
    cu = std::max(cu, 1);
    d.g_split_s = balanced_split(site_blocks(npol), nmm, cu);
    d.g_split_p = balanced_split(site_blocks(npol), npol, cu);
    d.g_split_m = balanced_split(site_blocks(nmm), npol, cu);
    // partials of the per-target sweeps, or of the both-ends static sweep
    // (one per block of the other end) and the dipole-dipole part
    const bool both = grad_both_ends(npol, nmm, s.amoeba);
    const size_t nstat = both ? site_blocks(nmm) : d.g_split_s, nmmp = both ? site_blocks(npol) : d.g_split_m;
    d.g_part_pol = dalloc<double>(d, (nstat + d.g_split_p) * 3 * npol);
    d.g_part_mm = dalloc<double>(d, nmmp * (s.amoeba ? 12 : 3) * nmm);
    d.g_pol = dalloc<double>(d, 3L * npol);
    d.g_tr = dalloc<double>(d, 3L * nmm);
    d.g_emm = dalloc<double>(d, 3L * nmm);
    d.g_out = dalloc<double>(d, 3L * nmm);
    d.grad_static_ready = true;
}

GradArgs grad_args(const System &s, DeviceState &d) {
    GradArgs a;
    a.nmm = s.nmm;
    a.npol = s.npol;
    a.amoeba = s.amoeba;
    a.ms = d.ms64;
    a.ps = d.ps64;
    a.ma = d.ma64;
    a.mq = d.mq;
    a.pa = d.pa64;
    a.ipd = d.ipd;
    a.q0 = d.q0;
    a.polar_mm = d.polar_mm;
    a.mm_polar = d.mm_polar;
    a.rev_ri = d.rev_ri;
    a.rev_ci = d.rev_ci;
    a.frames = d.frames;
    a.sp = d.sp;
    a.pp = d.pp;
    a.mp = d.mp;
    a.part_pol = d.g_part_pol;
    a.part_mm = d.g_part_mm;
    a.split_s = d.g_split_s;
    a.split_p = d.g_split_p;
    a.split_m = d.g_split_m;
    a.gpol = d.g_pol;
    a.gtr = d.g_tr;
    a.emm = d.g_emm;
    a.egrd = d.g_egrd;
    a.rotc = d.g_rotc;
    a.out = d.g_out;
    return a;
}

}  // namespace

// ommp_polelec_geomgrad (mod_interface.F90:671-683) -> polelec_geomgrad
// (mod_geomgrad.F90:113-273): grd[nmm][3] = 0, then d E_pol / d r when the
// system has polarizable sites; dipoles are solved first (default solver, no
// external field) if they are not current.
void polelec_geomgrad_host(System &s, double *grd) {
    std::fill(grd, grd + 3L * s.nmm, 0.0);
    if (s.npol == 0) return;
    if (!s.ipd_done) {
        prepare_polelec(s);
        polarization(s, s.dev->e_m2d, kSolverNone, kMatvNone, 3);
    }
    DeviceState &d = ensure_device(s);
    ensure_grad(s, d);
    GradArgs a = grad_args(s, d);
    const int pi = prof_begin(d, 5);
    hipEvent_t evb = pi >= 0 ? d.prof_pool[2 * pi] : nullptr, eve = pi >= 0 ? d.prof_pool[2 * pi + 1] : nullptr;
    if (s.fmm_requested) {
        // pair sums by fast multipoles (near field + far field), then the same
        // finalize, rotation and assembly as the direct sweeps
        ensure_fmm(s, d);
        if (!d.fmm_gpart_pol) {
            d.fmm_gpart_pol = fmm_alloc<double>(d, (size_t)fmm_grad_parts_pol(s.amoeba) * 3 * s.npol);
            d.fmm_gpart_mm = fmm_alloc<double>(d, (size_t)fmm_grad_parts_mm() * (s.amoeba ? 12 : 3) * s.nmm);
        }
        a.part_pol = d.fmm_gpart_pol;
        a.part_mm = d.fmm_gpart_mm;
        if (evb) HCHECK(hipEventRecord(evb, d.stream));
        fmm_launch_grad_fields(d.stream, d.fmm, s.amoeba, s.ld_cart, d.ipd, a.part_pol, a.part_mm);
        if (eve) HCHECK(hipEventRecord(eve, d.stream));
        launch_geomgrad_finish(d.stream, a, fmm_grad_parts_pol(s.amoeba), fmm_grad_parts_mm());
    } else {
        if (rhs_pair_once(s.npol)) {  // the dipole-dipole part pair-once, as the static field
            ensure_sym(s, d);
            a.sym_pairs = d.sym_pairs;
            a.sympart = static_cast<double *>(d.sympart);
        }
        a.fused = grad_both_ends(s.npol, s.nmm, s.amoeba);
        launch_geomgrad(d.stream, a, evb, eve);
    }
    HCHECK(hipMemcpyAsync(grd, d.g_out, sizeof(double) * 3L * s.nmm, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
}

// ommp_rotation_geomgrad (mod_interface.F90:699-710) -> rotation_geomgrad
// (rotate_multipoles.F90:3-230): grd = 0, then the gradient of the rotated
// multipoles' energy in the field E[nmm][3] and field gradient Egrd[nmm][6]
// through their frame atoms.  Without AMOEBA multipoles there is nothing to
// rotate and grd stays 0.
void rotation_geomgrad_host(System &s, const double *E, const double *Egrd, double *grd) {
    std::fill(grd, grd + 3L * s.nmm, 0.0);
    if (!s.amoeba || s.nmm == 0) return;
    DeviceState &d = ensure_device(s);
    ensure_grad(s, d);
    HCHECK(hipMemcpyAsync(d.g_emm, E, sizeof(double) * 3L * s.nmm, hipMemcpyHostToDevice, d.stream));
    HCHECK(hipMemcpyAsync(d.g_egrd, Egrd, sizeof(double) * 6L * s.nmm, hipMemcpyHostToDevice, d.stream));
    launch_rotation_grad(d.stream, grad_args(s, d));
    HCHECK(hipMemcpyAsync(grd, d.g_out, sizeof(double) * 3L * s.nmm, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
}

const double *device_ipd(System &s) {
    DeviceState &d = ensure_device(s);
    return d.ipd;
}

double tmatvec_host(System &s, int kind, int ncols, const double *x, double *y, int repeat) {
    if (ncols < 1 || ncols > 2) fatal("tmatvec supports 1 or 2 columns");
    if (kind < 0 || kind > 8) fatal("Unknown matvec kind");
    if (s.npol == 0) return 0.0;
    DeviceState &d = ensure_device(s);
    const long n = d.n;
    HCHECK(hipMemcpyAsync(d.vd, x, sizeof(double) * ncols * n, hipMemcpyHostToDevice, d.stream));
    matvec(s, d, kind, ncols, d.vd, d.vr2, true, nullptr, nullptr);  // builds packed T if needed
    if (repeat < 1) repeat = 1;
    HCHECK(hipEventRecord(d.ev[0], d.stream));
    for (int r = 0; r < repeat; r++) matvec(s, d, kind, ncols, d.vd, d.vr2, true, nullptr, nullptr);
    HCHECK(hipEventRecord(d.ev[1], d.stream));
    HCHECK(hipMemcpyAsync(y, d.vr2, sizeof(double) * ncols * n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, d.ev[0], d.ev[1]);
    return ms / repeat;
}

// y = T x for ncols host vectors through the inner matvec of the multi-RHS
// batch (bf16 matrix cores, three-part split), with the diagonal; the columns
// are grouped by dipole set as in a batch, so ncols need not be a multiple of
// it.  Returns the mean device time per launch of repeat launches (ms).
double block_tmatvec_host(System &s, int ncols, const double *x, double *y, int repeat) {
    if (ncols < 1 || ncols > kMfmaCols) fatal("The block matvec takes 1 to 128 columns");
    if (s.npol == 0) return 0.0;
    DeviceView v = device_view(s);
    DeviceState &d = *s.dev;
    const long n = d.n;
    const int nsets = s.n_ipd, nrhs = (ncols + nsets - 1) / nsets;
    BatchEngine E(d.stream, s.amoeba, true, 1, nrhs * nsets);
    if (E.add(v, nrhs) < 0) fatal("Cannot allocate device memory for the block matvec");
    const SegInfo &S = E.segs[0];
    std::vector<BCol> cols(nrhs * nsets);  // value-initialized: phase DONE
    for (int c = 0; c < ncols; c++) cols[c].s[1].phase = kPhaseInner;
    HCHECK(hipMemcpyAsync(E.d_cols, cols.data(), sizeof(BCol) * cols.size(), hipMemcpyHostToDevice, d.stream));
    HCHECK(hipMemcpyAsync(S.dev.p, x, sizeof(double) * ncols * n, hipMemcpyHostToDevice, d.stream));
    E.inner_matvec();
    if (repeat < 1) repeat = 1;
    HCHECK(hipEventRecord(d.ev[0], d.stream));
    for (int r = 0; r < repeat; r++) E.inner_matvec();
    HCHECK(hipEventRecord(d.ev[1], d.stream));
    HCHECK(hipMemcpyAsync(y, S.dev.h, sizeof(double) * ncols * n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, d.ev[0], d.ev[1]);
    return ms / repeat;
}

// z = M^-1 r of the inner solves' near-field preconditioner for ncols host
// vectors (the kernel the inner solves start with); returns 1 when it is in
// use, 0 when it is switched off for this geometry (then z = alpha r).
int inner_precond_apply_host(System &s, int ncols, const double *r, double *z) {
    if (ncols < 1 || ncols > 2) fatal("The preconditioner applies to 1 or 2 columns");
    if (s.npol == 0) return 0;
    DeviceState &d = ensure_device(s);
    ensure_nb(s, d);
    const long n = d.n;
    const double te[4] = {0.0, 0.0, 0.0, 0.0};
    HCHECK(hipMemcpyAsync(d.tol_dev, te, sizeof(te), hipMemcpyHostToDevice, d.stream));
    HCHECK(hipMemcpyAsync(d.vr2, r, sizeof(double) * ncols * n, hipMemcpyHostToDevice, d.stream));
    launch_cg_init_nb(d.stream, s.npol, ncols, d.vr2, d.vd, d.vr, d.vz, d.vp, d.alpha, d.nb, d.red_a, d.cs,
                      d.tol_dev, nullptr);
    int off = 0;
    HCHECK(hipMemcpyAsync(z, d.vz, sizeof(double) * ncols * n, hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipMemcpyAsync(&off, d.nb.off_flag, sizeof(int), hipMemcpyDeviceToHost, d.stream));
    HCHECK(hipStreamSynchronize(d.stream));
    return off ? 0 : 1;
}

void profile_enable(System &s, int every) {
    DeviceState &d = ensure_device(s);
    if (every <= 0) prof_flush(d);
    d.prof = every > 0;
    d.prof_every = every > 0 ? every : 1;
    for (long &q : d.prof_seq) q = 0;
}

void profile_stats(System &s, double *out, bool reset) {
    DeviceState &d = ensure_device(s);
    prof_flush(d);
    for (int k = 0; k < 6; k++) {
        out[2 * k] = (double)d.prof_count[k];
        out[2 * k + 1] = d.prof_ms[k];
        if (reset) {
            d.prof_count[k] = 0;
            d.prof_ms[k] = 0.0;
        }
    }
}

// ommp_set_external_field[_nomm] (mod_interface.F90:184-233)
void set_external_field(System &s, const double *ext, bool ext_on_device, int solver, int matv, bool nomm) {
    s.ipd_done = false;
    if (s.npol == 0) return;
    DeviceState &d = ensure_device(s);
    const long n = d.n;
    HCHECK(hipEventRecord(d.ev[0], d.stream));
    if (ext_on_device) launch_copy(d.stream, n, ext, d.ext);
    else HCHECK(hipMemcpyAsync(d.ext, ext, sizeof(double) * n, hipMemcpyHostToDevice, d.stream));
    double *ef = d.vr;  // staging: [n_ipd][n]
    if (!nomm) {
        prepare_polelec(s);
        launch_add_field(d.stream, n, s.n_ipd, d.e_m2d, d.ext, ef);
        // the MM-field branch never forwards matv (mod_interface.F90:221)
        polarization(s, ef, solver, kMatvNone, 3);
    } else {
        launch_add_field(d.stream, n, 1, nullptr, d.ext, ef);
        if (s.n_ipd > 1) HCHECK(hipMemsetAsync(ef + n, 0, sizeof(double) * n, d.stream));
        polarization(s, ef, solver, matv, 1);
    }
    HCHECK(hipEventRecord(d.ev[3], d.stream));
    if (!ext_on_device) refresh_host_ipd(s);
    HCHECK(hipEventSynchronize(d.ev[3]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, d.ev[0], d.ev[1]);
    s.last_ms[0] = ms;
    (void)hipEventElapsedTime(&ms, d.ev[1], d.ev[2]);
    s.last_ms[1] = ms;
    (void)hipEventElapsedTime(&ms, d.ev[2], d.ev[3]);
    s.last_ms[2] = ms;
}

}  // namespace ommp
