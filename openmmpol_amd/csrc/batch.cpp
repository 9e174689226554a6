// BatchEngine: host bookkeeping of the batched solver (batch_engine.h).
#include "batch_engine.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>

#include "ommp_internal.h"

namespace ommp {

#define BCHECK(x) HIP_CHECK(x)

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct SegLayout {
    int nsplit, chunk, tblocks, vblocks, sym32, sym64;
    size_t vec, xs, part, red_hp, red_rz, ext, ered, eout, charges, total;
};

SegLayout seg_layout(int npol, int nrhs, int nsets, int m32, int num_cu, int extra) {
    SegLayout L;
    const long n = 3L * npol;
    const int ncols = nrhs * nsets + extra;
    L.tblocks = site_blocks(npol);
    L.vblocks = vec_blocks(n);
    if (m32 == kBatchM32) {
        // matrix-core FP32 matvec: 32-site target tiles x 64-column groups
        const int items = ((npol + kMfmaSites - 1) / kMfmaSites) * ((ncols + kMfmaCols - 1) / kMfmaCols);
        L.nsplit = balanced_split(items, npol, num_cu);
    } else {
        const int groups = (ncols + m32 - 1) / m32;
        L.nsplit = balanced_split(L.tblocks * groups, npol, num_cu);
    }
    L.chunk = (npol + L.nsplit - 1) / L.nsplit;
    if (m32 == kBatchM32) {
        // the matrix-core tiles read 32 sources at a time from 16-byte aligned rows
        L.chunk = (L.chunk + kBxSrc - 1) / kBxSrc * kBxSrc;
        L.nsplit = (npol + L.chunk - 1) / L.chunk;
    }
    L.vec = align256(sizeof(double) * ncols * n);
    L.xs = m32 == kBatchM32 ? align256(sizeof(unsigned short) * kBxParts * ncols * (size_t)bx_ld(npol)) : 0;
    // the pair-once sweep at the sizes where the single-system solver uses it;
    // wide batches run both matvecs on the matrix cores with their own tiling
    // (partials per source split)
    L.sym32 = m32 == 2 && tmv_pair_once(npol, 32, ncols, size_t(8) << 30);
    L.sym64 = m32 == 2 && tmv_pair_once(npol, 64, ncols, size_t(8) << 30);
    size_t part = sizeof(double) * L.nsplit * ncols * n;
    if (L.sym32 || L.sym64) part = std::max(part, tmv_sym_part_bytes(npol, ncols, 64));
    L.part = align256(part);
    L.red_hp = align256(sizeof(double) * ncols * L.tblocks);
    L.red_rz = align256(sizeof(double) * ncols * L.vblocks);
    L.ext = align256(sizeof(double) * nrhs * n);
    L.ered = align256(sizeof(double) * nrhs * L.vblocks);
    L.eout = align256(sizeof(double) * nrhs);
    L.charges = align256(sizeof(double4) * kMaxCharges);
    // 8 vectors (b, x, r, z, p, h, d, r2), two kinds of r partials
    L.total = 8 * L.vec + L.xs + L.part + L.red_hp + 2 * L.red_rz + L.ext + L.ered + L.eout + L.charges;
    return L;
}

}  // namespace

BatchEngine::BatchEngine(hipStream_t st, bool amo, bool wide, int max_segs, int max_cols)
    : stream(st), amoeba(amo), m32(wide ? kBatchM32 : 2), m64(wide ? kBatchM64 : 2) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        num_cu = prop.multiProcessorCount;
    if (num_cu <= 0) num_cu = 256;
    segs.resize(max_segs);
    col_owner.assign(max_cols, -1);
    BCHECK(hipMalloc((void **)&d_cols, sizeof(BCol) * std::max(max_cols, 1)));
    BCHECK(hipMemsetAsync(d_cols, 0, sizeof(BCol) * std::max(max_cols, 1), stream));
    BCHECK(hipMalloc((void **)&d_segs_, sizeof(BSeg) * std::max(max_segs, 1)));
    BCHECK(hipMalloc((void **)&d_act_, sizeof(int) * std::max(max_cols, 1)));
    BCHECK(hipMalloc((void **)&d_actn_, sizeof(int) * std::max(max_segs, 1)));
    for (int k = 0; k < 2; k++) {
        BCHECK(hipHostMalloc((void **)&snap_[k], sizeof(BCol) * std::max(max_cols, 1), hipHostMallocDefault));
        BCHECK(hipEventCreateWithFlags(&snap_ev_[k], hipEventDisableTiming));
    }
    BCHECK(hipEventCreateWithFlags(&stage_ev_, hipEventDisableTiming));
    tables_.segs = d_segs_;
    tables_.cols = d_cols;
    tables_.act = d_act_;
    tables_.act_n = d_actn_;
}

BatchEngine::~BatchEngine() {
    (void)hipStreamSynchronize(stream);
    for (void *p : owned_) (void)hipFree(p);
    (void)hipFree(d_cols);
    (void)hipFree(d_segs_);
    (void)hipFree(d_act_);
    (void)hipFree(d_actn_);
    int4 *lists[] = {d_pair32_, d_fin32_, d_pair64_, d_fin64_, d_vec_, d_sym32_, d_sym64_, d_pair64m_};
    for (auto p : lists)
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; k++) {
        (void)hipHostFree(snap_[k]);
        (void)hipEventDestroy(snap_ev_[k]);
    }
    if (stage_) (void)hipHostFree(stage_);
    (void)hipEventDestroy(stage_ev_);
}

void batch_release(BatchEngine *&e) {
    delete e;
    e = nullptr;
}

size_t BatchEngine::segment_bytes(int npol, int nrhs, int nsets, bool wide, int ncu, int extra) {
    return seg_layout(npol, nrhs, nsets, wide ? kBatchM32 : 2, ncu, extra).total;
}

void *BatchEngine::pool_get(size_t bytes) {
    int best = -1;
    for (int i = 0; i < (int)free_.size(); i++)
        if (free_[i].second >= bytes && (best < 0 || free_[i].second < free_[best].second)) best = i;
    if (best >= 0) {
        void *p = free_[best].first;
        free_.erase(free_.begin() + best);
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    owned_.push_back(p);
    pool_bytes_ += bytes;
    return p;
}

void BatchEngine::pool_put(void *p, size_t bytes) { free_.push_back({p, bytes}); }

int BatchEngine::add(const DeviceView &v, int nrhs, int extra) {
    const int nsets = v.n_ipd;
    const int ncols = nrhs * nsets + extra;
    int slot = -1;
    for (int i = 0; i < (int)segs.size(); i++)
        if (!segs[i].used) {
            slot = i;
            break;
        }
    if (slot < 0 || ncols <= 0) return -1;
    // first fit of a contiguous column range
    int col0 = -1;
    for (int c = 0; c + ncols <= (int)col_owner.size(); c++) {
        bool ok = true;
        for (int k = 0; k < ncols; k++)
            if (col_owner[c + k] >= 0) {
                ok = false;
                c += k;
                break;
            }
        if (ok) {
            col0 = c;
            break;
        }
    }
    if (col0 < 0) return -1;
    const SegLayout L = seg_layout(v.npol, nrhs, nsets, m32, num_cu, extra);
    // keep the pool's block size so that it can be reused by a larger system
    char *mem = static_cast<char *>(pool_get(L.total));
    if (!mem) return -1;
    SegInfo &S = segs[slot];
    S.used = true;
    S.nrhs = nrhs;
    S.nsets = nsets;
    S.ncols = ncols;
    S.col0 = col0;
    S.bytes = L.total;
    S.mem = mem;
    S.view = v;
    BSeg &d = S.dev;
    d.npol = v.npol;
    d.ncols = ncols;
    d.nsplit = L.nsplit;
    d.chunk = L.chunk;
    d.sym32 = L.sym32;
    d.sym64 = L.sym64;
    d.tblocks = L.tblocks;
    d.vblocks = L.vblocks;
    d.col0 = col0;
    d.amoeba = v.amoeba;
    d.ps32 = v.ps32;
    d.pa32 = v.pa32;
    d.ps64 = v.ps64;
    d.pa64 = v.pa64;
    d.alpha = v.alpha;
    d.pp_ri = v.pp.ri;
    d.pp_ci = v.pp.ci;
    d.pp_f = v.pp.f;
    d.nb_row_ptr = v.nb_row_ptr;
    d.nb_ci = v.nb_ci;
    d.nb_b = v.nb_b;
    double **vecs[] = {&d.b, &d.x, &d.r, &d.z, &d.p, &d.h, &d.d, &d.r2};
    size_t off = 0;
    for (auto pv : vecs) {
        *pv = reinterpret_cast<double *>(mem + off);
        off += L.vec;
    }
    d.ldx = bx_ld(v.npol);
    d.xs = L.xs ? reinterpret_cast<unsigned short *>(mem + off) : nullptr;
    off += L.xs;
    d.part = mem + off;
    off += L.part;
    d.red_hp = reinterpret_cast<double *>(mem + off);
    off += L.red_hp;
    d.red_rz = reinterpret_cast<double *>(mem + off);
    off += L.red_rz;
    d.red_ra = reinterpret_cast<double *>(mem + off);
    off += L.red_rz;
    S.ext = reinterpret_cast<double *>(mem + off);
    off += L.ext;
    S.ered = reinterpret_cast<double *>(mem + off);
    off += L.ered;
    S.eout = reinterpret_cast<double *>(mem + off);
    off += L.eout;
    S.charges = reinterpret_cast<double4 *>(mem + off);
    for (int k = 0; k < ncols; k++) col_owner[col0 + k] = slot;
    dirty_ = true;
    return slot;
}

void BatchEngine::remove(int seg) {
    SegInfo &S = segs[seg];
    if (!S.used) return;
    for (int k = 0; k < S.ncols; k++) col_owner[S.col0 + k] = -1;
    // the columns are DONE; later launches never touch this memory, and any new
    // owner of it is written in stream order after the launches already queued
    pool_put(S.mem, S.bytes);
    S = SegInfo();
    dirty_ = true;
}

void BatchEngine::ensure_capacity(int4 *&buf, int &cap, int need) {
    if (need <= cap) return;
    // queued launches may still read the old list
    BCHECK(hipStreamSynchronize(stream));
    if (buf) BCHECK(hipFree(buf));
    cap = std::max(need, 2 * cap);
    BCHECK(hipMalloc((void **)&buf, sizeof(int4) * cap));
}

void BatchEngine::commit() {
    if (!dirty_) return;
    std::vector<int> order;
    for (int i = 0; i < (int)segs.size(); i++)
        if (segs[i].used) order.push_back(i);
    // largest systems first: their blocks start early, small ones fill the tail
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return segs[a].dev.npol > segs[b].dev.npol; });
    std::vector<int4> p32, f32, p64, f64, vec, s32, s64, p64m;
    for (int si : order) {
        const BSeg &d = segs[si].dev;
        const int g32 = (d.ncols + m32 - 1) / m32, g64 = (d.ncols + m64 - 1) / m64;
        // pair-once items: every block pair (I <= J), row-major (index order)
        auto sym_items = [&](std::vector<int4> &out, int groups) {
            for (int g = 0; g < groups; g++) {
                int idx = 0;
                for (int I = 0; I < d.tblocks; I++)
                    for (int J = I; J < d.tblocks; J++) out.push_back(make_int4(si, I | (J << 16), idx++, g));
            }
        };
        if (d.sym32) {
            sym_items(s32, g32);
        } else if (m32 == kBatchM32) {
            const int ttiles = (d.npol + kMfmaSites - 1) / kMfmaSites;
            const int gm = (d.ncols + kMfmaCols - 1) / kMfmaCols;
            for (int g = 0; g < gm; g++)
                for (int sp = 0; sp < d.nsplit; sp++)
                    for (int tt = 0; tt < ttiles; tt++) p32.push_back(make_int4(si, tt, sp, g));
        } else {
            for (int g = 0; g < g32; g++)
                for (int sp = 0; sp < d.nsplit; sp++)
                    for (int tb = 0; tb < d.tblocks; tb++) p32.push_back(make_int4(si, tb, sp, g));
        }
        for (int g = 0; g < g32; g++)
            for (int tb = 0; tb < d.tblocks; tb++) f32.push_back(make_int4(si, tb, g, 0));
        if (m32 == kBatchM32) {
            // FP64 refinement on the matrix cores, over the FP32 path's splits
            const int ttiles = (d.npol + kMfmaSites - 1) / kMfmaSites;
            const int gm = (d.ncols + kD64Cols - 1) / kD64Cols;
            for (int g = 0; g < gm; g++)
                for (int sp = 0; sp < d.nsplit; sp++)
                    for (int tt = 0; tt < ttiles; tt++) p64m.push_back(make_int4(si, tt, sp, g));
        } else if (d.sym64)
            sym_items(s64, g64);
        else
            for (int g = 0; g < g64; g++)
                for (int sp = 0; sp < d.nsplit; sp++)
                    for (int tb = 0; tb < d.tblocks; tb++) p64.push_back(make_int4(si, tb, sp, g));
        for (int g = 0; g < g64; g++)
            for (int tb = 0; tb < d.tblocks; tb++) f64.push_back(make_int4(si, tb, g, 0));
        for (int c = 0; c < d.ncols; c++)
            for (int vb = 0; vb < d.vblocks; vb++) vec.push_back(make_int4(si, c, vb, d.col0 + c));
    }
    ensure_capacity(d_pair32_, cap_pair32_, (int)p32.size());
    ensure_capacity(d_fin32_, cap_fin32_, (int)f32.size());
    ensure_capacity(d_pair64_, cap_pair64_, (int)p64.size());
    ensure_capacity(d_fin64_, cap_fin64_, (int)f64.size());
    ensure_capacity(d_vec_, cap_vec_, (int)vec.size());
    ensure_capacity(d_sym32_, cap_sym32_, (int)s32.size());
    ensure_capacity(d_sym64_, cap_sym64_, (int)s64.size());
    ensure_capacity(d_pair64m_, cap_pair64m_, (int)p64m.size());
    const size_t nseg = segs.size();
    const size_t need = sizeof(BSeg) * nseg + sizeof(int4) * (p32.size() + f32.size() + p64.size() + f64.size() +
                                                              vec.size() + s32.size() + s64.size() + p64m.size());
    if (stage_pending_) {
        BCHECK(hipEventSynchronize(stage_ev_));
        stage_pending_ = false;
    }
    if (need > stage_bytes_) {
        if (stage_) BCHECK(hipHostFree(stage_));
        stage_bytes_ = std::max(need, 2 * stage_bytes_);
        BCHECK(hipHostMalloc(&stage_, stage_bytes_, hipHostMallocDefault));
    }
    char *h = static_cast<char *>(stage_);
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes, void *dst) {
        if (bytes == 0) return;
        std::memcpy(h + off, src, bytes);
        BCHECK(hipMemcpyAsync(dst, h + off, bytes, hipMemcpyHostToDevice, stream));
        off += bytes;
    };
    std::vector<BSeg> dsegs(nseg);
    for (size_t i = 0; i < nseg; i++)
        if (segs[i].used) dsegs[i] = segs[i].dev;
    put(dsegs.data(), sizeof(BSeg) * nseg, d_segs_);
    put(p32.data(), sizeof(int4) * p32.size(), d_pair32_);
    put(f32.data(), sizeof(int4) * f32.size(), d_fin32_);
    put(p64.data(), sizeof(int4) * p64.size(), d_pair64_);
    put(f64.data(), sizeof(int4) * f64.size(), d_fin64_);
    put(vec.data(), sizeof(int4) * vec.size(), d_vec_);
    put(s32.data(), sizeof(int4) * s32.size(), d_sym32_);
    put(s64.data(), sizeof(int4) * s64.size(), d_sym64_);
    put(p64m.data(), sizeof(int4) * p64m.size(), d_pair64m_);
    BCHECK(hipEventRecord(stage_ev_, stream));
    stage_pending_ = true;
    tables_.segs = d_segs_;
    tables_.cols = d_cols;
    tables_.pair32 = d_pair32_;
    tables_.npair32 = (int)p32.size();
    tables_.fin32 = d_fin32_;
    tables_.nfin32 = (int)f32.size();
    tables_.pair64 = d_pair64_;
    tables_.npair64 = (int)p64.size();
    tables_.fin64 = d_fin64_;
    tables_.nfin64 = (int)f64.size();
    tables_.vec = d_vec_;
    tables_.nvec = (int)vec.size();
    tables_.sym32 = d_sym32_;
    tables_.nsym32 = (int)s32.size();
    tables_.sym64 = d_sym64_;
    tables_.nsym64 = (int)s64.size();
    tables_.pair64m = d_pair64m_;
    tables_.npair64m = (int)p64m.size();
    tables_.act = d_act_;
    tables_.act_n = d_actn_;
    tables_.nseg = (int)nseg;
    dirty_ = false;
}

void BatchEngine::tick(double eta, double frac) {
    commit();
    launch_batch_tick(stream, tables_, amoeba, m32, m64, eta, frac);
}

void BatchEngine::inner_matvec() {
    if (m32 != kBatchM32) fatal("The inner matvec alone runs on wide batches only");
    commit();
    launch_batch_inner_matvec(stream, tables_, amoeba);
}

long BatchEngine::enqueue_readback() {
    const long seq = snap_seq_++;
    const int k = (int)(seq & 1);
    BCHECK(hipMemcpyAsync(snap_[k], d_cols, sizeof(BCol) * col_owner.size(), hipMemcpyDeviceToHost, stream));
    BCHECK(hipEventRecord(snap_ev_[k], stream));
    return seq;
}

const BCol *BatchEngine::wait_readback(long seq) {
    const int k = (int)(seq & 1);
    BCHECK(hipEventSynchronize(snap_ev_[k]));
    return snap_[k];
}

void BatchEngine::set_rhs(int seg) {
    const SegInfo &S = segs[seg];
    launch_batch_rhs(stream, 3L * S.dev.npol, S.nsets, S.nrhs * S.nsets, S.view.e_m2d, S.ext, S.dev.b);
}

void BatchEngine::shift_rhs(int seg) {
    const SegInfo &S = segs[seg];
    launch_batch_shift_rhs(stream, 3L * S.dev.npol, S.nsets, S.nrhs * S.nsets, S.view.e_m2d, S.ext, S.dev.b,
                           S.dev.r);
}

void BatchEngine::start(int seg, double tol, int maxit, int rinit) {
    const SegInfo &S = segs[seg];
    launch_batch_start(stream, d_cols, S.col0, S.nrhs * S.nsets, tol, maxit, rinit);
}

void BatchEngine::start_cols(int seg, int c, int nc, double tol, int maxit, int rinit, int oneshot) {
    const SegInfo &S = segs[seg];
    launch_batch_start(stream, d_cols, S.col0 + c, nc, tol, maxit, rinit, oneshot);
}

void BatchEngine::energies(int seg) {
    const SegInfo &S = segs[seg];
    launch_batch_energy(stream, 3L * S.dev.npol, S.nsets, S.nrhs, S.dev.x, S.view.e_m2d, S.ered, S.eout);
}

}  // namespace ommp
