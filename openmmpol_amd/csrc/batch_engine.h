// Host side of the batched solver: segment memory, work lists, ticks and the
// column-state readback ring.  Used by the multi-RHS entry point (one segment)
// and by the serving loop (one segment per admitted system).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "batch.h"

namespace ommp {

constexpr int kMaxCharges = 4096;  // point charges held per segment (all steps of a request)

struct SegInfo {
    bool used = false;
    int nrhs = 0, nsets = 1, ncols = 0, col0 = 0;  // ncols = nrhs * nsets + extra
    size_t bytes = 0;
    void *mem = nullptr;
    BSeg dev{};              // the descriptor uploaded to the device
    DeviceView view{};
    double *ext = nullptr;   // [nrhs][3 npol] external field
    double *ered = nullptr;  // [nrhs][vblocks] energy partials
    double *eout = nullptr;  // [nrhs] energies
    double4 *charges = nullptr;  // [kMaxCharges] (x, y, z, q)
};

class BatchEngine {
  public:
    BatchEngine(hipStream_t stream, bool amoeba, bool wide, int max_segs, int max_cols);
    ~BatchEngine();
    BatchEngine(const BatchEngine &) = delete;
    BatchEngine &operator=(const BatchEngine &) = delete;

    // Segment for a system with nrhs right-hand sides (n_ipd columns each)
    // plus `extra` work columns after them.  Returns the slot, or -1 when the
    // segment or column capacity is full.
    int add(const DeviceView &v, int nrhs, int extra = 0);
    void remove(int seg);
    // Bytes a segment for this system would take (for admission decisions).
    static size_t segment_bytes(int npol, int nrhs, int nsets, bool wide, int num_cu, int extra = 0);

    // Uploads descriptors and work lists when the composition changed.
    void commit();
    void tick(double eta, double frac);
    // h = T p of the inner phase alone (wide engines only; see launch_batch_inner_matvec).
    void inner_matvec();
    // Copies the column states to the next pinned snapshot; returns its number.
    long enqueue_readback();
    // Waits for snapshot number seq and returns its column array.
    const BCol *wait_readback(long seq);
    // Number the next enqueue_readback will return: a column started now is
    // first visible in that snapshot.
    long next_readback() const { return snap_seq_; }

    // b = E_M2D + ext for every column of the segment.
    void set_rhs(int seg);
    // Same, keeping r = b - T x of the last solve in step (see launch_batch_shift_rhs).
    void shift_rhs(int seg);
    // Starts every right-hand-side column of the segment (phase OUTER);
    // rinit as launch_batch_start.
    void start(int seg, double tol, int maxit, int rinit);
    // Starts columns [c, c + nc) of the segment.
    void start_cols(int seg, int c, int nc, double tol, int maxit, int rinit, int oneshot);
    // Energies of the segment's right-hand sides into seg.eout (device).
    void energies(int seg);

    hipStream_t stream;
    bool amoeba;
    int m32, m64;
    int num_cu = 256;
    std::vector<SegInfo> segs;
    std::vector<int> col_owner;  // segment per column slot, -1 = free
    BCol *d_cols = nullptr;
    size_t bytes_allocated() const { return pool_bytes_; }

  private:
    void *pool_get(size_t bytes);
    void pool_put(void *p, size_t bytes);
    void ensure_capacity(int4 *&buf, int &cap, int need);

    BSeg *d_segs_ = nullptr;
    int *d_act_ = nullptr, *d_actn_ = nullptr;  // active-column lists (BatchTables::act)
    int4 *d_pair32_ = nullptr, *d_fin32_ = nullptr, *d_pair64_ = nullptr, *d_fin64_ = nullptr, *d_vec_ = nullptr;
    int4 *d_sym32_ = nullptr, *d_sym64_ = nullptr, *d_pair64m_ = nullptr;
    int cap_pair32_ = 0, cap_fin32_ = 0, cap_pair64_ = 0, cap_fin64_ = 0, cap_vec_ = 0;
    int cap_sym32_ = 0, cap_sym64_ = 0, cap_pair64m_ = 0;
    BatchTables tables_{};
    bool dirty_ = true;
    // pinned staging for table uploads (reused only after its copy finished)
    void *stage_ = nullptr;
    size_t stage_bytes_ = 0;
    hipEvent_t stage_ev_ = nullptr;
    bool stage_pending_ = false;
    // column-state snapshots
    BCol *snap_[2] = {nullptr, nullptr};
    hipEvent_t snap_ev_[2] = {nullptr, nullptr};
    long snap_seq_ = 0;
    // segment memory pool
    std::vector<std::pair<void *, size_t>> free_;
    std::vector<void *> owned_;
    size_t pool_bytes_ = 0;
};

void batch_release(BatchEngine *&e);

}  // namespace ommp
