// Batched solver: many right-hand-side columns, of one system or of several
// systems of different sizes, solved together by the mixed-precision scheme
// (FP32 T matvec inside FP64 iterative refinement) with every launch covering
// every column of the batch.
//
// A column is one dipole set (D or P) of one solve.  Columns of a system are
// stored [column][3 * npol] in that system's segment.  Each column carries its
// own CG scalars and a phase:
//   OUTER  next tick computes r = b - T x in FP64 and tests convergence;
//   INNER  next tick runs one FP32-matvec PCG iteration on T d = r;
//   DONE   converged (or stopped), skipped by every kernel.
// One tick = [FP64 matvec, residual, decide] for OUTER columns followed by
// [FP32 matvec, update, direction] for INNER columns.  A column's arithmetic
// depends only on its own data and on its segment's fixed launch geometry, so
// its result is the same whatever else is in the batch.
//
// Phase and scalars are double-buffered: the two kernels that change them
// (decide and direction) read copy 0 / write copy 1 and read copy 1 / write
// copy 0, so no block ever sees a value written by another block of the same
// launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device.h"

namespace ommp {

constexpr int kPhaseDone = 0, kPhaseOuter = 1, kPhaseInner = 2;
constexpr int kBatchMaxOuter = 30;
// Where the first FP64 residual of a solve comes from:
constexpr int kResMatvec = 0;  // r = b - T x (FP64 matvec)
constexpr int kResXZero = 1;   // x = 0: r = b
constexpr int kResKnown = 2;   // r already holds b - T x (same T, x, updated b)

struct BColSub {
    double g;        // inner r.z
    double tin;      // inner target rms
    double rms;      // last outer (FP64) preconditioned residual rms
    int phase;
    int it;          // iteration within the current inner solve
    int outer;       // outer steps taken
    int inner_total; // inner iterations summed over outer steps
    int status;      // 0 converged, 2 not converged
    int pad;
};

struct BCol {
    BColSub s[2];
    double tol;
    int maxit;       // inner iterations per refinement step
    int rinit;       // first residual: kResMatvec, kResXZero or kResKnown
    int oneshot;     // stop after the first inner solve, x = its correction (no FP64 check)
    int nb_off;      // 1: this column's inner solves use alpha alone (set on a breakdown)
};

// One system in a batch; device pointers only.
struct BSeg {
    int npol, ncols, nsplit, chunk;
    int tblocks;     // ceil(npol / kBlock): pair/finalize target blocks
    int vblocks;     // ceil(3 npol / kBlock): vector blocks per column
    int col0;        // first batch column of this segment
    int amoeba;
    const float4 *ps32;
    const float *pa32;
    const double4 *ps64;
    const double *pa64;
    const double *alpha;
    const int *pp_ri;
    const int *pp_ci;
    const double *pp_f;
    double *b, *x, *r, *z, *p, *h, *d;  // [ncols][3 npol]
    // matrix-core path: p of the inner solves split in kBxParts bf16 parts,
    // [part][ncols][ldx], ldx = 3 * (npol rounded up to kBxSrc); the elements
    // past 3 npol are zero
    unsigned short *xs;
    int ldx;
    // partial sums of the matvec: per-target sweep [nsplit][ncols][3 npol]; pair-once
    // sweep (sym32 / sym64 set for that precision, tmv_sym.h) [block pair][ncols][3][256]
    // target sides, then as many source sides; FP64-sized
    void *part;
    int sym32, sym64;
    double *red_hp;                      // [ncols][tblocks]
    double *red_rz;                      // [ncols][vblocks]: r.z
    double *red_ra;                      // [ncols][vblocks]: r.(alpha r), the convergence measure
    // near-field preconditioner of the inner solves (the system's lists and
    // tensors, precond.hip); nb_row_ptr == nullptr: z = alpha r
    const int *nb_row_ptr;
    const int *nb_ci;
    const float *nb_b;
    double *r2;                          // [ncols][3 npol]: the inner residual alternates r, r2
};

// Work lists (built on the host when the batch composition changes):
//   pair items     int4 {segment, target block, source split, column group}
//   pair-once items int4 {segment, I | J << 16, block pair index, column group}
//   finalize items int4 {segment, target block, column group, 0}
//   vector items   int4 {segment, column in segment, vector block, batch column}
// In a wide batch the FP64 pair items' column group g is the g-th run of
// kBatchM64 entries of the segment's active-column list (built by a small
// kernel before those launches), so that the columns still refining share as
// few groups as possible; otherwise group g is columns [g M, g M + M).
struct BatchTables {
    const BSeg *segs;
    BCol *cols;
    const int4 *pair32;
    int npair32;
    const int4 *fin32;
    int nfin32;
    const int4 *pair64;
    int npair64;
    const int4 *fin64;
    int nfin64;
    const int4 *sym32;
    int nsym32;
    const int4 *sym64;
    int nsym64;
    // wide batches: FP64 matrix-core items {segment, 32-site target tile, split,
    // group of kD64Cols entries of the active-column list}
    const int4 *pair64m;
    int npair64m;
    const int4 *vec;
    int nvec;
    // the active columns of each segment, in increasing order, when the FP64
    // pair kernels of a wide batch pack their column groups:
    // act[col0 + k] = k-th active column of the segment, act_n[segment] = count
    int *act;
    int *act_n;
    int nseg;
};

constexpr int kBatchM32 = 8;  // columns per FP32 finalize work item (multi-RHS)
// Multi-RHS inner matvec on the bf16 matrix cores: work item = 32 target sites x
// 128 columns x one source split (a multiple of kBxSrc sites); T tiles of 32 x 32
// sites built in LDS.
constexpr int kMfmaSites = 32;
constexpr int kMfmaCols = 128;
constexpr int kBxSrc = 32;
constexpr int kBxParts = 3;  // bf16 parts per FP32 value (8 significant bits each)
inline int bx_ld(int npol) { return 3 * ((npol + kBxSrc - 1) / kBxSrc) * kBxSrc; }
constexpr int kBatchM64 = 4;  // columns per FP64 pair work item (multi-RHS)
constexpr int kD64Cols = 64;  // refining columns per FP64 matrix-core work item (multi-RHS)

// The inner matvec of the wide (matrix-core) batch alone: h = T p for every
// column whose inner phase is set (a diagnostic of its accuracy and speed).
void launch_batch_inner_matvec(hipStream_t st, const BatchTables &t, bool amoeba);

// Column-group width used for a segment of ncols columns.
inline int batch_m32(int ncols) { return ncols <= 2 ? 2 : kBatchM32; }
inline int batch_m64(int ncols) { return ncols <= 2 ? 2 : kBatchM64; }

// One tick over every column of the batch (see the file comment).
// m32/m64: column-group widths of this batch (all segments share them).
// Inner solves stop at max(eta * outer rms, frac * tol).
void launch_batch_tick(hipStream_t st, const BatchTables &t, bool amoeba, int m32, int m64, double eta,
                       double frac);

// b[col] = E_M2D[set of col] + ext[rhs of col] for columns [c0, c0 + ncols) of a
// segment whose columns are ordered (rhs, set); ext [nrhs][3 npol] on device.
void launch_batch_rhs(hipStream_t st, long n, int nsets, int ncols, const double *e_m2d, const double *ext,
                      double *b);
// Field of point charges (x, y, z, q) at the polarizable sites, undamped and
// unscreened (E_n2p, mod_qm_helper.F90:376-391); ext[3 npol].
void launch_charges_field(hipStream_t st, int npol, const double4 *ps64, int nq, const double4 *charges,
                          double *ext);
// Start (or restart) columns [c0, c0 + nc) of the batch column array: phase
// OUTER, counters cleared, tolerance set; rinit says how the first residual
// is formed (kResMatvec, kResXZero, kResKnown).
void launch_batch_start(hipStream_t st, BCol *cols, int c0, int nc, double tol, int maxit, int rinit,
                        int oneshot = 0);
// Shared correction of a warm SCF step (only the external field changed, by
// the same amount for every dipole set): ext_new in bs on entry; on exit
// bs = ext_new - ext, ext = ext_new, b[set] = E_M2D[set] + ext_new.
void launch_batch_shared_rhs(hipStream_t st, long n, int nsets, const double *e_m2d, double *ext, double *b,
                             double *bs);
// x[set] += xs for every set (the shared correction applied to each set).
void launch_batch_feed(hipStream_t st, long n, int nsets, double *x, const double *xs);
// New right-hand sides for columns whose x and FP64 residual r = b - T x are
// those of their last solve (same geometry): b' = E_M2D + ext, r += b' - b,
// b = b'.  The next solve then starts without an FP64 matvec (kResKnown).
void launch_batch_shift_rhs(hipStream_t st, long n, int nsets, int ncols, const double *e_m2d, const double *ext,
                            double *b, double *r);
// E_pol per right-hand side, -1/2 mu_D . E_M2D,P (AMOEBA) or -1/2 mu . E_M2D:
// x columns ordered (rhs, set); red scratch [nrhs][vblocks]; out[nrhs] on device.
void launch_batch_energy(hipStream_t st, long n, int nsets, int nrhs, const double *x, const double *e_m2d,
                         double *red, double *out);

// Device arrays of one system that the batched engine reads (engine.cpp).
struct DeviceView {
    int device;
    hipStream_t stream;
    int npol, amoeba, n_ipd;
    const double4 *ps64;
    const double *pa64;
    const float4 *ps32;
    const float *pa32;
    const double *alpha;
    CsrDev pp;
    const double *e_m2d;
    double *ipd;
    size_t bytes;    // device memory held by the system
    // near-field preconditioner (nullptr: the system's inner solves use alpha)
    const int *nb_row_ptr = nullptr;
    const int *nb_ci = nullptr;
    const float *nb_b = nullptr;
};

}  // namespace ommp
