// HBM layout and kernel launchers of the MI355X induced-dipole path.
//
// Layout (per system, all resident in HBM for the lifetime of the system):
//   pol sites   ps64[npol]  double4 {x, y, z, 1/thole}    ps32: float4 copy
//               pa64[npol]  Thole damping a                pa32: float copy
//               alpha[npol] polarizability (FP64)
//   MM sources  ms64[nmm]   double4 {x, y, z, 1/thole} (1/thole = +inf when not
//               polarizable: the damping test then fails and the kernel is the
//               bare Coulomb one, as in the reference when thole_i*thole_j < eps)
//               ma64[nmm]   damping a (+inf when not polarizable, so that
//               min(a_i, a_j) picks the polarizable end)
//               mq[nmm][ld] lab-frame multipoles
//   correction lists, rows = polarizable target (CSR, 0-based):
//               sp_*: sources whose D or P scale differs from 1, f = 1 - s
//               pp_*: P-P pairs of the on-the-fly matvec, f = 1 - s
//   vectors     [column][3*npol] FP64, column = D (0) / P (1)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace ommp {

// Errors of the HIP runtime end the process the way every other error of the
// library does: "[OMMP]   [stop] ..." and exit(1) (io.cpp, the reference's
// fatal_error, src/mod_io.F90:155-175).
[[noreturn]] void fatal(const std::string &msg);
[[noreturn]] void hip_fatal(hipError_t e, const char *what, const char *file, int line);
#define HIP_CHECK(x)                                                 \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) ::ommp::hip_fatal(e_, #x, __FILE__, __LINE__); \
    } while (0)

constexpr int kBlock = 256;      // threads per block for every kernel here
constexpr int kSymBlock = 256;   // sites per block of the pair-once matvec (= kBlock)
constexpr int kMaxCols = 2;      // D and P sets
constexpr double kEpsRpDev = 2.220446049250313080847e-16 * 100.0;

struct CgState {
    double g[2];        // r.z, double-buffered by iteration parity
    double rms;
    double last_beta;
    double tol;
    double eps;         // |p.Tp| below this ends the solve (reference: eps_rp)
    int done;
    int status;         // 0 converged, 1 small step (|p.Tp| < eps), 2 running/not converged
    int iters;
    int pad;
};

struct CsrDev {
    int *ri = nullptr;
    int *ci = nullptr;
    double *f = nullptr;   // one value per entry (pp lists)
    double2 *f2 = nullptr; // two values per entry (sp list: x = D set, y = P set)
    int nnz = 0;
};

// ----------------------------------------------------------------- launchers
// Optional events evb / eve are recorded on the stream right before and after
// the pair (or packed-stream) kernel only, for live per-kernel timing.
// Static field of the multipoles at the polarizable sites (E_M2D), FP64.
void launch_rhs_m2d(hipStream_t st, bool amoeba, int nmm, int npol, const double4 *ms64,
                    const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                    const int *polar_mm, const CsrDev &sp, double *part, int nsplit, double *e_m2d,
                    hipEvent_t evb = nullptr, hipEvent_t eve = nullptr);

// The same one-sided over the MM sources src[0..nsrc) (the non-polarizable
// ones), without finalize: part[s][3 * npol] for the nsplit source splits.
void launch_rhs_pairs_list(hipStream_t st, bool amoeba, int nsrc, const int *src, int npol, const double4 *ms64,
                           const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                           double *part, int nsplit);
// Static field with each unordered pair of polarizable sites evaluated once
// (rhs_sym.hip): pairs[] and sympart as for launch_tmv_sym; the nmm - npol
// non-polarizable sources np_src are swept one-sided into part_np (nsplit_np
// splits of 3 * npol).  rhs_pair_once() says from which size it is used.
bool rhs_pair_once(int npol);
void launch_rhs_sym(hipStream_t st, bool amoeba, int nmm, int npol, const int2 *pairs, const double4 *ms64,
                    const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                    const int *polar_mm, const int *np_src, const CsrDev &sp, void *sympart, double *part_np,
                    int nsplit_np, double *e_m2d, hipEvent_t evb = nullptr, hipEvent_t eve = nullptr);

// T matvec on the fly: y[m] = T x[m] (dodiag) or (T - diag) x[m], m < M.
// prec 64 or 32 selects the pair arithmetic; corrections and the final sum are FP64.
// If dotpart != nullptr, per-block partial sums of y.x go to dotpart[m * nblk + b]
// where nblk = ceil(npol / kBlock).
void launch_tmv_otf(hipStream_t st, int prec, bool amoeba, int npol, int M, const double4 *ps64,
                    const double *pa64, const float4 *ps32, const float *pa32, const double *alpha,
                    const CsrDev &pp, const double *x, long xstride, double *y, long ystride,
                    bool dodiag, double *dotpart, void *part, int nsplit, const int *skip_flag,
                    hipEvent_t evb = nullptr, hipEvent_t eve = nullptr);
// Finalize alone, for FP64 partial fields computed elsewhere (fast multipoles).
void launch_tmv_finalize64(hipStream_t st, bool amoeba, int npol, int M, const double *part, int nsplit,
                           const double4 *ps64, const double *pa64, const double *alpha, const CsrDev &pp,
                           const double *x, long xstride, double *y, long ystride, bool dodiag, double *dotpart,
                           const int *skip_flag);
// T matvec on the fly with each unordered pair evaluated once (tmv_sym.hip):
// pairs[] = the block pairs (I <= J) of kSymBlock-site blocks, row-major;
// part holds tmv_sym_part_bytes(npol, M, prec).
int tmv_sym_blocks(int npol);
// Per-column y.x partials written by the pair-once matvec's finalize.
int tmv_sym_dot_parts(int npol);
size_t tmv_sym_part_bytes(int npol, int M, int prec);
// Whether the on-the-fly matvec of prec (32 / 64) with ncols columns uses the
// pair-once sweep at npol sites (its partial sums at most max_bytes).
bool tmv_pair_once(int npol, int prec, int ncols, size_t max_bytes = size_t(2) << 30);
void launch_tmv_sym(hipStream_t st, int prec, bool amoeba, int npol, int M, const int2 *pairs, const double4 *ps64,
                    const double *pa64, const float4 *ps32, const float *pa32, const double *alpha,
                    const CsrDev &pp, const double *x, long xstride, double *y, long ystride, bool dodiag,
                    double *dotpart, void *part, const int *skip_flag, hipEvent_t evb = nullptr,
                    hipEvent_t eve = nullptr);
void launch_rhs_finalize(hipStream_t st, bool amoeba, int npol, int nsplit, const double *part, const double4 *ms64,
                         const double *ma64, const double *mq, const double4 *ps64, const double *pa64,
                         const CsrDev &sp, double *e_m2d);
int tmv_split(int npol);
int balanced_split(int tb, int nsrc, int num_cu);
size_t tmv_part_bytes(int npol, int M, int prec, int nsplit);
// FMA rate (TFLOP/s, FMA = 2) of non-packed vector code on the current device,
// prec 32 or 64: eight independent chains per thread, 8,192 x 256 threads.
double vector_fma_tflops(int prec);
int vec_blocks(long n);     // blocks used by the element-wise kernels
int site_blocks(int npol);  // blocks used by the per-site kernels

// Packed symmetric in-core T (upper tile pairs, 64x64 sites per tile).
size_t tpack_bytes(int npol, int prec);
void launch_tpack_build(hipStream_t st, int prec, bool amoeba, int npol, const double4 *ps64,
                        const double *pa64, void *tpack, const int *pd_ri, const int *pd_ci,
                        const double *pd_s);
void launch_tpack_mv(hipStream_t st, int prec, int npol, int M, const void *tpack, const double *alpha,
                     const double *x, long xstride, double *y, long ystride, bool dodiag,
                     double *dotpart, void *part, const int *skip_flag, hipEvent_t evb = nullptr,
                     hipEvent_t eve = nullptr);
size_t tpack_part_bytes(int npol, int M);

// Potential and field at n external points cext[n][3] of the multipoles (mm)
// and/or of the induced dipoles (pol), FP64: out[n] = (V, Ex, Ey, Ez); part
// holds (smm + spol) * n double4; *bad is set if a point sits on a source.
int x2e_splits(int n, int nsrc);
void launch_x2e(hipStream_t st, int n, const double *cext, bool mm, bool pol, bool amoeba, int ld, int nmm,
                const double4 *ms, const double *mq, int npol, const double4 *ps, const double *ipd, double4 *part,
                int smm, int spol, double4 *out, int *bad);

// The same for the field only, with no allocation: the points cext[n][3] and
// the field e[n][3] (host, e += result) are staged in a static per-device
// scratch array (n <= x2e_noalloc_max_points()).  Returns the coincident-point flag.
int x2e_noalloc_max_points();
size_t x2e_noalloc_scratch_bytes();
int x2e_field_noalloc(hipStream_t st, int device, int n, const double *cext, bool mm, bool pol, bool amoeba, int ld,
                      int nmm, const double4 *ms, const double *mq, int npol, const double4 *ps, const double *ipd,
                      double *e);

// Geometric gradient of the polarization energy.  part_pol holds
// (split_s + split_p) * 3 * npol doubles, part_mm split_m * 12 * nmm (AMOEBA) or
// split_m * 3 * nmm; mp: correction list per MM target (polarizable sources,
// f = 1 - s for the D and P sets, self pairs excluded); frames[nmm] = (convention,
// z, x, y atom; 0-based, -1 = none); rev_ri/rev_ci list, per atom, the entries
// 4 * site + role (0 self, 1 x, 2 y, 3 z) whose frame uses that atom.
struct GradArgs {
    int nmm = 0, npol = 0;
    bool amoeba = false;
    const double4 *ms = nullptr, *ps = nullptr;
    const double *ma = nullptr, *mq = nullptr, *pa = nullptr, *ipd = nullptr, *q0 = nullptr;
    const int *polar_mm = nullptr, *mm_polar = nullptr, *rev_ri = nullptr, *rev_ci = nullptr;
    const int4 *frames = nullptr;
    CsrDev sp, pp, mp;
    double *part_pol = nullptr, *part_mm = nullptr;
    int split_s = 1, split_p = 1, split_m = 1;
    double *gpol = nullptr, *gtr = nullptr, *emm = nullptr, *egrd = nullptr, *rotc = nullptr, *out = nullptr;
    // set: the dipole-dipole part runs pair-once over these block pairs of
    // polarizable sites, its block-pair partials in sympart
    const int2 *sym_pairs = nullptr;
    double *sympart = nullptr;
    // set (grad_both_ends): the static part evaluates each (polarizable, MM)
    // pair once for both ends; part_pol then holds ceil(nmm / kBlock) static
    // partials and part_mm ceil(npol / kBlock)
    bool fused = false;
};
bool grad_both_ends(int npol, int nmm, bool amoeba);
// out[nmm][3] = d E_pol / d r (pair sweeps + rotation term); emm/egrd receive
// E_D2M and Egrd_D2M.  evb/eve bracket the three pair sweeps.
void launch_geomgrad(hipStream_t st, const GradArgs &a, hipEvent_t evb = nullptr, hipEvent_t eve = nullptr);
// The same after the pair sums: finalize (sum of nparts partials in the layouts
// above, screened-pair corrections), rotation term and per-atom assembly, for
// partials computed elsewhere (fast multipoles).
// direct: the MM partials carry the translational gradient (k_gstat_both).
void launch_geomgrad_finish(hipStream_t st, const GradArgs &a, int nparts_pol, int nparts_mm, bool direct = false);
// out[nmm][3] = rotation term only, for the field a.emm and gradient a.egrd.
void launch_rotation_grad(hipStream_t st, const GradArgs &a);

// CG pieces (columns m < M, vectors [m * n + k]).
void launch_norm2_guess(hipStream_t st, long n, int M, double *x, const double *alpha, double *red);
void launch_cg_init(hipStream_t st, long n, int M, const double *b, const double *h, double *x,
                    double *r, double *z, double *p, const double *alpha, double *red, CgState *cs,
                    bool zero_guess, const double *tol_eps /* [tol0, tol1, eps0, eps1] */,
                    const CgState *pre = nullptr /* columns with pre[m].done start done */);
// Second half of launch_cg_init alone: CG state from nred per-block r.z partials.
void launch_cg_init2(hipStream_t st, int M, const double *red, int nred, CgState *cs, const double *tol,
                     const CgState *pre);

// Near-field preconditioner of the inner FP32 PCG (precond.hip): neighbour
// lists within kNbCut by target (row_ptr, ci; rowi = target of each entry),
// B_ij = alpha_i alpha_j s_ij (k3 d d^T - k2 I) as 6 floats (xx, xy, xz, yy, yz,
// zz) per entry; off_flag = 1 switches it off (set by the direction kernel on a
// breakdown, cleared when B is recomputed).
constexpr double kNbCut = 3.0 * 1.8897261245650;  // 3 Angstrom in bohr
struct NbDev {
    int *row_ptr = nullptr, *ci = nullptr, *rowi = nullptr;
    float *b = nullptr;
    int *off_flag = nullptr;
    long nnz = 0, cap = 0;
    bool list_valid = false;  // lists of the current coordinates uploaded
    bool valid = false;       // B computed
};
int nb_blocks(int npol);  // per-column partial sums of the preconditioned CG kernels
void launch_nb_tensors(hipStream_t st, bool amoeba, const NbDev &nb, const float4 *ps32, const float *pa32,
                       const double *alpha, const CsrDev &pp);
// Preconditioned CG pieces (columns m < M, vectors [m * 3 npol + k]); the
// per-block partials have nb_blocks(npol) entries per column.
void launch_cg_init_nb(hipStream_t st, int npol, int M, const double *b, double *x, double *r, double *z, double *p,
                       const double *alpha, const NbDev &nb, double *red, CgState *cs, const double *tol,
                       const CgState *pre);
void launch_cg_update_nb(hipStream_t st, int npol, int M, int it, const double *h, double *x, const double *rold,
                         double *rnew, double *z, const double *p, const double *alpha, const double *red_hp,
                         int nred_hp, double *red_rz, double *red_ra, CgState *cs, const NbDev &nb);
void launch_cg_direction_nb(hipStream_t st, int npol, int M, int it, const double *z, double *p,
                            const double *red_rz, const double *red_ra, CgState *cs, const NbDev &nb);

// Mixed-precision refinement step decided on the device (k_refine_decide).
void launch_refine_decide(hipStream_t st, int M, const double *rms, CgState *out, double tol, double eta,
                          double frac, double *inner_tol, int outer);
void launch_cg_update(hipStream_t st, long n, int M, int it, const double *h, double *x, double *r,
                      double *z, const double *p, const double *alpha, const double *red_hp,
                      int nred_hp, double *red_rz, CgState *cs);
void launch_cg_direction(hipStream_t st, long n, int M, int it, const double *z, double *p,
                         const double *red_rz, CgState *cs, int *all_done);

// Interaction energy of the static multipoles (energy_MM_MM), FP64: part holds
// nsplit * nmm doubles, red site_blocks(nmm); ss = S-S list by target with
// f = 1 - s; *bad is set if two sites coincide; *out = the energy.
void launch_mm_energy(hipStream_t st, bool amoeba, int nmm, const double4 *ms64, const double *mq, const CsrDev &ss,
                      double *part, int nsplit, double *red, int *bad, double *out);

// Element-wise helpers.
void launch_add_field(hipStream_t st, long n, int nsets, const double *e_m2d, const double *ext,
                      double *ef);
void launch_copy(hipStream_t st, long n, const double *src, double *dst);
void launch_axpy(hipStream_t st, long n, int M, double a, const double *x, double *y);
void launch_residual_rms(hipStream_t st, long n, int M, const double *b, const double *tx,
                         double *r, const double *alpha, double *red, double *rms_out);
void launch_dot(hipStream_t st, long n, const double *a, const double *b, double *red, double *out);
void launch_energy(hipStream_t st, long n, const double *ipd_d, const double *e_p, double *red,
                   double *out);

// Jacobi / DIIS pieces.
void launch_jacobi(hipStream_t st, long n, const double *rhs, const double *ox, const double *x,
                   const double *alpha, double *xnew, double *err);
void launch_multi_dot(hipStream_t st, long n, int k, const double *const *vecs, const double *e,
                      double *red, double *out);
void launch_lincomb(hipStream_t st, long n, int k, const double *const *vecs, const double *coef,
                    double *out);
void launch_maxabs_diff(hipStream_t st, long n, const double *a, const double *b, double *red,
                        double *out);

// Dense T and LU for the inversion solver.
void launch_tdense_build(hipStream_t st, bool amoeba, int npol, const double4 *ps64,
                         const double *pa64, const double *alpha, const int *pd_ri,
                         const int *pd_ci, const double *pd_s, double *T);
int launch_lu_factor(hipStream_t st, long n, double *A, int *piv, double *work);
void launch_lu_apply(hipStream_t st, long n, const double *A, const int *piv, double *rhs);

}  // namespace ommp
