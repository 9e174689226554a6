/*
 * ommp_amd.h -- extensions of the OpenMMPol C interface for the MI355X build.
 *
 * None of these has a counterpart in the reference's C interface; each says
 * which reference routine or behaviour it exposes or extends.  They use the
 * same opaque handle as openmmpol.h and only plain C types.
 */
#ifndef _OMMP_AMD_H
#define _OMMP_AMD_H

#include <stdbool.h>
#include <stdint.h>
#include "openmmpol.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the T matvec inside the solvers.
   FP64  : every step in FP64, reference algorithm and iterate sequence.
   MIXED : FP32 T matvec (pair arithmetic) inside an FP64 iterative-refinement
           loop; the residual r = E - T mu, CG dot products and step lengths,
           the dipole accumulation and the energy sum are FP64.  D and P sets
           are solved together (two right-hand sides per matvec). */
#define OMMP_AMD_PREC_FP64 0
#define OMMP_AMD_PREC_MIXED 1

/* Solver status codes returned by ommp_amd_get_last_status. */
#define OMMP_AMD_STATUS_CONVERGED 0   /* residual rms below tolerance                 */
#define OMMP_AMD_STATUS_SMALL_STEP 1  /* |p.Tp| < eps_rp quick return (CG)            */
#define OMMP_AMD_STATUS_NOT_CONVERGED 2

/* Build a system from arrays instead of a file.  This is what the reference
   does in memory between mmpol_init and mmpol_prepare (mod_mmpol.F90:49-237)
   when a front-end has already parsed its input.
     ff_type   OMMP_FF_AMOEBA or OMMP_FF_AMBER
     ff_rules  for non-AMOEBA: 0 = Wang AL, 1 = Wang DL (scale sets and Thole
               scale of mod_constants.F90:37-58); ignored for AMOEBA
     cmm       [mm_atoms][3] coordinates in bohr
     q         [mm_atoms][ld_cart] multipoles in the .mmp convention
               (AMOEBA quadrupoles are divided by 3 and rotated here)
     pol       [mm_atoms] polarizability in bohr^3 (|pol| <= eps: not polarizable)
     thole_a   [mm_atoms] AMOEBA Thole damping factor, or NULL for 0.39
     adj_ri, adj_ci  1-2 connectivity as a 0-based CSR (adj_ri has mm_atoms+1 entries)
     pgroup    [mm_atoms] 0-based polarization group (AMOEBA), or NULL
     frames    [mm_atoms][4] (mol_frame, iz, ix, iy; 1-based atom indices, 0 = none),
               or NULL for no rotation (all frames 0) */
OMMP_SYSTEM_PRT ommp_amd_init_arrays(int32_t ff_type, int32_t ff_rules, int32_t mm_atoms,
                                     const double *cmm, const double *q, const double *pol,
                                     const double *thole_a, const int32_t *adj_ri,
                                     const int32_t *adj_ci, const int32_t *pgroup,
                                     const int32_t *frames);

/* set_screening_parameters (mod_electrostatics.F90:440-467) followed by a
   rebuild of the screening lists; intra may be NULL for non-AMOEBA. */
void ommp_amd_set_screening_parameters(OMMP_SYSTEM_PRT s, const double *m, const double *p,
                                       const double *d, const double *u, const double *intra);

/* Solver tolerance and iteration limit (the reference hard-codes 1e-8 and 200,
   mod_solvers.F90:30-32; those remain the defaults). */
void ommp_amd_set_solver_tolerance(OMMP_SYSTEM_PRT s, double tol, int32_t max_iter);
void ommp_amd_set_precision(OMMP_SYSTEM_PRT s, int32_t mode);
int32_t ommp_amd_get_precision(OMMP_SYSTEM_PRT s);

/* Iterations of the last solve: out[0..n_ipd-1] matvec-driving iterations per
   set (FP64), and for MIXED out[2] = outer FP64 steps, out[3] = total inner
   FP32 iterations.  out must hold 4 entries. */
void ommp_amd_get_last_iterations(OMMP_SYSTEM_PRT s, int32_t *out);
int32_t ommp_amd_get_last_status(OMMP_SYSTEM_PRT s);

/* Cached static field E_M2D (prepare_polelec, mod_electrostatics.F90:1487-1596):
   returns [n_ipd][pol_atoms][3] on the host (computed on the device if needed). */
double *ommp_amd_get_e_m2d(OMMP_SYSTEM_PRT s);

/* Device-resident variants for callers that keep fields and dipoles in HBM.
   d_ext is a device pointer [pol_atoms][3]; semantics as ommp_set_external_field.
   The host copy of the dipoles is refreshed only when ommp_get_ipd is called. */
void ommp_amd_set_external_field_device(OMMP_SYSTEM_PRT s, const double *d_ext, int32_t solver);
/* Device pointer of the dipoles [n_ipd][pol_atoms][3] (valid until terminate). */
const double *ommp_amd_device_ipd(OMMP_SYSTEM_PRT s);
/* Stream used by the library for this system (a hipStream_t; NULL = default). */
void ommp_amd_set_stream(OMMP_SYSTEM_PRT s, void *hip_stream);
/* Drop the warm-start guess and, with also_static_field, the cached static
   field and the near-field preconditioner of the mixed solver, as the reference
   does after a coordinate update (mod_mmpol.F90:356-361), without changing
   coordinates (the packed in-core T stays). */
void ommp_amd_reset_guess(OMMP_SYSTEM_PRT s, bool also_static_field);
/* Preconditioner of the inner FP32 solves of the mixed solver (the plain FP64
   solvers always keep the reference's z = alpha r, mod_polarization.F90:535-557):
   0 = alpha (Jacobi); 1 = near-field (default): alpha plus the first-order
   Neumann term of the pairs closer than 3 Angstrom, z_i = 2 alpha_i r_i -
   alpha_i sum_j T_ij alpha_j r_j, rebuilt on the device for every geometry.
   Convergence is decided by the FP64 residual either way. */
void ommp_amd_set_inner_preconditioner(OMMP_SYSTEM_PRT s, int32_t kind);
/* z = M^-1 r of that near-field preconditioner for ncols (1 or 2) host vectors
   r, z [ncols][pol_atoms][3], with the device kernel the inner solves use.
   Returns 1, or 0 when it is switched off for this geometry (more than 64
   neighbours per site on average) and z = alpha r. */
int32_t ommp_amd_inner_precond_apply(OMMP_SYSTEM_PRT s, int32_t ncols, const double *r, double *z);

/* Host-side introspection of the prepared state (no device work):
   which = 0 list_S_P_P, 1 list_S_P_D, 2 list_P_P (rows = source, reference layout),
   3..6 exact bond distance 1-2..1-5, 7..10 polarization-group distance 0..3.
   Returns nnz; when ri/ci/val are non-NULL they receive rows+1, nnz, nnz entries. */
int32_t ommp_amd_get_list(OMMP_SYSTEM_PRT s, int32_t which, int32_t *ri, int32_t *ci, double *val);
/* Thole factors per MM site (thole_init, mod_electrostatics.F90:866-897). */
double *ommp_amd_get_thole(OMMP_SYSTEM_PRT s);
/* Polarizabilities per polarizable site (bohr^3). */
double *ommp_amd_get_pol(OMMP_SYSTEM_PRT s);

/* One T matvec y = T x on the device for ncols (1 or 2) host vectors
   x, y [ncols][pol_atoms][3] (TMatVec_otf / TMatVec_incore with the diagonal,
   mod_polarization.F90:437-504).  kind: 0 on-the-fly FP64, 1 on-the-fly FP32
   (each the sweep the solvers use at this size), 2 packed in-core FP32, 3 packed
   in-core FP64, 4 fast multipoles (FP64, with the system's FMM settings),
   5 / 6 on-the-fly FP64 / FP32 evaluating each unordered pair once, 7 / 8
   on-the-fly FP64 / FP32 per target.  repeat > 1 launches it that
   many times back to back and returns the mean device time per launch (ms). */
double ommp_amd_tmatvec(OMMP_SYSTEM_PRT s, int32_t kind, int32_t ncols, const double *x, double *y,
                        int32_t repeat);

/* The same y = T x for 1 to 128 host columns x, y [ncols][pol_atoms][3] through
   the inner matvec of the multi-RHS batched solver: damped tensors built in
   FP32, tensors and x split in three bf16 parts each, the six leading products
   summed on the matrix cores with FP32 accumulation.  No reference counterpart:
   it exposes that kernel alone so that its accuracy and speed can be checked.
   Returns the mean device time per launch of repeat launches (ms). */
double ommp_amd_block_tmatvec(OMMP_SYSTEM_PRT s, int32_t ncols, const double *x, double *y, int32_t repeat);

/* Live kernel timing on the system's stream: when enabled, every pair-kernel
   launch of the T matvec (on-the-fly, both precisions) and every packed-T
   stream launch is bracketed by HIP events.  stats receives, per class
   k = 0 on-the-fly FP64, 1 on-the-fly FP32, 2 packed FP32, 3 packed FP64,
   4 static field, 5 geometric-gradient pair sweeps: stats[2k] = launches,
   stats[2k+1] = total milliseconds.  stats must hold 12 entries; reset clears
   the counters. */
void ommp_amd_profile_kernels(OMMP_SYSTEM_PRT s, bool enable);
/* Same, bracketing only every `every`-th launch of each class (each bracketed
   launch costs a few microseconds of idle GPU for the two event records);
   every <= 0 disables.  The stats then count the bracketed launches. */
void ommp_amd_profile_kernels_every(OMMP_SYSTEM_PRT s, int32_t every);
void ommp_amd_get_kernel_stats(OMMP_SYSTEM_PRT s, double *stats, bool reset);

/* No counterpart in the reference.  The FMA rate, in TFLOP/s (FMA = 2), that
   non-packed vector code reaches on the current device: eight independent FMA
   chains per thread over 8,192 x 256 threads, prec 32 (FP32) or 64 (FP64),
   built like the library's kernels (without packed FP32 instructions).  The
   reference point for the vector-ALU-bound pair kernels; about 0.1 s. */
double ommp_amd_vector_fma_tflops(int32_t prec);

/* Timing of the last solve on the device, in milliseconds: out[0] static field,
   out[1] solver, out[2] energy/copy-back.  out must hold 3 entries. */
void ommp_amd_get_last_timings(OMMP_SYSTEM_PRT s, double *out);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory -- code under HIP graph capture, tight serving loops, harnesses that
   forbid allocation: ommp_field_mmpol2ext (reference include/openmmpol.h:46,
   mod_interface.F90:297-311) with no allocation of any kind.  The points
   cext[n][3], the per-split partial sums and the result are staged in a static
   device array of 34.5 MB that each GPU the library runs on holds until the
   process ends; the field of the multipoles and of the induced dipoles at the
   n points is ADDED to e[n][3], with the same kernels, split counts and bits
   as ommp_field_mmpol2ext.  Limits: n <= 8192 points per call (larger n is a
   fatal error), and the system's device state must exist already (after a
   solve, or any other call that ran on the GPU; fatal otherwise).  Calls that
   overlap on one GPU (several host threads or streams) share that GPU's array
   and race; each call returns after its own work has finished. */
void ommp_amd_field_mmpol2ext_noalloc(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e);

/* ---------------------------------------------------------------- batched */
/* Several external fields of one system solved together: nrhs right-hand sides
   E_M2D + ext[k], each with its D and P sets (AMOEBA) or its single set, by the
   mixed solver (FP32 block T matvec, FP64 refinement per column) with every
   launch covering all columns.  The reference has no such call; it is the
   batched form of ommp_set_external_field (mod_interface.F90:184-223) for QM
   codes that need the response to many fields (e.g. several excited states or
   finite-difference displacements).  The guess of every column is the
   system's current dipoles when a previous solve left them, else zero; the
   system's own dipoles are not changed.  Tolerance: ommp_amd_set_solver_tolerance.
     ext   [nrhs][pol_atoms][3]
     ipd   [nrhs][n_ipd][pol_atoms][3] output, or NULL
     epol  [nrhs] polarization energies (host), or NULL */
void ommp_amd_set_external_fields_multi(OMMP_SYSTEM_PRT s, int32_t nrhs, const double *ext, double *ipd,
                                        double *epol);
/* Same with ext and ipd in device memory (epol stays a host array). */
void ommp_amd_set_external_fields_multi_device(OMMP_SYSTEM_PRT s, int32_t nrhs, const double *d_ext,
                                               double *d_ipd, double *epol);
/* Field of nq point charges (x, y, z, q in bohr and e; [nq][4]) at the
   polarizable sites, undamped and unscreened, as the reference's E_n2p
   (mod_qm_helper.F90:376-391); out [pol_atoms][3] on the host. */
void ommp_amd_charges_field(OMMP_SYSTEM_PRT s, int32_t nq, const double *charges, double *out);

/* ---------------------------------------------------------------- serving */
/* A serving loop on the current GPU, in two lanes by system size (below 4,096
   polarizable sites, and the rest), each a worker thread with its own stream:
   requests are admitted in arrival order while their lane's batch slots
   (max_active per lane) and the shared HBM budget (bytes, counted as the
   device memory of each admitted system plus its batch state) allow, and each
   lane runs its admitted solves together in ticks of the batched mixed solver.
   A request
   is one system and an SCF sequence of nsteps solves; the external field of
   step k is that of the point charges charges[k][nq][4] (x, y, z, q).  Each
   solve starts from the previous step's dipoles.  Solves that converge leave
   the batch at the next tick and the request's next solve takes their place;
   requests that arrive while others run are admitted as slots and memory free
   up.  All systems of one server must use the same force field.  Tolerance
   and iteration limits are each system's (ommp_amd_set_solver_tolerance).
   While a request is pending or active its system must not be used by other
   calls; when it completes, the system's dipoles are those of its last solve. */
typedef void *OMMP_AMD_SERVER_PTR;
OMMP_AMD_SERVER_PTR ommp_amd_server_create(double hbm_budget_bytes, int32_t max_active);
/* Returns the request id. */
int32_t ommp_amd_server_submit(OMMP_AMD_SERVER_PTR srv, OMMP_SYSTEM_PRT s, int32_t nsteps, int32_t nq,
                               const double *charges);
/* Blocks until request id has completed (id < 0: every submitted request). */
void ommp_amd_server_wait(OMMP_AMD_SERVER_PTR srv, int32_t id);
/* Results of a completed request: energies[nsteps] (Hartree), latency_ms[nsteps]
   (from the time the solve could start -- arrival for the first -- to the time
   its convergence was seen), iters[nsteps][2] (FP64 refinement steps, FP32
   iterations); any pointer may be NULL.  Returns 0 when every solve
   converged, 2 otherwise, -1 if the request has not completed. */
int32_t ommp_amd_server_result(OMMP_AMD_SERVER_PTR srv, int32_t id, double *energies, double *latency_ms,
                               int32_t *iters);
/* out[6]: ticks, solves, mean columns in flight per tick, most requests admitted
   at once, peak admitted bytes, admission deferrals. */
void ommp_amd_server_stats(OMMP_AMD_SERVER_PTR srv, double *out);
/* Waits for every submitted request, then stops the worker. */
void ommp_amd_server_destroy(OMMP_AMD_SERVER_PTR srv);

#ifdef __cplusplus
}
#endif
#endif
