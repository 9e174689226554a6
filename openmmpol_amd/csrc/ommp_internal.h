// Internal data model of the MI355X induced-dipole library.
//
// Host side: the prepared system (what the reference keeps in ommp_system /
// ommp_electrostatics_type, src/mod_electrostatics.F90:23-241), built on the
// CPU in O(N * neighbours).  Device side: HBM-resident copies laid out for the
// pair kernels (see device.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ommp {

constexpr double kEpsRp = 2.220446049250313080847e-16 * 100.0;  // mod_constants.F90:61
constexpr double kAng2Au = 1.8897261245650;                     // openmmpol_const.h.in:34
constexpr int kSolverNone = 0, kSolverCG = 1, kSolverDIIS = 2, kSolverInversion = 3;
constexpr int kMatvNone = 0, kMatvIncore = 1, kMatvDirect = 2;

// Compressed sparse rows, 0-based.
struct Csr {
    int n = 0;
    std::vector<int> ri{0};
    std::vector<int> ci;
    std::vector<double> val;
    int nnz() const { return ri.empty() ? 0 : ri.back(); }
};

struct DeviceState;  // device.h

struct System {
    // ---------------------------------------------------------------- model
    bool amoeba = false;
    int ff_rules = 0;
    int nmm = 0, npol = 0, ld_cart = 1, n_ipd = 1;
    std::vector<double> cmm;       // [nmm][3] bohr
    std::vector<double> q0;        // [nmm][ld] unrotated, AMOEBA quadrupoles / 3
    std::vector<double> q;         // [nmm][ld] lab frame
    std::vector<int> mol_frame, ix, iy, iz;  // AMOEBA frames, 0-based atom or -1
    std::vector<double> pol;       // [npol] bohr^3
    std::vector<int> polar_mm;     // [npol]
    std::vector<int> mm_polar;     // [nmm], -1 if not polarizable
    std::vector<int> atz;          // [nmm] atomic numbers (0: not read, as the .mmp loader)
    std::vector<double> cpol;      // [npol][3]
    std::vector<double> thole;     // [nmm]
    std::vector<double> thole_damping_mm;  // [nmm] damping a per MM site
    double thole_scale = 0.0;
    Csr adj;                       // 1-2
    Csr conn[4];                   // exact bond distance 1..4
    std::vector<int> grp;          // [nmm] polarization group (AMOEBA)
    int ngrp = 0;
    Csr grp_members;
    Csr gconn[4];                  // group distance 0..3
    double mscale[4], pscale[4], dscale[4], uscale[4], pscale_intra[4];
    // screening lists, rows = source (reference layout), val = scale factor
    Csr list_S_P_P, list_S_P_D, list_P_P;
    // static -> static (m-scale by bond distance), rows = source
    Csr list_S_S;
    // P-P scale factors the dense/incore T uses: the reference builds that
    // matrix with screening_rules on every pair (mod_polarization.F90:348),
    // which for AMOEBA reaches group neighbours that are not bond neighbours.
    Csr list_P_P_dense;

    // ------------------------------------------------------------ API state
    int def_solver = kSolverCG, def_matv = kMatvDirect;
    bool fmm_requested = false;
    // fast multipoles (ommp_set_fmm_*; defaults openmmpol_const.h.in:46-49)
    int fmm_lmax_pol = 8, fmm_lmax_static = 8;
    double fmm_distance = 5.0 * kAng2Au, fmm_min_cell = 7.0 * kAng2Au;
    bool m2d_done = false, ipd_done = false, ipd_use_guess = false;
    bool m2m_done = false;          // fixed-multipole energy cached (prepare_fixedelec)
    double e_mm = 0.0;
    bool host_ipd_stale = false;
    std::vector<double> ipd;       // [n_ipd][npol][3] host mirror
    std::vector<double> e_m2d;     // [n_ipd][npol][3] host mirror
    bool host_m2d_stale = true;

    // -------------------------------------------------------------- options
    double tol = 1e-8;
    int max_iter = 200;
    int precision = 0;             // OMMP_AMD_PREC_*
    int inner_precond = 1;         // inner FP32 solves: 0 alpha (Jacobi), 1 near-field (default)
    int last_iters[4] = {0, 0, 0, 0};
    // iterations each CG call site took last time: the first batch of
    // iterations is enqueued without a host round trip (slot: 0/1 FP64 D/P,
    // 2.. mixed inner solves by refinement step)
    int pred_iters[16] = {0};
    int pred_outer = 0;            // FP64 residuals the last mixed solve needed
    int last_status = 0;
    double last_ms[3] = {0, 0, 0};
    void *user_stream = nullptr;

    DeviceState *dev = nullptr;
    bool dev_geometry_dirty = true;  // positions/multipoles need re-upload
};

// prep.cpp
System *read_mmp(const std::string &path);
void turn_pol_off(System &s, int n, const int *nopol);  // 1-based MM indices
System *build_from_arrays(int ff_type, int ff_rules, int nmm, const double *cmm, const double *q,
                          const double *pol, const double *thole_a, const int *adj_ri,
                          const int *adj_ci, const int *pgroup, const int *frames);
void prepare(System &s);                 // mmpol_prepare
void rotate_multipoles(System &s);
void make_screening_lists(System &s);
void check_coincident_sites(const System &s);
void set_default_scales(System &s);

// io.cpp
void message(const std::string &msg, int level, const char *prefix = nullptr);
[[noreturn]] void fatal(const std::string &msg);
// While a FatalToException lives on a thread, fatal() on that thread throws
// FatalError instead of printing and exiting: worker threads catch it, record
// the message and wake the caller's thread, which then calls fatal() itself,
// so the process exits from the thread that called the library.
struct FatalError {
    std::string msg;
};
class FatalToException {
  public:
    FatalToException();
    ~FatalToException();
    FatalToException(const FatalToException &) = delete;
    FatalToException &operator=(const FatalToException &) = delete;

  private:
    bool prev_;
};
int verbosity();

// engine.cpp
// Inner-solve target of the mixed refinement: max(eta * outer rms, frac * tol).
double refine_eta();
double refine_frac();
void device_release(System &s);
void prepare_polelec(System &s);                              // E_M2D on device
void polarization(System &s, const double *d_field, int solver, int matv, int mask);
double polelec_energy(System &s);
double fixedelec_energy(System &s);                           // energy_MM_MM
void refresh_host_ipd(System &s);
void set_external_field(System &s, const double *ext, bool ext_on_device, int solver, int matv,
                        bool nomm);
void invalidate_geometry(System &s);
void invalidate_precond(System &s);  // the near-field preconditioner is rebuilt before the next mixed solve
const double *device_ipd(System &s);
void refresh_host_m2d(System &s);
double tmatvec_host(System &s, int kind, int ncols, const double *x, double *y, int repeat);
double block_tmatvec_host(System &s, int ncols, const double *x, double *y, int repeat);
int inner_precond_apply_host(System &s, int ncols, const double *r, double *z);
void profile_enable(System &s, int every);  // 0 off, n > 0: every n-th launch per class
void profile_stats(System &s, double *out, bool reset);
void solve_multi(System &s, int nrhs, const double *ext, bool ext_dev, double *ipd_out, bool out_dev,
                 double *epol);
void charges_field_host(System &s, int nq, const double *charges, double *out);
void x2e_host(System &s, int n, const double *cext, bool mm, bool pol, double *V, double *E);
void x2e_field_noalloc_host(System &s, int n, const double *cext, double *E);
void fmm_settings_changed(System &s);  // drops the tree, the static field and the dipoles
void polelec_geomgrad_host(System &s, double *grd);
void rotation_geomgrad_host(System &s, const double *E, const double *Egrd, double *grd);

// kernels.hip: FMA rate (TFLOP/s) of non-packed vector code on the current
// device, prec 32 or 64 (fatal without a GPU)
double vector_fma_tflops(int prec);

// server.cpp
void *server_create(double budget, int max_active);
void server_destroy(void *srv);
int server_submit(void *srv, System *s, int nsteps, int nq, const double *charges);
void server_wait(void *srv, int id);
int server_result(void *srv, int id, double *energies, double *latency_ms, int *iters);
void server_stats(void *srv, double *out);

}  // namespace ommp
