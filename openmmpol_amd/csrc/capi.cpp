// extern "C" surface: the reference's ommp_* entry points for the
// polarization path (include/openmmpol.h) and the extensions in ommp_amd.h.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/ommp_amd.h"
#include "ommp_internal.h"

using ommp::System;

namespace {
System *sys(OMMP_SYSTEM_PRT p) {
    if (!p) ommp::fatal("NULL system handle");
    return static_cast<System *>(p);
}
}  // namespace

extern "C" {

OMMP_SYSTEM_PRT ommp_init_mmp(const char *filename) {
    ommp::message(std::string("OpenMMPol version: ") + OMMP_VERSION_STRING, 1);
    return ommp::read_mmp(filename ? filename : "");
}

void ommp_set_default_solver(OMMP_SYSTEM_PRT s, int32_t solver) {
    // set_def_solver (mod_electrostatics.F90)
    if (solver != OMMP_SOLVER_CG && solver != OMMP_SOLVER_INVERSION && solver != OMMP_SOLVER_DIIS)
        ommp::fatal("Unrecognized setting for default solver method");
    sys(s)->def_solver = solver;
}

void ommp_set_default_matv(OMMP_SYSTEM_PRT s, int32_t matv) {
    if (matv != OMMP_MATV_INCORE && matv != OMMP_MATV_DIRECT)
        ommp::fatal("Unrecognized setting for default matrix-vector method");
    sys(s)->def_matv = matv;
}

void ommp_terminate(OMMP_SYSTEM_PRT s) {
    if (!s) return;
    System *p = sys(s);
    ommp::device_release(*p);
    delete p;
}

double ommp_get_fixedelec_energy(OMMP_SYSTEM_PRT s) {
    ommp_time_push();
    double e = ommp::fixedelec_energy(*sys(s));
    ommp_time_pull("Fixed Electrostatic energy");
    return e;
}

// mod_interface.F90:388-398: fixed multipoles plus polarization
double ommp_get_full_ele_energy(OMMP_SYSTEM_PRT s) {
    return ommp_get_fixedelec_energy(s) + ommp_get_polelec_energy(s);
}

double ommp_get_polelec_energy(OMMP_SYSTEM_PRT s) {
    ommp_time_push();
    double e = ommp::polelec_energy(*sys(s));
    ommp_time_pull("Polarizable Electrostatic energy");
    return e;
}

void ommp_set_external_field(OMMP_SYSTEM_PRT s, const double *ext, int32_t solver, int32_t matv) {
    ommp::set_external_field(*sys(s), ext, false, solver, matv, false);
}

void ommp_set_external_field_nomm(OMMP_SYSTEM_PRT s, const double *ext, int32_t solver, int32_t matv) {
    ommp::set_external_field(*sys(s), ext, false, solver, matv, true);
}

bool ommp_ff_is_amoeba(OMMP_SYSTEM_PRT s) { return sys(s)->amoeba; }
int32_t ommp_get_n_ipd(OMMP_SYSTEM_PRT s) { return sys(s)->n_ipd; }
int32_t ommp_get_ld_cart(OMMP_SYSTEM_PRT s) { return sys(s)->ld_cart; }
int32_t ommp_get_mm_atoms(OMMP_SYSTEM_PRT s) { return sys(s)->nmm; }
int32_t ommp_get_pol_atoms(OMMP_SYSTEM_PRT s) { return sys(s)->npol; }
double *ommp_get_cmm(OMMP_SYSTEM_PRT s) { return sys(s)->cmm.data(); }
double *ommp_get_cpol(OMMP_SYSTEM_PRT s) { return sys(s)->cpol.data(); }
double *ommp_get_q(OMMP_SYSTEM_PRT s) { return sys(s)->q.data(); }

double *ommp_get_ipd(OMMP_SYSTEM_PRT s) {
    System *p = sys(s);
    ommp::refresh_host_ipd(*p);
    return p->ipd.data();
}

int32_t *ommp_get_polar_mm(OMMP_SYSTEM_PRT s) { return sys(s)->polar_mm.data(); }
int32_t *ommp_get_zmm(OMMP_SYSTEM_PRT s) { return sys(s)->atz.data(); }

void ommp_turn_pol_off(OMMP_SYSTEM_PRT s, int32_t n, const int32_t *nopol) {
    ommp::turn_pol_off(*sys(s), n, nopol);
}

void ommp_update_coordinates(OMMP_SYSTEM_PRT s, const double *new_c) {
    // update_coordinates (mod_mmpol.F90:325-372)
    System &p = *sys(s);
    std::memcpy(p.cmm.data(), new_c, sizeof(double) * 3 * p.nmm);
    for (int i = 0; i < p.npol; i++)
        for (int c = 0; c < 3; c++) p.cpol[3 * i + c] = p.cmm[3 * p.polar_mm[i] + c];
    p.m2d_done = false;
    p.m2m_done = false;
    p.ipd_done = false;
    p.ipd_use_guess = false;
    if (p.amoeba) ommp::rotate_multipoles(p);
    ommp::invalidate_geometry(p);
}

// mod_c_interface.F90:1781-1880
void ommp_enable_fmm(OMMP_SYSTEM_PRT s) {
    sys(s)->fmm_requested = true;
    ommp::fmm_settings_changed(*sys(s));
}

void ommp_disable_fmm(OMMP_SYSTEM_PRT s) {
    sys(s)->fmm_requested = false;
    ommp::fmm_settings_changed(*sys(s));
}

bool ommp_use_fmm(OMMP_SYSTEM_PRT s) { return sys(s)->fmm_requested; }

static int fmm_order(int32_t l) {
    if (l < 0 || l > 20) ommp::fatal("FMM expansion order must be between 0 and 20");
    return l;
}

void ommp_set_fmm_lmax_pol(OMMP_SYSTEM_PRT s, int32_t l) {
    sys(s)->fmm_lmax_pol = fmm_order(l);
    ommp::fmm_settings_changed(*sys(s));
}

void ommp_set_fmm_lmax(OMMP_SYSTEM_PRT s, int32_t l) {
    sys(s)->fmm_lmax_static = fmm_order(l);
    ommp::fmm_settings_changed(*sys(s));
}

void ommp_set_fmm_distance(OMMP_SYSTEM_PRT s, double d) {
    if (!(d >= 0.0)) ommp::fatal("FMM near-field distance must be non-negative");
    sys(s)->fmm_distance = d;
    ommp::fmm_settings_changed(*sys(s));
}

void ommp_set_fmm_min_cell_size(OMMP_SYSTEM_PRT s, double d) {
    if (!(d > 0.0)) ommp::fatal("FMM minimum cell size must be positive");
    sys(s)->fmm_min_cell = d;
    ommp::fmm_settings_changed(*sys(s));
}

// ------------------------------------------------------------ extensions
OMMP_SYSTEM_PRT ommp_amd_init_arrays(int32_t ff_type, int32_t ff_rules, int32_t mm_atoms, const double *cmm,
                                     const double *q, const double *pol, const double *thole_a,
                                     const int32_t *adj_ri, const int32_t *adj_ci, const int32_t *pgroup,
                                     const int32_t *frames) {
    if (ff_type != OMMP_FF_AMOEBA && ff_type != OMMP_FF_AMBER) ommp::fatal("Cannot initialize an UNKNOWN forcefield!");
    return ommp::build_from_arrays(ff_type, ff_rules, mm_atoms, cmm, q, pol, thole_a, adj_ri, adj_ci, pgroup,
                                   frames);
}

void ommp_amd_set_screening_parameters(OMMP_SYSTEM_PRT s, const double *m, const double *p, const double *d,
                                       const double *u, const double *intra) {
    // set_screening_parameters (mod_electrostatics.F90:440-467)
    System &sy = *sys(s);
    if (intra && !sy.amoeba) ommp::fatal("Scale factors for atoms of the same group cannot be set outside AMOEBA FF");
    if (!intra && sy.amoeba) ommp::fatal("Scale factors for atoms of the same group should be defined in AMOEBA FF");
    std::memcpy(sy.mscale, m, sizeof(sy.mscale));
    std::memcpy(sy.pscale, p, sizeof(sy.pscale));
    std::memcpy(sy.dscale, d, sizeof(sy.dscale));
    std::memcpy(sy.uscale, u, sizeof(sy.uscale));
    if (intra) std::memcpy(sy.pscale_intra, intra, sizeof(sy.pscale_intra));
    ommp::make_screening_lists(sy);
    // device copies of the lists and every cached result depend on the scales
    ommp::device_release(sy);
    sy.m2d_done = false;
    sy.m2m_done = false;
    sy.ipd_done = false;
    sy.dev_geometry_dirty = true;
}

void ommp_amd_set_solver_tolerance(OMMP_SYSTEM_PRT s, double tol, int32_t max_iter) {
    if (!(tol > 0.0)) ommp::fatal("Solver tolerance must be positive");
    if (max_iter < 1) ommp::fatal("Maximum number of iterations must be positive");
    sys(s)->tol = tol;
    sys(s)->max_iter = max_iter;
}

void ommp_amd_set_precision(OMMP_SYSTEM_PRT s, int32_t mode) {
    if (mode != OMMP_AMD_PREC_FP64 && mode != OMMP_AMD_PREC_MIXED) ommp::fatal("Unknown precision mode");
    sys(s)->precision = mode;
}

int32_t ommp_amd_get_precision(OMMP_SYSTEM_PRT s) { return sys(s)->precision; }

void ommp_amd_get_last_iterations(OMMP_SYSTEM_PRT s, int32_t *out) {
    for (int k = 0; k < 4; k++) out[k] = sys(s)->last_iters[k];
}

int32_t ommp_amd_get_last_status(OMMP_SYSTEM_PRT s) { return sys(s)->last_status; }

double *ommp_amd_get_e_m2d(OMMP_SYSTEM_PRT s) {
    System &p = *sys(s);
    if (p.npol > 0) ommp::refresh_host_m2d(p);
    return p.e_m2d.data();
}

void ommp_amd_set_external_field_device(OMMP_SYSTEM_PRT s, const double *d_ext, int32_t solver) {
    ommp::set_external_field(*sys(s), d_ext, true, solver, OMMP_MATV_NONE, false);
}

const double *ommp_amd_device_ipd(OMMP_SYSTEM_PRT s) { return ommp::device_ipd(*sys(s)); }

void ommp_amd_set_stream(OMMP_SYSTEM_PRT s, void *hip_stream) {
    System &p = *sys(s);
    if (p.dev) ommp::fatal("ommp_amd_set_stream must be called before the first device operation");
    p.user_stream = hip_stream;
}

void ommp_amd_reset_guess(OMMP_SYSTEM_PRT s, bool also_static_field) {
    System &p = *sys(s);
    p.ipd_use_guess = false;
    p.ipd_done = false;
    if (also_static_field) {
        p.m2d_done = false;
        ommp::invalidate_precond(p);
    }
}

int32_t ommp_amd_inner_precond_apply(OMMP_SYSTEM_PRT s, int32_t ncols, const double *r, double *z) {
    return ommp::inner_precond_apply_host(*sys(s), ncols, r, z);
}

void ommp_amd_set_inner_preconditioner(OMMP_SYSTEM_PRT s, int32_t kind) {
    if (kind != 0 && kind != 1) ommp::fatal("Unknown preconditioner for the inner solves");
    sys(s)->inner_precond = kind;
}

int32_t ommp_amd_get_list(OMMP_SYSTEM_PRT s, int32_t which, int32_t *ri, int32_t *ci, double *val) {
    System &p = *sys(s);
    const ommp::Csr *c = nullptr;
    if (which == 0) c = &p.list_S_P_P;
    else if (which == 1) c = &p.list_S_P_D;
    else if (which == 2) c = &p.list_P_P;
    else if (which >= 3 && which < 7) c = &p.conn[which - 3];
    else if (which >= 7 && which < 11) c = &p.gconn[which - 7];
    else ommp::fatal("Unknown list id");
    if (c->ri.size() < 2 && c->n == 0) return 0;
    const int nnz = c->nnz();
    if (ri) std::memcpy(ri, c->ri.data(), sizeof(int32_t) * c->ri.size());
    if (ci && nnz) std::memcpy(ci, c->ci.data(), sizeof(int32_t) * nnz);
    if (val && nnz && !c->val.empty()) std::memcpy(val, c->val.data(), sizeof(double) * nnz);
    return nnz;
}

double *ommp_amd_get_thole(OMMP_SYSTEM_PRT s) { return sys(s)->thole.data(); }
double *ommp_amd_get_pol(OMMP_SYSTEM_PRT s) { return sys(s)->pol.data(); }

double ommp_amd_tmatvec(OMMP_SYSTEM_PRT s, int32_t kind, int32_t ncols, const double *x, double *y,
                        int32_t repeat) {
    return ommp::tmatvec_host(*sys(s), kind, ncols, x, y, repeat);
}

double ommp_amd_block_tmatvec(OMMP_SYSTEM_PRT s, int32_t ncols, const double *x, double *y, int32_t repeat) {
    return ommp::block_tmatvec_host(*sys(s), ncols, x, y, repeat);
}

void ommp_amd_profile_kernels(OMMP_SYSTEM_PRT s, bool enable) { ommp::profile_enable(*sys(s), enable ? 1 : 0); }

void ommp_amd_profile_kernels_every(OMMP_SYSTEM_PRT s, int32_t every) {
    ommp::profile_enable(*sys(s), every > 0 ? every : 0);
}

void ommp_amd_get_kernel_stats(OMMP_SYSTEM_PRT s, double *stats, bool reset) {
    ommp::profile_stats(*sys(s), stats, reset);
}

double ommp_amd_vector_fma_tflops(int32_t prec) {
    if (prec != 32 && prec != 64) ommp::fatal("vector_fma_tflops: precision must be 32 or 64");
    return ommp::vector_fma_tflops(prec);
}

void ommp_amd_get_last_timings(OMMP_SYSTEM_PRT s, double *out) {
    for (int k = 0; k < 3; k++) out[k] = sys(s)->last_ms[k];
}

// mod_interface.F90:250-341 (potential_M2E / D2E, field_M2E / D2E), results
// added to the caller's arrays.
void ommp_potential_mmpol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v) {
    ommp::x2e_host(*sys(s), n, cext, true, true, v, nullptr);
}
void ommp_potential_mm2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v) {
    ommp::x2e_host(*sys(s), n, cext, true, false, v, nullptr);
}
void ommp_potential_pol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v) {
    ommp::x2e_host(*sys(s), n, cext, false, true, v, nullptr);
}
void ommp_field_mmpol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e) {
    ommp::x2e_host(*sys(s), n, cext, true, true, nullptr, e);
}

void ommp_amd_field_mmpol2ext_noalloc(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e) {
    ommp::x2e_field_noalloc_host(*sys(s), n, cext, e);
}
void ommp_field_mm2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e) {
    ommp::x2e_host(*sys(s), n, cext, true, false, nullptr, e);
}
void ommp_field_pol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e) {
    ommp::x2e_host(*sys(s), n, cext, false, true, nullptr, e);
}

void ommp_polelec_geomgrad(OMMP_SYSTEM_PRT s, double *grd) { ommp::polelec_geomgrad_host(*sys(s), grd); }

void ommp_rotation_geomgrad(OMMP_SYSTEM_PRT s, const double *E, const double *Egrd, double *grd) {
    ommp::rotation_geomgrad_host(*sys(s), E, Egrd, grd);
}

void ommp_amd_set_external_fields_multi(OMMP_SYSTEM_PRT s, int32_t nrhs, const double *ext, double *ipd,
                                        double *epol) {
    ommp::solve_multi(*sys(s), nrhs, ext, false, ipd, false, epol);
}

void ommp_amd_set_external_fields_multi_device(OMMP_SYSTEM_PRT s, int32_t nrhs, const double *d_ext,
                                               double *d_ipd, double *epol) {
    ommp::solve_multi(*sys(s), nrhs, d_ext, true, d_ipd, true, epol);
}

void ommp_amd_charges_field(OMMP_SYSTEM_PRT s, int32_t nq, const double *charges, double *out) {
    ommp::charges_field_host(*sys(s), nq, charges, out);
}

OMMP_AMD_SERVER_PTR ommp_amd_server_create(double hbm_budget_bytes, int32_t max_active) {
    return ommp::server_create(hbm_budget_bytes, max_active);
}

int32_t ommp_amd_server_submit(OMMP_AMD_SERVER_PTR srv, OMMP_SYSTEM_PRT s, int32_t nsteps, int32_t nq,
                               const double *charges) {
    return ommp::server_submit(srv, sys(s), nsteps, nq, charges);
}

void ommp_amd_server_wait(OMMP_AMD_SERVER_PTR srv, int32_t id) { ommp::server_wait(srv, id); }

int32_t ommp_amd_server_result(OMMP_AMD_SERVER_PTR srv, int32_t id, double *energies, double *latency_ms,
                               int32_t *iters) {
    return ommp::server_result(srv, id, energies, latency_ms, iters);
}

void ommp_amd_server_stats(OMMP_AMD_SERVER_PTR srv, double *out) { ommp::server_stats(srv, out); }

void ommp_amd_server_destroy(OMMP_AMD_SERVER_PTR srv) {
    ommp::server_wait(srv, -1);
    ommp::server_destroy(srv);
}

}  // extern "C"
