/*
 * openmmpol.h -- drop-in C interface of the induced-dipole polarization path.
 *
 * Every entry point below has the name, signature, argument meaning, array
 * layout and error behaviour of the reference function it replaces (cited as
 * reference include/openmmpol.h:<line> and the Fortran routine behind it).
 * Arrays are C row-major views of the reference's Fortran column-major
 * storage: ipd is [n_ipd][pol_atoms][3], fields are [pol_atoms][3],
 * coordinates [mm_atoms][3], multipoles [mm_atoms][ld_cart].  All lengths are
 * in atomic units.  Errors print "[OMMP]   [stop] ..." and exit(1), as the
 * reference's fatal_error (src/mod_io.F90:155-175).
 *
 * Pointers returned by the getters point into library-owned host memory that
 * stays valid until ommp_terminate (or until ommp_update_coordinates for the
 * coordinate-derived arrays), and is refreshed in place after every solve.
 */
#ifndef _OPENMMPOL
#define _OPENMMPOL

#include <stdbool.h>
#include <stdint.h>
#include "openmmpol_const.h"

typedef void *OMMP_SYSTEM_PRT;

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/openmmpol.h:17; mod_c_interface.F90:216-231 -> mod_inputloader.F90:15-257 */
OMMP_SYSTEM_PRT ommp_init_mmp(const char *filename);
/* :19; mod_electrostatics.F90 set_def_solver (fatal on unknown id) */
void ommp_set_default_solver(OMMP_SYSTEM_PRT s, int32_t solver);
/* :20; mod_electrostatics.F90 set_def_matv (fatal on unknown id) */
void ommp_set_default_matv(OMMP_SYSTEM_PRT s, int32_t matv);
/* :23; mod_interface.F90:134-168 + remove_null_pol (mod_electrostatics.F90:469-536):
   the MM atoms nopol[0..n) (1-based; fatal outside 1..mm_atoms) lose their
   polarizability and the polarizable set is compacted in order; the dipoles
   are dropped, and the pointers of the polarizable-site getters (ipd, cpol,
   polar_mm, pol) must be fetched again.  The Thole factors keep their values.
   The reference's loop reads its counter where it means the list entry (:156);
   the two agree for the lists 1..n its callers pass. */
void ommp_turn_pol_off(OMMP_SYSTEM_PRT s, int32_t n, const int32_t *nopol);
/* :24; mod_interface.F90:161-170 */
void ommp_terminate(OMMP_SYSTEM_PRT s);
/* :28; mod_io.F90:92-103 (clamped to 0..3) */
void ommp_set_verbose(int32_t level);
/* :29-30; mod_io.F90 set_iof_mmpol / close_output */
void ommp_set_outputfile(const char *filename);
void ommp_close_outputfile(void);
/* :31-32; mod_io.F90:105-153 and :155-175 */
void ommp_message(const char *msg, int32_t level, const char *prefix);
void ommp_fatal(const char *msg);
/* :33-34; mod_profiling.F90:20-71 */
void ommp_time_pull(const char *label);
void ommp_time_push(void);

/* :38; mod_interface.F90:343-370 -> energy_MM_pol (mod_electrostatics.F90:943-976).
   Solves with the default solver first if no dipoles are available. */
double ommp_get_polelec_energy(OMMP_SYSTEM_PRT s);
/* :39; mod_c_interface.F90:390-400 -> mod_interface.F90:372-383 -> energy_MM_MM
   (mod_electrostatics.F90:899-941) on prepare_fixedelec / elec_prop_M2M
   (:1407-1485, 1678-1952): interaction energy of the static multipoles
   (Hartree), m-scaled by bond distance; cached until the coordinates change.
   Computed by the exact direct sums also when fast multipoles are enabled. */
double ommp_get_fixedelec_energy(OMMP_SYSTEM_PRT s);
/* :63; mod_interface.F90:388-398: ommp_get_fixedelec_energy + ommp_get_polelec_energy. */
double ommp_get_full_ele_energy(OMMP_SYSTEM_PRT s);
/* :40; mod_c_interface.F90:289-305 -> mod_interface.F90:184-223.  ext is
   [pol_atoms][3]; the MM field is added and both D and P sets are solved with
   the default matrix-vector method (the matv argument is ignored, as in the
   reference). */
void ommp_set_external_field(OMMP_SYSTEM_PRT s, const double *ext, int32_t solver, int32_t matv);
/* :41; mod_interface.F90:224-232: field = ext only, solves the D set only. */
void ommp_set_external_field_nomm(OMMP_SYSTEM_PRT s, const double *ext, int32_t solver, int32_t matv);

/* :43-48; mod_interface.F90:250-341.  Potential v[n] / field e[n][3] at n
   external points cext[n][3] (bohr) of the static multipoles (mm), of the
   induced dipoles (pol: D set for the potential, (D + P) / 2 for the field
   with AMOEBA) or of both, undamped and unscreened; ADDED to v / e.  The
   dipole terms need a completed solve (fatal otherwise). */
void ommp_potential_mmpol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v);
void ommp_potential_mm2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v);
void ommp_potential_pol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *v);
void ommp_field_mmpol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e);
void ommp_field_mm2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e);
void ommp_field_pol2ext(OMMP_SYSTEM_PRT s, int32_t n, const double *cext, double *e);

/* :91; mod_c_interface.F90:833-846 -> mod_interface.F90:671-683 ->
   polelec_geomgrad (mod_geomgrad.F90:113-273).  grd[mm_atoms][3] is set to the
   gradient of the polarization energy with respect to the MM coordinates
   (Hartree/bohr), AMOEBA multipole rotation included; 0 without polarizable
   atoms.  Solves with the default solver (no external field) first if no
   dipoles are available.  The external field's own position dependence is the
   caller's (as in the reference). */
void ommp_polelec_geomgrad(OMMP_SYSTEM_PRT s, double *grd);
/* :92; mod_c_interface.F90:606-624 -> mod_interface.F90:699-710 ->
   rotation_geomgrad (rotate_multipoles.F90:3-230).  grd[mm_atoms][3] is set to
   the gradient, through the frame atoms, of the energy of the rotated
   multipoles in a field E[mm_atoms][3] and field gradient Egrd[mm_atoms][6]
   (xx, xy, yy, xz, yz, zz; -dE_b/dr_a).  0 for non-AMOEBA systems. */
void ommp_rotation_geomgrad(OMMP_SYSTEM_PRT s, const double *E, const double *Egrd, double *grd);

/* :65; mod_c_interface.F90 C_ommp_ff_is_amoeba */
bool ommp_ff_is_amoeba(OMMP_SYSTEM_PRT s);
/* :68-71 */
int32_t ommp_get_n_ipd(OMMP_SYSTEM_PRT s);
int32_t ommp_get_ld_cart(OMMP_SYSTEM_PRT s);
int32_t ommp_get_mm_atoms(OMMP_SYSTEM_PRT s);
int32_t ommp_get_pol_atoms(OMMP_SYSTEM_PRT s);
/* :73 [mm_atoms][3] */
double *ommp_get_cmm(OMMP_SYSTEM_PRT s);
/* :76 [pol_atoms][3] */
double *ommp_get_cpol(OMMP_SYSTEM_PRT s);
/* :77 [mm_atoms][ld_cart], lab-frame (rotated) multipoles */
double *ommp_get_q(OMMP_SYSTEM_PRT s);
/* :78; mod_c_interface.F90:999-1008: [n_ipd][pol_atoms][3] */
double *ommp_get_ipd(OMMP_SYSTEM_PRT s);
/* :79; mod_c_interface.F90:1010-1026: 0-based MM index of each polarizable site */
int32_t *ommp_get_polar_mm(OMMP_SYSTEM_PRT s);
/* :74; mod_c_interface.F90:950-959: atomic number of each MM atom [mm_atoms];
   zero for systems read from .mmp files, whose loader skips that section
   (mod_inputloader.F90:131) after topology_init zeroed it */
int32_t *ommp_get_zmm(OMMP_SYSTEM_PRT s);

/* :86; mod_mmpol.F90:325-372: new coordinates [mm_atoms][3]; clears the cached
   static field, the dipoles and the warm-start guess, re-rotates multipoles. */
void ommp_update_coordinates(OMMP_SYSTEM_PRT s, const double *new_c);

/* :153-160; mod_c_interface.F90:1781-1880.  Fast multipoles for the static
   field and the T matvec of the solvers (CG, DIIS; inversion stays dense): a
   uniform octree whose leaves are at least max(min cell size, near-field
   distance) wide (defaults 7 and 5 Angstrom, given in bohr), pairs in the 27
   leaves around a site computed directly with damping and screening, farther
   pairs through complex solid-harmonic expansions of order lmax (static) and
   lmax_pol (dipoles), default 8, at most 20.  Any change drops the static field
   and the dipoles.  With fast multipoles the solve runs in FP64 whatever
   ommp_amd_set_precision says, and the gradient goes through them too; the
   batched solvers stay direct.
   Unlike the reference, which switches them on by itself above 1,000 MM atoms
   (mod_electrostatics.F90:294-300), this library computes every interaction
   directly until ommp_enable_fmm is called. */
void ommp_enable_fmm(OMMP_SYSTEM_PRT s);
void ommp_disable_fmm(OMMP_SYSTEM_PRT s);
bool ommp_use_fmm(OMMP_SYSTEM_PRT s);
void ommp_set_fmm_lmax_pol(OMMP_SYSTEM_PRT s, int32_t l);
void ommp_set_fmm_lmax(OMMP_SYSTEM_PRT s, int32_t l);
void ommp_set_fmm_distance(OMMP_SYSTEM_PRT s, double d);
void ommp_set_fmm_min_cell_size(OMMP_SYSTEM_PRT s, double d);

#ifdef __cplusplus
}
#endif
#endif
