// Fast multipole state of one system (device pointers) and its launchers.
//
// Uniform octree over the bounding cube of the MM atoms, `levels` = L below the
// root (leaves: 2^L per side, side >= max(min cell size, near-field distance)),
// dense cell numbering per level (ix + 2^l (iy + 2^l iz)), cells of all levels in
// one array at offsets cell_off[l].  Polarizable sites and MM atoms are each
// sorted by leaf; *_leaf_start[leaf] gives the sorted range of a leaf.
#pragma once

#include <hip/hip_runtime.h>

namespace ommp {

constexpr int kFmmMaxLevels = 7;   // 2^21 leaves at most (dense per-level arrays)
constexpr int kFmmMaxOrder = 20;   // (p + 1)^2 <= 512 coefficients: two per thread in the M2L/M2M kernels
constexpr int kFmmNearSplits = 27; // near-field partials, one per neighbour leaf; the far field is one more

struct FmmDev {
    int npol = 0, nmm = 0;
    int levels = 0;            // L
    int p_pol = 8, p_static = 8, p_tab = 8;
    int ncoef_tab = 81, ncoef2_tab = 289;
    double leaf_h = 0.0;       // leaf side (bohr)
    double3 origin = {0.0, 0.0, 0.0};
    long cell_off[kFmmMaxLevels + 2] = {0};
    // sorted polarizable sites
    int *pol_order = nullptr;      // sorted -> polarizable index
    int *pol_leaf = nullptr;       // sorted -> leaf
    int *pol_self = nullptr;       // sorted -> sorted (self exclusion in the dipole near field)
    int *pol_mmself = nullptr;     // sorted -> sorted MM position of the same atom
    int *pol_leaf_start = nullptr; // [nleaf + 1]
    double4 *pol_pos = nullptr;    // (x, y, z, 1/thole)
    double *pol_a = nullptr;       // damping a
    // sorted MM atoms (static sources, gradient targets)
    int *mm_order = nullptr;       // sorted -> MM index
    int *mm_leaf = nullptr;        // sorted -> leaf
    int *mm_leaf_start = nullptr;
    double4 *mm_pos = nullptr;
    double *mm_a = nullptr;
    double *mm_q = nullptr;        // [nmm][ld]
    // occupancy per cell of every level
    int *cnt_pol = nullptr, *cnt_mm = nullptr;
    // expansions [cell][coef] (coef stride of the order in use), translation tables
    double2 *mexp = nullptr, *lexp = nullptr;
    double2 *tab_m2m = nullptr;    // [level][8][ncoef_tab]   R(c_child - c_parent), parent level
    double2 *tab_l2l = nullptr;    // [level][8][ncoef_tab]   S(c_child - c_parent), parent level
    double2 *tab_m2l = nullptr;    // [level][343][ncoef2_tab] T(c_source - c_target)
    double *xs = nullptr;          // sorted dipole columns [3][npol][3]
    double2 *mstatic = nullptr;    // [nleaf][ncoef_tab] static multipole moments of the leaves (order p_tab)
};

// Moments of the static multipoles in every leaf, order p_tab, into f.mstatic:
// computed once per tree (the multipoles change only with the geometry) and
// used by the static field and the gradient.  Nothing to do without a far field.
void fmm_launch_static_moments(hipStream_t st, const FmmDev &f, int ld);

// Near (splits 0 .. 26, one per neighbour leaf) and far (split 27) dipole fields
// of the M columns x at the polarizable sites, in the partial layout of the
// on-the-fly matvec: part[(split * M + m) * 3 npol + 3 j + c].
void fmm_launch_matvec_fields(hipStream_t st, const FmmDev &f, bool amoeba, int M, const double *x, long xstride,
                              double *part);
// Near (splits 0 .. 26) and far (split 27) fields of the static multipoles at
// the polarizable sites: part[split * 3 npol + 3 j + c].  Needs f.mstatic.
void fmm_launch_static_fields(hipStream_t st, const FmmDev &f, bool amoeba, int ld, double *part);
// Pair sums of the polarization-energy gradient, near (splits 0 .. 26) and far,
// in the partial layouts of the direct gradient sweeps:
//   part_pol[split][3][npol]: sum_b w_b Egrd_ab at the polarizable sites, far
//     splits 27 and 28 (AMOEBA: static + mu_P sources against 1/2 mu_D, static +
//     mu_D against 1/2 mu_P) or 27 (AMBER: static + mu against mu);
//   part_mm[split][na][nmm]: E, Egrd and Q : EHes (na = 12, AMOEBA) or E (na = 3)
//     of the dipoles nu = 1/2 (mu_D + mu_P) (AMOEBA) or mu at the MM atoms, far
//     split 27.
// Every pair within the 27 leaves around a target is in the near field with
// scale 1 and the damped kernels; ipd = [D or mu][npol][3] (+ [P][npol][3]).
// Needs f.mstatic.
void fmm_launch_grad_fields(hipStream_t st, const FmmDev &f, bool amoeba, int ld, const double *ipd,
                            double *part_pol, double *part_mm);
constexpr int fmm_grad_parts_pol(bool amoeba) { return kFmmNearSplits + (amoeba ? 2 : 1); }
constexpr int fmm_grad_parts_mm() { return kFmmNearSplits + 1; }

}  // namespace ommp
