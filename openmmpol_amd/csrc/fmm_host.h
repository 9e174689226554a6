// Host-built octree and translation tables of the fast multipole path (see
// fmm.h for the device layout they are uploaded into).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fmm.h"

namespace ommp {

struct System;

struct FmmHost {
    int levels = 0, p_pol = 8, p_static = 8, p_tab = 8, ncoef_tab = 81, ncoef2_tab = 289;
    double side = 0.0, leaf_h = 0.0, origin[3] = {0.0, 0.0, 0.0};
    std::vector<long> cell_off;
    std::vector<int> pol_order, pol_leaf, pol_mmself, pol_leaf_start;
    std::vector<int> mm_order, mm_leaf, mm_leaf_start;
    std::vector<int> cnt_pol, cnt_mm;
    std::vector<double2> tab_m2m, tab_l2l, tab_m2l;
};

// Octree with leaves of side >= max(min_cell, distance) (bohr) over the MM
// atoms of s, at most kFmmMaxLevels levels below the root.
FmmHost fmm_build(const System &s, int p_pol, int p_static, double min_cell, double distance);

}  // namespace ommp
