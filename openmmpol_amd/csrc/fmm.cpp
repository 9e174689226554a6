// Host side of the fast multipole path: the octree over the current
// coordinates (counting sort of sites and atoms by leaf, occupancy per cell) and
// the translation tables of each level, built once per geometry and uploaded by
// engine.cpp.  The harmonics are the scaled solid harmonics of fmm_kernels.hip,
// from the same Cartesian recurrences.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "fmm_host.h"
#include "ommp_internal.h"

namespace ommp {

namespace {

using cplx = std::complex<double>;

inline int hidx(int n, int m) { return n * n + n + m; }

cplx ipow(int e) {
    switch (e & 3) {
        case 0: return {1.0, 0.0};
        case 1: return {0.0, 1.0};
        case 2: return {-1.0, 0.0};
        default: return {0.0, -1.0};
    }
}

// Phi_n^m(v), m >= 0 (see fmm_kernels.hip), into phi[hidx(n, m)].
void phi_all(int p, const double v[3], std::vector<cplx> &phi) {
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const cplx w(v[0], v[1]);
    phi.assign((p + 1) * (p + 1), 0.0);
    cplx pmm = 1.0;
    for (int m = 0; m <= p; m++) {
        if (m > 0) pmm = pmm * w / (2.0 * m);
        cplx prev2 = 0.0, prev1 = pmm;
        for (int n = m; n <= p; n++) {
            cplx cur;
            if (n == m) cur = pmm;
            else if (n == m + 1) cur = v[2] * pmm;
            else cur = ((2.0 * n - 1.0) * v[2] * prev1 - r2 * prev2) / ((double)n * n - (double)m * m);
            if (n > m) {
                prev2 = prev1;
                prev1 = cur;
            }
            phi[hidx(n, m)] = cur;
        }
    }
}

// Psi_n^m(v), m >= 0.
void psi_all(int p, const double v[3], std::vector<cplx> &psi) {
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2], ir2 = 1.0 / r2;
    const cplx w(v[0], v[1]);
    psi.assign((p + 1) * (p + 1), 0.0);
    cplx pmm = 1.0 / std::sqrt(r2);
    for (int m = 0; m <= p; m++) {
        if (m > 0) pmm = pmm * (2.0 * m - 1.0) * w * ir2;
        cplx prev2 = 0.0, prev1 = pmm;
        for (int n = m; n <= p; n++) {
            cplx cur;
            if (n == m) cur = pmm;
            else if (n == m + 1) cur = (2.0 * m + 1.0) * v[2] * pmm * ir2;
            else
                cur = ((2.0 * n - 1.0) * v[2] * prev1 - (double)(n + m - 1) * (double)(n - m - 1) * prev2) * ir2;
            if (n > m) {
                prev2 = prev1;
                prev1 = cur;
            }
            psi[hidx(n, m)] = cur;
        }
    }
}

// R_n^+-m (source moments), S_n^+-m (local evaluation), T_n^+-m (transfer).
void r_all(int p, const double v[3], double2 *out) {
    std::vector<cplx> phi;
    phi_all(p, v, phi);
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            const cplx f = ipow(-m);
            const cplx a = f * std::conj(phi[hidx(n, m)]), b = f * phi[hidx(n, m)];
            out[hidx(n, m)] = make_double2(a.real(), a.imag());
            out[hidx(n, -m)] = make_double2(b.real(), b.imag());
        }
}
void s_all(int p, const double v[3], double2 *out) {
    std::vector<cplx> phi;
    phi_all(p, v, phi);
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            const cplx f = ipow(-m) * ((n & 1) ? -1.0 : 1.0);
            const cplx a = f * phi[hidx(n, m)], b = f * std::conj(phi[hidx(n, m)]);
            out[hidx(n, m)] = make_double2(a.real(), a.imag());
            out[hidx(n, -m)] = make_double2(b.real(), b.imag());
        }
}
void t_all(int p, const double v[3], double2 *out) {
    std::vector<cplx> psi;
    psi_all(p, v, psi);
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            const cplx f = ipow(m) * ((n & 1) ? -1.0 : 1.0);
            const cplx a = f * psi[hidx(n, m)], b = f * std::conj(psi[hidx(n, m)]);
            out[hidx(n, m)] = make_double2(a.real(), a.imag());
            out[hidx(n, -m)] = make_double2(b.real(), b.imag());
        }
}

}  // namespace

FmmHost fmm_build(const System &s, int p_pol, int p_static, double min_cell, double distance) {
    FmmHost f;
    const int nmm = s.nmm, npol = s.npol;
    f.p_pol = p_pol;
    f.p_static = p_static;
    f.p_tab = std::max(p_pol, p_static);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < nmm; i++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], s.cmm[3 * i + c]);
            hi[c] = std::max(hi[c], s.cmm[3 * i + c]);
        }
    double side = 0.0;
    for (int c = 0; c < 3; c++) side = std::max(side, hi[c] - lo[c]);
    side = side * (1.0 + 1e-9) + 1e-6;
    const double leaf_min = std::max(min_cell, distance);
    int L = 0;
    while (L < kFmmMaxLevels && side / (double)(1 << (L + 1)) >= leaf_min) L++;
    f.levels = L;
    f.side = side;
    for (int c = 0; c < 3; c++) f.origin[c] = lo[c] - 0.5 * (side - (hi[c] - lo[c]));
    const int ns = 1 << L;
    f.leaf_h = side / ns;
    const long nleaf = (long)ns * ns * ns;
    auto leaf_of = [&](int i) {
        int ix[3];
        for (int c = 0; c < 3; c++) {
            int k = (int)std::floor((s.cmm[3 * i + c] - f.origin[c]) / f.leaf_h);
            ix[c] = std::min(std::max(k, 0), ns - 1);
        }
        return ((long)ix[2] * ns + ix[1]) * ns + ix[0];
    };
    // MM atoms by leaf (stable counting sort)
    std::vector<long> mleaf(nmm);
    f.mm_leaf_start.assign(nleaf + 1, 0);
    for (int i = 0; i < nmm; i++) {
        mleaf[i] = leaf_of(i);
        f.mm_leaf_start[mleaf[i] + 1]++;
    }
    for (long k = 0; k < nleaf; k++) f.mm_leaf_start[k + 1] += f.mm_leaf_start[k];
    f.mm_order.assign(nmm, 0);
    f.mm_leaf.assign(nmm, 0);
    std::vector<int> mm_rank(nmm), fill(f.mm_leaf_start.begin(), f.mm_leaf_start.end() - 1);
    for (int i = 0; i < nmm; i++) {
        const int t = fill[mleaf[i]]++;
        f.mm_order[t] = i;
        f.mm_leaf[t] = (int)mleaf[i];
        mm_rank[i] = t;
    }
    // polarizable sites by leaf
    f.pol_leaf_start.assign(nleaf + 1, 0);
    for (int j = 0; j < npol; j++) f.pol_leaf_start[mleaf[s.polar_mm[j]] + 1]++;
    for (long k = 0; k < nleaf; k++) f.pol_leaf_start[k + 1] += f.pol_leaf_start[k];
    f.pol_order.assign(npol, 0);
    f.pol_leaf.assign(npol, 0);
    f.pol_mmself.assign(npol, 0);
    std::vector<int> pfill(f.pol_leaf_start.begin(), f.pol_leaf_start.end() - 1);
    for (int j = 0; j < npol; j++) {
        const long lf = mleaf[s.polar_mm[j]];
        const int t = pfill[lf]++;
        f.pol_order[t] = j;
        f.pol_leaf[t] = (int)lf;
        f.pol_mmself[t] = mm_rank[s.polar_mm[j]];
    }
    // occupancy of every level
    f.cell_off.assign(L + 2, 0);
    for (int l = 0; l <= L; l++) f.cell_off[l + 1] = f.cell_off[l] + (1L << (3 * l));
    f.cnt_pol.assign(f.cell_off[L + 1], 0);
    f.cnt_mm.assign(f.cell_off[L + 1], 0);
    for (long k = 0; k < nleaf; k++) {
        f.cnt_pol[f.cell_off[L] + k] = f.pol_leaf_start[k + 1] - f.pol_leaf_start[k];
        f.cnt_mm[f.cell_off[L] + k] = f.mm_leaf_start[k + 1] - f.mm_leaf_start[k];
    }
    for (int l = L - 1; l >= 0; l--) {
        const int n = 1 << l, nc = 2 * n;
        for (int iz = 0; iz < nc; iz++)
            for (int iy = 0; iy < nc; iy++)
                for (int ix = 0; ix < nc; ix++) {
                    const long c = ((long)iz * nc + iy) * nc + ix;
                    const long pc = ((long)(iz >> 1) * n + (iy >> 1)) * n + (ix >> 1);
                    f.cnt_pol[f.cell_off[l] + pc] += f.cnt_pol[f.cell_off[l + 1] + c];
                    f.cnt_mm[f.cell_off[l] + pc] += f.cnt_mm[f.cell_off[l + 1] + c];
                }
    }
    // translation tables, indexed by the parent (M2M, L2L) or the cell (M2L) level
    const int pt = f.p_tab;
    f.ncoef_tab = (pt + 1) * (pt + 1);
    f.ncoef2_tab = (2 * pt + 1) * (2 * pt + 1);
    f.tab_m2m.assign((size_t)(L + 1) * 8 * f.ncoef_tab, make_double2(0.0, 0.0));
    f.tab_l2l.assign((size_t)(L + 1) * 8 * f.ncoef_tab, make_double2(0.0, 0.0));
    f.tab_m2l.assign((size_t)(L + 1) * 343 * f.ncoef2_tab, make_double2(0.0, 0.0));
    for (int l = 0; l <= L; l++) {
        const double h = side / (double)(1 << l);
        if (l < L)
            for (int oct = 0; oct < 8; oct++) {
                const double v[3] = {((oct & 1) ? 0.25 : -0.25) * h, (((oct >> 1) & 1) ? 0.25 : -0.25) * h,
                                     ((oct >> 2) ? 0.25 : -0.25) * h};
                r_all(pt, v, &f.tab_m2m[((size_t)l * 8 + oct) * f.ncoef_tab]);
                s_all(pt, v, &f.tab_l2l[((size_t)l * 8 + oct) * f.ncoef_tab]);
            }
        if (l >= 2)
            for (int dz = -3; dz <= 3; dz++)
                for (int dy = -3; dy <= 3; dy++)
                    for (int dx = -3; dx <= 3; dx++) {
                        if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                        const double v[3] = {dx * h, dy * h, dz * h};
                        const int o = (dz + 3) * 49 + (dy + 3) * 7 + (dx + 3);
                        t_all(2 * pt, v, &f.tab_m2l[((size_t)l * 343 + o) * f.ncoef2_tab]);
                    }
    }
    return f;
}

}  // namespace ommp
