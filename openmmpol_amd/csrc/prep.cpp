// Host-side construction of a prepared system: .mmp reader, multipole
// rotation, connectivity levels, polarization groups and screening lists.
// All of it is O(N * neighbours) and runs once per system.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "ommp_internal.h"

namespace ommp {

namespace {

// Default scale factors (mod_constants.F90:37-52), order m, p, d, u, intra.
const double kScalesAmoeba[5][4] = {{0.0, 0.0, 0.4, 0.8}, {0.0, 0.0, 1.0, 1.0}, {0.0, 1.0, 1.0, 1.0},
                                     {1.0, 1.0, 1.0, 1.0}, {0.0, 0.0, 0.5, 1.0}};
const double kScalesWangAL[4][4] = {{0.0, 0.0, 1.0, 1.0}, {0.0, 0.0, 1.0, 1.0}, {0.0, 0.0, 1.0, 1.0},
                                     {0.0, 0.0, 1.0, 1.0}};
const double kTholeScaleWangAL = 2.5874;  // mod_constants.F90:55
const double kTholeScaleWangDL = 2.0580;  // mod_constants.F90:57

// Fortran list-directed READ over lines: each read starts on a fresh record
// and may continue over the following ones.
class ListReader {
  public:
    explicit ListReader(std::vector<std::string> lines) : lines_(std::move(lines)) {}
    void skip(size_t n) { pos_ += n; }
    std::vector<std::string> read(size_t k) {
        std::vector<std::string> out;
        while (out.size() < k) {
            if (pos_ >= lines_.size()) fatal("Unexpected end of MMP file");
            std::string line = lines_[pos_++];
            for (auto &c : line)
                if (c == ',') c = ' ';
            std::istringstream is(line);
            std::string tok;
            while (is >> tok) out.push_back(tok);
        }
        out.resize(k);
        return out;
    }
    double real() { return to_real(read(1)[0]); }
    long integer() { return std::stol(read(1)[0]); }
    static double to_real(std::string t) {
        for (auto &c : t)
            if (c == 'd' || c == 'D') c = 'E';
        return std::stod(t);
    }

  private:
    std::vector<std::string> lines_;
    size_t pos_ = 0;
};

void rows_to_csr(const std::vector<std::vector<std::pair<int, double>>> &rows, Csr &out, bool keep_val) {
    out.n = (int)rows.size();
    out.ri.assign(out.n + 1, 0);
    for (int i = 0; i < out.n; i++) out.ri[i + 1] = out.ri[i] + (int)rows[i].size();
    out.ci.resize(out.ri[out.n]);
    out.val.clear();
    if (keep_val) out.val.resize(out.ri[out.n]);
    for (int i = 0; i < out.n; i++)
        for (size_t k = 0; k < rows[i].size(); k++) {
            out.ci[out.ri[i] + k] = rows[i][k].first;
            if (keep_val) out.val[out.ri[i] + k] = rows[i][k].second;
        }
}

// Exact shortest-path distance levels.  The reference computes
// C_n = C_{n-1} C_1 andnot (C_{n-1} or C_{n-2}) with sparse boolean products
// (mod_adjacency_mat.F90:367-422), i.e. the nodes at distance exactly n; a
// bounded breadth-first search from every node yields the same sets in
// O(N * deg^n).  Level 0 (identity) is included when start_id.
void distance_levels(const Csr &adj, int nlev, bool start_id, Csr *out) {
    const int n = adj.n;
    const int nout = start_id ? nlev + 1 : nlev;
    std::vector<std::vector<std::vector<std::pair<int, double>>>> rows(
        nout, std::vector<std::vector<std::pair<int, double>>>(n));
    std::vector<int> dist(n, -1), queue;
    queue.reserve(64);
    for (int s = 0; s < n; s++) {
        queue.clear();
        queue.push_back(s);
        dist[s] = 0;
        for (size_t h = 0; h < queue.size(); h++) {
            int u = queue[h];
            if (dist[u] >= nlev) continue;
            for (int k = adj.ri[u]; k < adj.ri[u + 1]; k++) {
                int v = adj.ci[k];
                if (dist[v] < 0) {
                    dist[v] = dist[u] + 1;
                    queue.push_back(v);
                }
            }
        }
        for (int v : queue) {
            int d = dist[v];
            if (start_id) rows[d][s].push_back({v, 1.0});
            else if (d >= 1) rows[d - 1][s].push_back({v, 1.0});
            dist[v] = -1;
        }
    }
    for (int l = 0; l < nout; l++) {
        for (auto &r : rows[l]) std::sort(r.begin(), r.end());
        rows_to_csr(rows[l], out[l], false);
    }
}

bool row_has(const Csr &c, int row, int col) {
    for (int k = c.ri[row]; k < c.ri[row + 1]; k++)
        if (c.ci[k] == col) return true;
    return false;
}

// Polarization-group distance between the groups of MM atoms i and j, -1 if > 3.
int group_distance(const System &s, int i, int j) {
    int gi = s.grp[i], gj = s.grp[j];
    for (int l = 0; l < 4; l++)
        if (row_has(s.gconn[l], gi, gj)) return l;
    return -1;
}

}  // namespace

void set_default_scales(System &s) {
    if (s.amoeba) {
        std::memcpy(s.mscale, kScalesAmoeba[0], sizeof(s.mscale));
        std::memcpy(s.pscale, kScalesAmoeba[1], sizeof(s.pscale));
        std::memcpy(s.dscale, kScalesAmoeba[2], sizeof(s.dscale));
        std::memcpy(s.uscale, kScalesAmoeba[3], sizeof(s.uscale));
        std::memcpy(s.pscale_intra, kScalesAmoeba[4], sizeof(s.pscale_intra));
        s.thole_scale = 0.0;
    } else if (s.ff_rules == 1) {  // Wang DL: everything 1
        for (int k = 0; k < 4; k++)
            s.mscale[k] = s.pscale[k] = s.dscale[k] = s.uscale[k] = s.pscale_intra[k] = 1.0;
        s.thole_scale = kTholeScaleWangDL;
    } else {  // Wang AL
        std::memcpy(s.mscale, kScalesWangAL[0], sizeof(s.mscale));
        std::memcpy(s.pscale, kScalesWangAL[1], sizeof(s.pscale));
        std::memcpy(s.dscale, kScalesWangAL[2], sizeof(s.dscale));
        std::memcpy(s.uscale, kScalesWangAL[3], sizeof(s.uscale));
        for (int k = 0; k < 4; k++) s.pscale_intra[k] = 1.0;
        s.thole_scale = kTholeScaleWangAL;
    }
}

// .mmp revisions 2 and 3 (mod_inputloader.F90:15-257).
System *read_mmp(const std::string &path) {
    std::ifstream f(path);
    if (!f.good()) fatal("Input file '" + path + "' does not exist.");
    std::vector<std::string> lines;
    for (std::string line; std::getline(f, line);) lines.push_back(line);
    ListReader rd(std::move(lines));
    long rev = rd.integer();
    if (rev != 2 && rev != 3) fatal("MMP file is only supported with version 2 or 3.");
    rd.skip(2);
    long ff_type = rd.integer();
    long ff_rules = rd.integer();
    rd.skip(rev == 3 ? 17 : 15);
    long nmm = rd.integer();
    const int ld = ff_type == 1 ? 10 : 1;
    rd.skip(nmm + 1);  // a zero and the atomic numbers
    std::vector<double> cmm(3 * nmm), q(ld * nmm), pol(nmm);
    for (long i = 0; i < nmm; i++) {
        auto t = rd.read(3);
        for (int c = 0; c < 3; c++) cmm[3 * i + c] = ListReader::to_real(t[c]) * kAng2Au;
    }
    rd.skip(nmm);  // residue numbers
    for (long i = 0; i < nmm; i++) {
        auto t = rd.read(ld);
        for (int c = 0; c < ld; c++) q[ld * i + c] = ListReader::to_real(t[c]);
    }
    for (long i = 0; i < nmm; i++) {
        double p = rd.real();
        // mod_inputloader.F90:159-172 (threshold on the file value), :205-207 (AMOEBA A^3 -> bohr^3)
        if (!(p > 1e-8)) p = 0.0;
        if (ff_type == 1) p *= kAng2Au * kAng2Au * kAng2Au;
        pol[i] = p;
    }
    std::vector<int> adj_ri(nmm + 1, 0), adj_ci;
    {
        std::vector<std::vector<int>> nb(nmm);
        for (long i = 0; i < nmm; i++) {
            auto t = rd.read(8);
            for (auto &v : t) {
                long j = std::stol(v);
                if (j != 0) nb[i].push_back((int)j - 1);
            }
            std::sort(nb[i].begin(), nb[i].end());
        }
        for (long i = 0; i < nmm; i++) {
            adj_ri[i + 1] = adj_ri[i] + (int)nb[i].size();
            adj_ci.insert(adj_ci.end(), nb[i].begin(), nb[i].end());
        }
    }
    std::vector<int> grp, frames;
    if (ff_type == 1) {
        // polgroup11_to_mm2pg (mod_inputloader.F90:259-299)
        grp.assign(nmm, -1);
        int g = 0;
        std::vector<std::vector<int>> members(nmm);
        for (long i = 0; i < nmm; i++) {
            auto t = rd.read(120);
            for (auto &v : t) {
                long m = std::stol(v);
                if (m == 0) break;
                members[i].push_back((int)m - 1);
            }
        }
        for (long i = 0; i < nmm; i++) {
            if (grp[i] >= 0) continue;
            for (int m : members[i]) {
                if (m < 0 || m >= nmm || grp[m] >= 0) fatal("Unexpected error in polgroup11_to_mm2pg");
                grp[m] = g;
            }
            g++;
        }
        for (long i = 0; i < nmm; i++)
            if (grp[i] < 0) fatal("Unexpected error in polgroup11_to_mm2pg");
        frames.resize(4 * nmm);
        for (long i = 0; i < nmm; i++) {
            auto t = rd.read(4);
            for (int c = 0; c < 4; c++) frames[4 * i + c] = (int)std::stol(t[c]);
        }
        message("With deprecated .mmp input, thole damping factors are assumed to be 0.39", 2);
    }
    return build_from_arrays((int)ff_type == 1 ? 1 : 0, (int)ff_rules, (int)nmm, cmm.data(), q.data(),
                             pol.data(), nullptr, adj_ri.data(), adj_ci.empty() ? nullptr : adj_ci.data(),
                             grp.empty() ? nullptr : grp.data(), frames.empty() ? nullptr : frames.data());
}

System *build_from_arrays(int ff_type, int ff_rules, int nmm, const double *cmm, const double *q,
                          const double *pol, const double *thole_a, const int *adj_ri, const int *adj_ci,
                          const int *pgroup, const int *frames) {
    if (ff_type != 1 && ff_type != 0) fatal("Cannot initialize an UNKNOWN forcefield!");
    if (nmm < 0) fatal("Number of MM atoms should be non-negative");
    System *sp = new System();
    System &s = *sp;
    s.amoeba = ff_type == 1;
    s.ff_rules = ff_rules;
    s.nmm = nmm;
    s.ld_cart = s.amoeba ? 10 : 1;
    s.n_ipd = s.amoeba ? 2 : 1;
    s.cmm.assign(cmm, cmm + 3 * (size_t)nmm);
    s.atz.assign(nmm, 0);
    s.q0.assign(q, q + (size_t)s.ld_cart * nmm);
    if (s.amoeba)
        for (int i = 0; i < nmm; i++)
            for (int c = 4; c < 10; c++) s.q0[10 * i + c] /= 3.0;  // mod_mmpol.F90:209-215
    s.q = s.q0;
    s.thole_damping_mm.assign(nmm, 0.39);
    if (thole_a)
        for (int i = 0; i < nmm; i++) s.thole_damping_mm[i] = thole_a[i];
    // polarizable sites (remove_null_pol keeps |pol| > eps_rp, mod_electrostatics.F90:469-536)
    s.mm_polar.assign(nmm, -1);
    for (int i = 0; i < nmm; i++)
        if (std::fabs(pol[i]) > kEpsRp) {
            s.mm_polar[i] = (int)s.polar_mm.size();
            s.polar_mm.push_back(i);
            s.pol.push_back(pol[i]);
        }
    s.npol = (int)s.polar_mm.size();
    s.adj.n = nmm;
    s.adj.ri.assign(adj_ri, adj_ri + nmm + 1);
    s.adj.ci.assign(adj_ci ? adj_ci : adj_ri, (adj_ci ? adj_ci : adj_ri) + (adj_ci ? adj_ri[nmm] : 0));
    for (int i = 0; i < nmm; i++)
        for (int k = s.adj.ri[i]; k < s.adj.ri[i + 1]; k++)
            if (s.adj.ci[k] < 0 || s.adj.ci[k] >= nmm) fatal("Connectivity index out of range");
    if (s.amoeba) {
        if (!pgroup) fatal("Polarization groups are required for AMOEBA systems");
        s.grp.assign(pgroup, pgroup + nmm);
        s.mol_frame.assign(nmm, 0);
        s.ix.assign(nmm, -1);
        s.iy.assign(nmm, -1);
        s.iz.assign(nmm, -1);
        if (frames)
            for (int i = 0; i < nmm; i++) {
                s.mol_frame[i] = frames[4 * i];
                s.iz[i] = frames[4 * i + 1] - 1;
                s.ix[i] = frames[4 * i + 2] - 1;
                s.iy[i] = frames[4 * i + 3] - 1;
                if (s.mol_frame[i] < 0 || s.mol_frame[i] > 5) fatal("Unknown molecular frame convention");
            }
    }
    set_default_scales(s);
    prepare(s);
    return sp;
}

// mmpol_prepare (mod_mmpol.F90:139-237): connectivity, maps, Thole factors,
// polarization-group graph, rotation, screening lists.
void prepare(System &s) {
    distance_levels(s.adj, 4, false, s.conn);
    s.cpol.resize(3 * (size_t)s.npol);
    for (int p = 0; p < s.npol; p++)
        for (int c = 0; c < 3; c++) s.cpol[3 * p + c] = s.cmm[3 * s.polar_mm[p] + c];
    // thole_init (mod_electrostatics.F90:866-897)
    s.thole.assign(s.nmm, 0.0);
    for (int p = 0; p < s.npol; p++) s.thole[s.polar_mm[p]] = std::pow(s.pol[p], 1.0 / 6.0);
    if (!s.amoeba) {
        if (s.thole_scale > kEpsRp) {
            double f = std::sqrt(s.thole_scale);
            for (auto &t : s.thole) t *= f;
        } else {
            fatal("Scale factor for Thole damping should be greater than 0.0 when non-AMOEBA FF are used");
        }
    }
    if (s.amoeba) {
        int ng = 0;
        for (int i = 0; i < s.nmm; i++) {
            if (s.grp[i] < 0) fatal("Unexpected error in polgroup11_to_mm2pg");
            ng = std::max(ng, s.grp[i] + 1);
        }
        s.ngrp = ng;
        // reverse_grp_tab (mod_adjacency_mat.F90:424-493)
        std::vector<std::vector<std::pair<int, double>>> rows(ng);
        for (int i = 0; i < s.nmm; i++) rows[s.grp[i]].push_back({i, 1.0});
        rows_to_csr(rows, s.grp_members, false);
        // build_pg_adjacency_matrix (mod_mmpol.F90:272-323)
        std::vector<std::vector<std::pair<int, double>>> grows(ng);
        for (int g = 0; g < ng; g++)
            for (int m = s.grp_members.ri[g]; m < s.grp_members.ri[g + 1]; m++) {
                int a1 = s.grp_members.ci[m];
                for (int k = s.adj.ri[a1]; k < s.adj.ri[a1 + 1]; k++) {
                    int g2 = s.grp[s.adj.ci[k]];
                    if (g2 == g) continue;
                    bool dup = false;
                    for (auto &e : grows[g]) dup |= e.first == g2;
                    if (!dup) grows[g].push_back({g2, 1.0});
                }
            }
        Csr gadj;
        rows_to_csr(grows, gadj, false);
        distance_levels(gadj, 3, true, s.gconn);
        rotate_multipoles(s);
    }
    make_screening_lists(s);
    s.ipd.assign((size_t)s.n_ipd * s.npol * 3, 0.0);
    s.e_m2d.assign((size_t)s.n_ipd * s.npol * 3, 0.0);
}

// ommp_turn_pol_off (mod_interface.F90:134-168) and the remove_null_pol it
// ends with (mod_electrostatics.F90:469-536): the listed MM atoms (1-based)
// lose their polarizability, the polarizable set is compacted in order and the
// dipoles are dropped (zero, no guess).  The Thole factors keep their values
// (thole_init does not run again), so a switched-off atom's pairs stay damped
// as sources, while its damping parameter a no longer enters (it is not
// polarizable in the combination rule).  Two points where the reference is not
// self-consistent are resolved: its loop reads mm_polar at the loop counter
// (:156) rather than at the listed atom, the same atom for the lists 1..n its
// callers pass; and it keeps the screening lists built for the old polarizable
// numbering, which are rebuilt here.  The device state is released and built
// again for the new set at the next use.
void turn_pol_off(System &s, int n, const int *nopol) {
    if (n <= 0) return;
    const int mn = *std::min_element(nopol, nopol + n), mx = *std::max_element(nopol, nopol + n);
    if (mn < 1) fatal("Atom indexes are 1-based, so no index should be less than 1");
    if (mx > s.nmm) fatal("Atom indexes should be inside the topology range");
    for (int k = 0; k < n; k++) {
        const int p = s.mm_polar[nopol[k] - 1];
        if (p >= 0) s.pol[p] = 0.0;
        else message("Atom " + std::to_string(nopol[k]) + " is not polarizable. Ignoring.", 1);
    }
    std::vector<int> keep;
    for (int p = 0; p < s.npol; p++)
        if (std::fabs(s.pol[p]) > kEpsRp) keep.push_back(p);
    if ((int)keep.size() == s.npol) return;
    message("Removing " + std::to_string(s.npol - (int)keep.size()) + " polarizabilities out of " +
                std::to_string(s.npol), 1);
    std::vector<double> pol;
    std::vector<int> polar_mm;
    for (int p : keep) {
        pol.push_back(s.pol[p]);
        polar_mm.push_back(s.polar_mm[p]);
    }
    s.pol.swap(pol);
    s.polar_mm.swap(polar_mm);
    s.npol = (int)s.polar_mm.size();
    s.mm_polar.assign(s.nmm, -1);
    for (int p = 0; p < s.npol; p++) s.mm_polar[s.polar_mm[p]] = p;
    s.cpol.resize(3 * (size_t)s.npol);
    for (int p = 0; p < s.npol; p++)
        for (int c = 0; c < 3; c++) s.cpol[3 * p + c] = s.cmm[3 * s.polar_mm[p] + c];
    make_screening_lists(s);
    device_release(s);
    s.dev_geometry_dirty = true;
    s.ipd.assign((size_t)s.n_ipd * s.npol * 3, 0.0);
    s.e_m2d.assign((size_t)s.n_ipd * s.npol * 3, 0.0);
    s.m2d_done = s.ipd_done = s.ipd_use_guess = false;
    s.host_ipd_stale = false;
    s.host_m2d_stale = true;
    for (int &k : s.pred_iters) k = 0;
    s.pred_outer = 0;
    message("Removing null polarizable sites done", 1);
}

// rotation_matrix without derivatives (rotate_multipoles.F90:298-463) and
// rotate_multipoles (:232-295).
void rotate_multipoles(System &s) {
    if (!s.amoeba) return;
    for (int j = 0; j < s.nmm; j++) {
        double r[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const int frame = s.mol_frame[j];
        if (frame != 0) {
            const int jz = s.iz[j] >= 0 ? s.iz[j] : j;
            const int jx = s.ix[j] >= 0 ? s.ix[j] : j;
            const int jy = s.iy[j] >= 0 ? s.iy[j] : j;
            const double *c = &s.cmm[3 * j];
            double xi[3], eta[3], zeta[3], u[3] = {0, 0, 0}, v[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                xi[k] = s.cmm[3 * jz + k] - c[k];
                eta[k] = s.cmm[3 * jx + k] - c[k];
                zeta[k] = s.cmm[3 * jy + k] - c[k];
            }
            auto nrm = [](const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
            const double xn = nrm(xi), en = nrm(eta), zn = nrm(zeta);
            switch (frame) {
                case 1:  // z-then-x
                    for (int k = 0; k < 3; k++) { u[k] = xi[k]; v[k] = eta[k]; }
                    break;
                case 2:  // bisector
                    for (int k = 0; k < 3; k++) { u[k] = en * xi[k] + xn * eta[k]; v[k] = eta[k]; }
                    break;
                case 3: {  // z-only
                    for (int k = 0; k < 3; k++) u[k] = xi[k];
                    double dot = u[2] / xn;
                    if (dot <= 0.866 && std::fabs(u[1]) > kEpsRp) v[0] = 1.0;
                    else v[1] = 1.0;
                    break;
                }
                case 4:  // z-bisector
                    for (int k = 0; k < 3; k++) { u[k] = xi[k]; v[k] = zn * eta[k] + en * zeta[k]; }
                    break;
                case 5:  // 3-fold
                    for (int k = 0; k < 3; k++) {
                        u[k] = en * zn * xi[k] + xn * zn * eta[k] + xn * en * zeta[k];
                        v[k] = eta[k];
                    }
                    break;
            }
            if (nrm(u) < kEpsRp || nrm(v) < kEpsRp)
                fatal("Ill defined rotation axis (either u or v vectors have 0 norm.");
            double un = nrm(u), ez[3], ex[3], ey[3];
            for (int k = 0; k < 3; k++) ez[k] = u[k] / un;
            double vez = v[0] * ez[0] + v[1] * ez[1] + v[2] * ez[2];
            for (int k = 0; k < 3; k++) ex[k] = v[k] - vez * ez[k];
            double exn = nrm(ex);
            for (int k = 0; k < 3; k++) ex[k] /= exn;
            ey[0] = ez[1] * ex[2] - ez[2] * ex[1];
            ey[1] = ez[2] * ex[0] - ez[0] * ex[2];
            ey[2] = ez[0] * ex[1] - ez[1] * ex[0];
            for (int k = 0; k < 3; k++) {
                r[k][0] = ex[k];
                r[k][1] = ey[k];
                r[k][2] = ez[k];
            }
        }
        const double *a = &s.q0[10 * j];
        double *b = &s.q[10 * j];
        b[0] = a[0];
        for (int k = 0; k < 3; k++) b[1 + k] = r[k][0] * a[1] + r[k][1] * a[2] + r[k][2] * a[3];
        // storage xx, xy, yy, xz, yz, zz (f_cart_components.h:13-18)
        const double Q[3][3] = {{a[4], a[5], a[7]}, {a[5], a[6], a[8]}, {a[7], a[8], a[9]}};
        double RQ[3][3], RQR[3][3];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) RQ[i][k] = r[i][0] * Q[0][k] + r[i][1] * Q[1][k] + r[i][2] * Q[2][k];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) RQR[i][k] = RQ[i][0] * r[k][0] + RQ[i][1] * r[k][1] + RQ[i][2] * r[k][2];
        b[4] = RQR[0][0];
        b[5] = RQR[0][1];
        b[6] = RQR[1][1];
        b[7] = RQR[0][2];
        b[8] = RQR[1][2];
        b[9] = RQR[2][2];
    }
}

// make_screening_lists, direct branch (mod_electrostatics.F90:538-864), with
// the scale of every entry given by screening_rules (:3338-3487).  Only
// factors different from 1 are stored.
void make_screening_lists(System &s) {
    const int n = s.nmm, npol = s.npol;
    std::vector<std::vector<std::pair<int, double>>> rows;
    auto keep = [](double sc) { return std::fabs(sc - 1.0) > kEpsRp; };
    // static -> static: m-scale by bond distance (:620-656)
    rows.assign(n, {});
    for (int i = 0; i < n; i++)
        for (int l = 0; l < 4; l++)
            for (int k = s.conn[l].ri[i]; k < s.conn[l].ri[i + 1]; k++)
                if (keep(s.mscale[l])) rows[i].push_back({s.conn[l].ci[k], s.mscale[l]});
    rows_to_csr(rows, s.list_S_S, true);
    // static -> polarizable, polarization field ('P'): bond distance
    rows.assign(n, {});
    for (int i = 0; i < n; i++)
        for (int l = 0; l < 4; l++)
            for (int k = s.conn[l].ri[i]; k < s.conn[l].ri[i + 1]; k++) {
                const int j = s.conn[l].ci[k], jp = s.mm_polar[j];
                if (jp < 0) continue;
                double sc = (s.amoeba && s.grp[i] == s.grp[j]) ? s.pscale_intra[l] : s.pscale[l];
                if (keep(sc)) rows[i].push_back({jp, sc});
            }
    rows_to_csr(rows, s.list_S_P_P, true);
    // polarizable -> polarizable: the list scans bond neighbours only
    rows.assign(npol, {});
    for (int ip = 0; ip < npol; ip++) {
        const int i = s.polar_mm[ip];
        for (int l = 0; l < 4; l++)
            for (int k = s.conn[l].ri[i]; k < s.conn[l].ri[i + 1]; k++) {
                const int j = s.conn[l].ci[k], jp = s.mm_polar[j];
                if (jp < 0) continue;
                double sc;
                if (s.amoeba) {
                    int g = group_distance(s, i, j);
                    sc = g < 0 ? 1.0 : s.uscale[g];
                } else {
                    sc = s.uscale[l];
                }
                if (keep(sc)) rows[ip].push_back({jp, sc});
            }
    }
    rows_to_csr(rows, s.list_P_P, true);
    if (s.amoeba) {
        // static -> polarizable, direct field ('D'): polarization-group distance
        rows.assign(n, {});
        for (int i = 0; i < n; i++) {
            const int pg = s.grp[i];
            for (int l = 0; l < 4; l++)
                for (int k = s.gconn[l].ri[pg]; k < s.gconn[l].ri[pg + 1]; k++) {
                    const int g = s.gconn[l].ci[k];
                    for (int m = s.grp_members.ri[g]; m < s.grp_members.ri[g + 1]; m++) {
                        const int jp = s.mm_polar[s.grp_members.ci[m]];
                        if (jp < 0) continue;
                        if (keep(s.dscale[l])) rows[i].push_back({jp, s.dscale[l]});
                    }
                }
        }
        rows_to_csr(rows, s.list_S_P_D, true);
        // dense T: screening_rules on every pair = group distance for AMOEBA
        rows.assign(npol, {});
        for (int ip = 0; ip < npol; ip++) {
            const int pg = s.grp[s.polar_mm[ip]];
            for (int l = 0; l < 4; l++) {
                if (!keep(s.uscale[l])) continue;
                for (int k = s.gconn[l].ri[pg]; k < s.gconn[l].ri[pg + 1]; k++) {
                    const int g = s.gconn[l].ci[k];
                    for (int m = s.grp_members.ri[g]; m < s.grp_members.ri[g + 1]; m++) {
                        const int jp = s.mm_polar[s.grp_members.ci[m]];
                        if (jp < 0 || jp == ip) continue;
                        rows[ip].push_back({jp, s.uscale[l]});
                    }
                }
            }
        }
        rows_to_csr(rows, s.list_P_P_dense, true);
    } else {
        s.list_S_P_D = Csr();
        s.list_S_P_D.n = n;
        s.list_S_P_D.ri.assign(n + 1, 0);
        s.list_P_P_dense = s.list_P_P;
    }
}

// coulomb_kernel refuses two sites closer than eps_rp (mod_electrostatics.F90:1003-1008);
// every pair with a polarizable end is evaluated, so check those up front with a grid.
void check_coincident_sites(const System &s) {
    if (s.npol == 0) return;
    struct Key {
        long x, y, z;
        bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct Hash {
        size_t operator()(const Key &k) const {
            return (size_t)(k.x * 73856093L) ^ (size_t)(k.y * 19349663L) ^ (size_t)(k.z * 83492791L);
        }
    };
    std::unordered_map<Key, std::vector<int>, Hash> grid;
    grid.reserve(2 * s.nmm);
    auto key = [&](int i) {
        return Key{(long)std::floor(s.cmm[3 * i]), (long)std::floor(s.cmm[3 * i + 1]),
                   (long)std::floor(s.cmm[3 * i + 2])};
    };
    for (int i = 0; i < s.nmm; i++) grid[key(i)].push_back(i);
    for (int p = 0; p < s.npol; p++) {
        const int i = s.polar_mm[p];
        Key k = key(i);
        for (long dx = -1; dx <= 1; dx++)
            for (long dy = -1; dy <= 1; dy++)
                for (long dz = -1; dz <= 1; dz++) {
                    auto it = grid.find(Key{k.x + dx, k.y + dy, k.z + dz});
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        if (j == i) continue;
                        double d2 = 0;
                        for (int c = 0; c < 3; c++) {
                            double d = s.cmm[3 * i + c] - s.cmm[3 * j + c];
                            d2 += d * d;
                        }
                        if (std::sqrt(d2) < kEpsRp)
                            fatal("Requesting Coulomb kernel for two atoms placed in the same point, this "
                                  "could be an internal bug or a problem in your input file, please check.");
                    }
                }
    }
}

}  // namespace ommp
