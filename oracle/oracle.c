/*
 * oracle.c -- CPU FP64 restatement of OpenMMPol's induced-point-dipole path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file holds the reference results the GPU library is
 * compared against.  It is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py -- never by the product library, which has
 * no CPU fallback.
 *
 * Every routine cites the reference routine it restates (paths relative to the
 * OpenMMPol source tree).  Arithmetic is plain FP64, loop order follows the
 * reference so that CG/DIIS iterates track it to rounding.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - damped kernels, fields, M2D RHS, OTF matvec, CG, energy: the reference's
 *     own golden IPD_*.ref / ENE_*.ref fixtures (N-methylacetamide, 1CRN) and
 *     closed forms (single site, two-site 6x6 solve, traceless/symmetric T,
 *     finite differences of the potential);
 *   - inversion and Jacobi/DIIS: same solution as CG and as numpy's dense
 *     solve (the reference's own tests never select them, so those two are
 *     pinned by the mathematics only);
 *   - screening lists: the reference's summary_*.ref fixtures for the
 *     topology levels they print, plus brute-force rules on tiny systems.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS_RP (2.220446049250313080847e-16 * 100.0) /* mod_constants.F90:61 */

enum { SOLVER_NONE = 0, SOLVER_CG = 1, SOLVER_DIIS = 2, SOLVER_INVERSION = 3 };
enum { MATV_NONE = 0, MATV_INCORE = 1, MATV_DIRECT = 2 };

typedef struct {
    int n, nnz;
    int *ri; /* n+1, 0-based */
    int *ci;
    double *val; /* optional */
} csr_t;

typedef struct {
    int nmm, npol, amoeba, ld_cart, n_ipd;
    double *cmm;          /* [nmm][3] bohr */
    double *q;            /* [nmm][ld_cart] lab-frame multipoles */
    int *polar_mm;        /* [npol] */
    int *mm_polar;        /* [nmm], -1 when not polarizable */
    double *pol;          /* [npol] bohr^3 */
    double *thole;        /* [nmm] */
    double *thole_damping;/* [npol] */
    csr_t conn[4];        /* exact bond distance 1..4 (1-2 .. 1-5) */
    int ngrp;
    int *grp;             /* [nmm] */
    csr_t grp_members;    /* group -> MM atoms */
    csr_t gconn[4];       /* group distance 0..3 */
    double mscale[4], pscale[4], dscale[4], uscale[4], pscale_intra[4];
    /* screening lists, row = source (reference layout) */
    csr_t list_S_P_P, list_S_P_D, list_P_P, list_S_S;
    /* transposed (row = target) for O(1) lookups in the pair loops */
    csr_t tS_P_P, tS_P_D, tP_P;
    int fatal;            /* set instead of exit(1) so tests can observe it */
    char msg[256];
} osys;

static void oracle_fatal(osys *s, const char *m)
{
    if (!s->fatal) {
        s->fatal = 1;
        snprintf(s->msg, sizeof(s->msg), "%s", m);
    }
}

/* ------------------------------------------------------------------ CSR */
static void csr_free(csr_t *c)
{
    free(c->ri); free(c->ci); free(c->val);
    memset(c, 0, sizeof(*c));
}

/* Build CSR from per-row dynamic arrays. */
typedef struct { int n, cap; int *idx; double *val; } rowbuf_t;

static void rb_push(rowbuf_t *r, int i, double v)
{
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 8;
        r->idx = (int *)realloc(r->idx, sizeof(int) * r->cap);
        r->val = (double *)realloc(r->val, sizeof(double) * r->cap);
    }
    r->idx[r->n] = i; r->val[r->n] = v; r->n++;
}

static void rows_to_csr(int n, rowbuf_t *rows, csr_t *out, int with_val)
{
    out->n = n;
    out->ri = (int *)malloc(sizeof(int) * (n + 1));
    out->ri[0] = 0;
    for (int i = 0; i < n; i++) out->ri[i + 1] = out->ri[i] + rows[i].n;
    out->nnz = out->ri[n];
    out->ci = (int *)malloc(sizeof(int) * (out->nnz ? out->nnz : 1));
    out->val = with_val ? (double *)malloc(sizeof(double) * (out->nnz ? out->nnz : 1)) : NULL;
    for (int i = 0; i < n; i++) {
        memcpy(out->ci + out->ri[i], rows[i].idx, sizeof(int) * rows[i].n);
        if (with_val) memcpy(out->val + out->ri[i], rows[i].val, sizeof(double) * rows[i].n);
        free(rows[i].idx); free(rows[i].val);
    }
    free(rows);
}

static void csr_transpose(const csr_t *a, int ncol, csr_t *t)
{
    rowbuf_t *rows = (rowbuf_t *)calloc(ncol, sizeof(rowbuf_t));
    for (int i = 0; i < a->n; i++)
        for (int k = a->ri[i]; k < a->ri[i + 1]; k++)
            rb_push(&rows[a->ci[k]], i, a->val ? a->val[k] : 1.0);
    rows_to_csr(ncol, rows, t, 1);
}

/*
 * Exact graph distance levels 1..nlev from an adjacency CSR.  The reference
 * builds C_n = C_{n-1} C_1 andnot (C_{n-1} or C_{n-2}) (mod_adjacency_mat.F90
 * build_conn_upto_n :367-422), which is exactly the set of nodes at shortest
 * path distance n; a breadth-first search gives the same sets.  With
 * start_id the identity is level 0 (mod_mmpol.F90:221).
 */
static void conn_levels(const csr_t *adj, int nlev, int start_id, csr_t *out)
{
    int n = adj->n;
    int nout = start_id ? nlev + 1 : nlev;
    rowbuf_t **rows = (rowbuf_t **)malloc(sizeof(rowbuf_t *) * nout);
    for (int l = 0; l < nout; l++) rows[l] = (rowbuf_t *)calloc(n, sizeof(rowbuf_t));
    int *dist = (int *)malloc(sizeof(int) * n);
    int *queue = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) dist[i] = -1;
    for (int s = 0; s < n; s++) {
        int qh = 0, qt = 0;
        queue[qt++] = s; dist[s] = 0;
        while (qh < qt) {
            int u = queue[qh++];
            if (dist[u] >= nlev) continue;
            for (int k = adj->ri[u]; k < adj->ri[u + 1]; k++) {
                int v = adj->ci[k];
                if (dist[v] < 0) { dist[v] = dist[u] + 1; queue[qt++] = v; }
            }
        }
        for (int k = 0; k < qt; k++) {
            int v = queue[k], d = dist[v];
            if (start_id) rb_push(&rows[d][s], v, 1.0);
            else if (d >= 1) rb_push(&rows[d - 1][s], v, 1.0);
            dist[v] = -1;
        }
    }
    for (int l = 0; l < nout; l++) {
        /* sort each row ascending (the reference rows come out of sparse products in
           ascending column order); order only matters for first-match scans */
        for (int i = 0; i < n; i++) {
            rowbuf_t *r = &rows[l][i];
            for (int a = 1; a < r->n; a++) {
                int x = r->idx[a], b = a - 1;
                while (b >= 0 && r->idx[b] > x) { r->idx[b + 1] = r->idx[b]; b--; }
                r->idx[b + 1] = x;
            }
        }
        rows_to_csr(n, rows[l], &out[l], 0);
    }
    free(rows); free(dist); free(queue);
}

static int csr_row_has(const csr_t *c, int row, int col)
{
    for (int k = c->ri[row]; k < c->ri[row + 1]; k++)
        if (c->ci[k] == col) return 1;
    return 0;
}

/* --------------------------------------------------------- screening rules */
/*
 * screening_rules (mod_electrostatics.F90:3338-3487).  kind 'P' indices are
 * polarizable indices, 'S' are MM indices; field is 'P', 'D' or '-'.
 */
static double screening_rule(osys *s, int i, char kind_i, int j, char kind_j, char in_field)
{
    int i_mm = (kind_i == 'P') ? s->polar_mm[i] : i;
    int j_mm = (kind_j == 'P') ? s->polar_mm[j] : j;
    const double *myscale, *myintra = NULL;
    char interaction, field = '-';
    if (kind_i == 'P' && kind_j == 'P') {
        myscale = s->uscale; interaction = 'P';
    } else if (kind_i == 'S' && kind_j == 'S') {
        myscale = s->mscale; interaction = 'S';
    } else {
        if (in_field == 'P' && s->amoeba) { myscale = s->pscale; myintra = s->pscale_intra; field = 'P'; }
        else if (in_field == 'D' && s->amoeba) { myscale = s->dscale; field = 'D'; }
        else if (!s->amoeba) { myscale = s->pscale; field = 'P'; }
        else { oracle_fatal(s, "Unexpected value of field in screening rules"); return 1.0; }
        interaction = 'M';
    }
    double scalf = 1.0;
    if (!s->amoeba || (interaction == 'M' && field == 'P') || interaction == 'S') {
        for (int l = 0; l < 4; l++) {
            if (csr_row_has(&s->conn[l], i_mm, j_mm)) {
                if (interaction == 'M' && s->amoeba)
                    scalf = (s->grp[i_mm] == s->grp[j_mm]) ? myintra[l] : myscale[l];
                else
                    scalf = myscale[l];
                break;
            }
        }
    } else {
        int pg_i = s->grp[i_mm];
        for (int l = 0; l < 4; l++) {
            int found = 0;
            for (int k = s->gconn[l].ri[pg_i]; k < s->gconn[l].ri[pg_i + 1] && !found; k++) {
                int g = s->gconn[l].ci[k];
                if (csr_row_has(&s->grp_members, g, j_mm)) found = 1;
            }
            if (found) { scalf = myscale[l]; break; }
        }
    }
    return scalf;
}

/* make_screening_lists (mod_electrostatics.F90:538-864), direct (no FMM) branch. */
static void make_screening_lists(osys *s)
{
    int n = s->nmm, npol = s->npol;
    rowbuf_t *rows;
    /* S-S (static-static, m-scale by bond distance; :620-656) */
    rows = (rowbuf_t *)calloc(n ? n : 1, sizeof(rowbuf_t));
    for (int i = 0; i < n; i++)
        for (int l = 0; l < 4; l++)
            for (int k = s->conn[l].ri[i]; k < s->conn[l].ri[i + 1]; k++) {
                int j = s->conn[l].ci[k];
                double sc = screening_rule(s, i, 'S', j, 'S', '-');
                if (fabs(sc - 1.0) > EPS_RP) rb_push(&rows[i], j, sc);
            }
    rows_to_csr(n, rows, &s->list_S_S, 1);
    /* S-P (polarization field) */
    rows = (rowbuf_t *)calloc(n, sizeof(rowbuf_t));
    for (int i = 0; i < n; i++)
        for (int l = 0; l < 4; l++)
            for (int k = s->conn[l].ri[i]; k < s->conn[l].ri[i + 1]; k++) {
                int jp = s->mm_polar[s->conn[l].ci[k]];
                if (jp < 0) continue;
                double sc = screening_rule(s, i, 'S', jp, 'P', 'P');
                if (fabs(sc - 1.0) > EPS_RP) rb_push(&rows[i], jp, sc);
            }
    rows_to_csr(n, rows, &s->list_S_P_P, 1);
    /* P-P (only topological neighbours are scanned, as in the reference) */
    rows = (rowbuf_t *)calloc(npol > 0 ? npol : 1, sizeof(rowbuf_t));
    for (int ip = 0; ip < npol; ip++) {
        int i = s->polar_mm[ip];
        for (int l = 0; l < 4; l++)
            for (int k = s->conn[l].ri[i]; k < s->conn[l].ri[i + 1]; k++) {
                int jp = s->mm_polar[s->conn[l].ci[k]];
                if (jp < 0) continue;
                double sc = screening_rule(s, ip, 'P', jp, 'P', '-');
                if (fabs(sc - 1.0) > EPS_RP) rb_push(&rows[ip], jp, sc);
            }
    }
    rows_to_csr(npol, rows, &s->list_P_P, 1);
    /* S-P (direct field, AMOEBA) over polarization-group neighbours */
    if (s->amoeba) {
        rows = (rowbuf_t *)calloc(n, sizeof(rowbuf_t));
        for (int i = 0; i < n; i++) {
            int pg = s->grp[i];
            for (int l = 0; l < 4; l++)
                for (int k = s->gconn[l].ri[pg]; k < s->gconn[l].ri[pg + 1]; k++) {
                    int g = s->gconn[l].ci[k];
                    for (int m = s->grp_members.ri[g]; m < s->grp_members.ri[g + 1]; m++) {
                        int jp = s->mm_polar[s->grp_members.ci[m]];
                        if (jp < 0) continue;
                        double sc = screening_rule(s, i, 'S', jp, 'P', 'D');
                        if (fabs(sc - 1.0) > EPS_RP) rb_push(&rows[i], jp, sc);
                    }
                }
        }
        rows_to_csr(n, rows, &s->list_S_P_D, 1);
        csr_transpose(&s->list_S_P_D, npol, &s->tS_P_D);
    }
    csr_transpose(&s->list_S_P_P, npol, &s->tS_P_P);
    csr_transpose(&s->list_P_P, npol, &s->tP_P);
}

/* ---------------------------------------------------------------- creation */
/*
 * Build the prepared system from raw inputs (what mmpol_init_from_mmp +
 * mmpol_prepare produce; mod_inputloader.F90:15-257, mod_mmpol.F90:139-237):
 *   cmm     [nmm][3] bohr
 *   q       [nmm][ld_cart] lab-frame multipoles, AMOEBA quadrupoles already /3
 *   pol_mm  [nmm] polarizability per MM site in bohr^3 (0 = not polarizable)
 *   td_mm   [nmm] Thole damping a per MM site (AMOEBA)
 *   adj_ri/adj_ci  1-2 adjacency, 0-based CSR
 *   grp     [nmm] polarization group, 0-based (AMOEBA; may be NULL otherwise)
 *   scales  [5][4] = mscale, pscale, dscale, uscale, pscale_intra
 */
void *oracle_create(int nmm, int amoeba, const double *cmm, const double *q,
                    const double *pol_mm, const double *td_mm, double thole_scale,
                    const int *adj_ri, const int *adj_ci, const int *grp,
                    const double *scales)
{
    osys *s = (osys *)calloc(1, sizeof(osys));
    s->nmm = nmm; s->amoeba = amoeba;
    s->ld_cart = amoeba ? 10 : 1;
    s->n_ipd = amoeba ? 2 : 1;
    s->cmm = (double *)malloc(sizeof(double) * 3 * nmm);
    memcpy(s->cmm, cmm, sizeof(double) * 3 * nmm);
    s->q = (double *)malloc(sizeof(double) * s->ld_cart * nmm);
    memcpy(s->q, q, sizeof(double) * s->ld_cart * nmm);
    memcpy(s->mscale, scales + 0, sizeof(double) * 4);
    memcpy(s->pscale, scales + 4, sizeof(double) * 4);
    memcpy(s->dscale, scales + 8, sizeof(double) * 4);
    memcpy(s->uscale, scales + 12, sizeof(double) * 4);
    memcpy(s->pscale_intra, scales + 16, sizeof(double) * 4);

    /* remove_null_pol (mod_electrostatics.F90:469-536): keep |pol| > eps_rp */
    s->mm_polar = (int *)malloc(sizeof(int) * nmm);
    s->polar_mm = (int *)malloc(sizeof(int) * (nmm ? nmm : 1));
    s->pol = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
    s->thole_damping = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
    s->npol = 0;
    for (int i = 0; i < nmm; i++) {
        s->mm_polar[i] = -1;
        if (fabs(pol_mm[i]) > EPS_RP) {
            s->polar_mm[s->npol] = i;
            s->pol[s->npol] = pol_mm[i];
            s->thole_damping[s->npol] = td_mm ? td_mm[i] : 0.39;
            s->mm_polar[i] = s->npol;
            s->npol++;
        }
    }
    /* thole_init (mod_electrostatics.F90:866-897) */
    s->thole = (double *)calloc(nmm ? nmm : 1, sizeof(double));
    for (int p = 0; p < s->npol; p++)
        s->thole[s->polar_mm[p]] = pow(s->pol[p], 1.0 / 6.0);
    if (!amoeba) {
        if (thole_scale > EPS_RP) {
            double f = sqrt(thole_scale);
            for (int i = 0; i < nmm; i++) s->thole[i] *= f;
        } else {
            oracle_fatal(s, "Scale factor for Thole damping should be greater than 0.0 when non-AMOEBA FF are used");
        }
    }
    /* connectivity 1-2..1-5 */
    csr_t adj = {nmm, adj_ri[nmm], (int *)adj_ri, (int *)adj_ci, NULL};
    conn_levels(&adj, 4, 0, s->conn);
    /* polarization groups (AMOEBA): reverse table, group adjacency, group levels 0..3 */
    s->grp = (int *)calloc(nmm ? nmm : 1, sizeof(int));
    if (amoeba && grp) {
        memcpy(s->grp, grp, sizeof(int) * nmm);
        int ng = 0;
        for (int i = 0; i < nmm; i++) if (grp[i] + 1 > ng) ng = grp[i] + 1;
        s->ngrp = ng;
        rowbuf_t *rows = (rowbuf_t *)calloc(ng ? ng : 1, sizeof(rowbuf_t));
        for (int i = 0; i < nmm; i++) rb_push(&rows[grp[i]], i, 1.0);
        rows_to_csr(ng, rows, &s->grp_members, 0);
        /* build_pg_adjacency_matrix (mod_mmpol.F90:272-323) */
        rows = (rowbuf_t *)calloc(ng ? ng : 1, sizeof(rowbuf_t));
        for (int g = 0; g < ng; g++)
            for (int m = s->grp_members.ri[g]; m < s->grp_members.ri[g + 1]; m++) {
                int a1 = s->grp_members.ci[m];
                for (int k = adj_ri[a1]; k < adj_ri[a1 + 1]; k++) {
                    int g2 = grp[adj_ci[k]];
                    if (g2 == g) continue;
                    int dup = 0;
                    for (int t = 0; t < rows[g].n; t++) if (rows[g].idx[t] == g2) dup = 1;
                    if (!dup) rb_push(&rows[g], g2, 1.0);
                }
            }
        csr_t gadj;
        rows_to_csr(ng, rows, &gadj, 0);
        conn_levels(&gadj, 3, 1, s->gconn);
        csr_free(&gadj);
    } else if (amoeba) {
        oracle_fatal(s, "AMOEBA system without polarization groups");
    }
    make_screening_lists(s);
    return s;
}

void oracle_destroy(void *h)
{
    osys *s = (osys *)h;
    if (!s) return;
    free(s->cmm); free(s->q); free(s->polar_mm); free(s->mm_polar); free(s->pol);
    free(s->thole); free(s->thole_damping); free(s->grp);
    for (int l = 0; l < 4; l++) { csr_free(&s->conn[l]); csr_free(&s->gconn[l]); }
    csr_free(&s->grp_members);
    csr_free(&s->list_S_P_P); csr_free(&s->list_S_P_D); csr_free(&s->list_P_P); csr_free(&s->list_S_S);
    csr_free(&s->tS_P_P); csr_free(&s->tS_P_D); csr_free(&s->tP_P);
    free(s);
}

/*
 * ommp_turn_pol_off (mod_interface.F90:134-168) followed by remove_null_pol
 * (mod_electrostatics.F90:469-536): the listed MM atoms (1-based) lose their
 * polarizability and the polarizable set is compacted in order; thole(:)
 * keeps its values (thole_init does not run again), so a switched-off atom's
 * pairs stay damped as sources, while its damping parameter a no longer enters
 * (it counts as not polarizable in the damped kernel's combination rule).  Two
 * points where the reference is not self-consistent are resolved: its loop
 * reads mm_polar at the loop counter (:156) rather than at the listed atom,
 * the same atom for the lists 1..n its callers pass; and it keeps the
 * screening lists and thole_damping of the old polarizable numbering, which
 * are rebuilt and compacted here.  Returns -1 after a fatal error.
 */
int oracle_turn_pol_off(void *h, int n, const int *nopol)
{
    osys *s = (osys *)h;
    if (n <= 0) return 0;
    int mn = nopol[0], mx = nopol[0];
    for (int k = 1; k < n; k++) {
        if (nopol[k] < mn) mn = nopol[k];
        if (nopol[k] > mx) mx = nopol[k];
    }
    if (mn < 1) {
        oracle_fatal(s, "Atom indexes are 1-based, so no index should be less than 1");
        return -1;
    }
    if (mx > s->nmm) {
        oracle_fatal(s, "Atom indexes should be inside the topology range");
        return -1;
    }
    for (int k = 0; k < n; k++) {
        const int p = s->mm_polar[nopol[k] - 1];
        if (p >= 0) s->pol[p] = 0.0;
    }
    int np = 0;
    for (int p = 0; p < s->npol; p++)
        if (fabs(s->pol[p]) > EPS_RP) {
            s->polar_mm[np] = s->polar_mm[p];
            s->pol[np] = s->pol[p];
            s->thole_damping[np] = s->thole_damping[p];
            np++;
        }
    if (np == s->npol) return 0;
    s->npol = np;
    for (int i = 0; i < s->nmm; i++) s->mm_polar[i] = -1;
    for (int p = 0; p < np; p++) s->mm_polar[s->polar_mm[p]] = p;
    csr_free(&s->list_S_P_P); csr_free(&s->list_S_P_D); csr_free(&s->list_P_P); csr_free(&s->list_S_S);
    csr_free(&s->tS_P_P); csr_free(&s->tS_P_D); csr_free(&s->tP_P);
    make_screening_lists(s);
    return 0;
}

int oracle_npol(void *h) { return ((osys *)h)->npol; }
int oracle_fatal_flag(void *h) { return ((osys *)h)->fatal; }
const char *oracle_fatal_msg(void *h) { return ((osys *)h)->msg; }
void oracle_get_polar_mm(void *h, int *out) { osys *s = (osys *)h; memcpy(out, s->polar_mm, sizeof(int) * s->npol); }
void oracle_get_thole(void *h, double *out) { osys *s = (osys *)h; memcpy(out, s->thole, sizeof(double) * s->nmm); }

/* Number of entries / copy of a screening list: which 0=S_P_P 1=S_P_D 2=P_P
   3+l=conn level l, 7+l=group level l, 11=S_S */
static const csr_t *list_by_id(osys *s, int which)
{
    return which == 0 ? &s->list_S_P_P : which == 1 ? &s->list_S_P_D : which == 2 ? &s->list_P_P :
           which < 7 ? &s->conn[which - 3] : which < 11 ? &s->gconn[which - 7] : &s->list_S_S;
}
int oracle_list_nnz(void *h, int which)
{
    const csr_t *c = list_by_id((osys *)h, which);
    return c->ri ? c->nnz : 0;
}
void oracle_list_get(void *h, int which, int *ri, int *ci, double *val)
{
    const csr_t *c = list_by_id((osys *)h, which);
    if (!c->ri) return;
    memcpy(ri, c->ri, sizeof(int) * (c->n + 1));
    memcpy(ci, c->ci, sizeof(int) * (c->nnz ? c->nnz : 0));
    if (val && c->val) memcpy(val, c->val, sizeof(double) * c->nnz);
}

double oracle_screening_rule(void *h, int i, int kind_i, int j, int kind_j, int field)
{
    return screening_rule((osys *)h, i, (char)kind_i, j, (char)kind_j, (char)field);
}

/* ------------------------------------------------------------- kernels */
/* coulomb_kernel (mod_electrostatics.F90:978-1016) */
static int coulomb_kernel(const double dr[3], int maxder, double *res)
{
    double r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    if (r < EPS_RP) return 1;
    res[0] = 1.0 / r;
    double inv2 = res[0] * res[0];
    for (int l = 1; l <= maxder; l++) res[l] = res[l - 1] * inv2;
    return 0;
}

/* damped_coulomb_kernel (mod_electrostatics.F90:1018-1095); i, j MM indices, dr = r_j - r_i */
static void damped_kernel(osys *s, int i, int j, int maxder, double *res, double dr[3])
{
    double sij = s->thole[i] * s->thole[j], a = 0.0;
    int pi = s->mm_polar[i], pj = s->mm_polar[j];
    if (pi >= 0 && pj >= 0) a = fmin(s->thole_damping[pi], s->thole_damping[pj]);
    else if (pi >= 0) a = s->thole_damping[pi];
    else if (pj >= 0) a = s->thole_damping[pj];
    else oracle_fatal(s, "Computing damped kernel between two mm atoms");
    for (int k = 0; k < 3; k++) dr[k] = s->cmm[3 * j + k] - s->cmm[3 * i + k];
    if (coulomb_kernel(dr, maxder, res)) {
        oracle_fatal(s, "Requesting Coulomb kernel for two atoms placed in the same point");
        for (int l = 0; l <= maxder; l++) res[l] = 0.0;
        return;
    }
    if (fabs(sij) < EPS_RP) return;
    double u = 1.0 / (res[0] * sij);
    double u3 = u * u * u;
    if (s->amoeba && u3 * a < 50.0) {
        double fexp = -a * u3, e = exp(fexp);
        if (maxder >= 1) res[1] *= 1.0 - e;
        if (maxder >= 2) res[2] *= 1.0 - (1.0 - fexp) * e;
        if (maxder >= 3) res[3] *= 1.0 - (1.0 - fexp + 0.6 * fexp * fexp) * e;
        if (maxder >= 4) res[4] *= 1.0 - (1.0 - fexp + 18.0 / 35.0 * fexp * fexp - 9.0 / 35.0 * fexp * fexp * fexp) * e;
    } else if (!s->amoeba && res[0] > 1.0 / sij) {
        double u4 = u3 * u;
        if (maxder >= 1) res[1] *= 4.0 * u3 - 3.0 * u4;
        if (maxder >= 2) res[2] *= u4;
        if (maxder >= 3) res[3] *= 0.2 * u4;
    }
}

/* Field parts of q_elec_prop / mu_elec_prop / quad_elec_prop (:1097-1405) */
static inline void q_field(double q, const double dr[3], const double *k, double E[3])
{
    for (int c = 0; c < 3; c++) E[c] += q * k[1] * dr[c];
}
static inline void mu_field(const double *mu, const double dr[3], const double *k, double E[3])
{
    double md = mu[0] * dr[0] + mu[1] * dr[1] + mu[2] * dr[2];
    for (int c = 0; c < 3; c++) E[c] += 3.0 * md * dr[c] * k[2] - mu[c] * k[1];
}
static inline void quad_field(const double *Q, const double dr[3], const double *k, double E[3])
{
    /* storage xx, xy, yy, xz, yz, zz (f_cart_components.h:13-18) */
    double qr[3];
    qr[0] = Q[0] * dr[0] + Q[1] * dr[1] + Q[3] * dr[2];
    qr[1] = Q[1] * dr[0] + Q[2] * dr[1] + Q[4] * dr[2];
    qr[2] = Q[3] * dr[0] + Q[4] * dr[1] + Q[5] * dr[2];
    double qrr = qr[0] * dr[0] + qr[1] * dr[1] + qr[2] * dr[2];
    for (int c = 0; c < 3; c++) E[c] += 15.0 * qrr * dr[c] * k[3] - 6.0 * qr[c] * k[2];
}

/* Potential parts (used only by the closed-form finite-difference checks). */
double oracle_multipole_potential_field(const double *qsrc, int ld_cart, const double *dr,
                                        const double *kernel, double *E)
{
    double V = kernel[0] * qsrc[0];
    E[0] = E[1] = E[2] = 0.0;
    q_field(qsrc[0], dr, kernel, E);
    if (ld_cart == 10) {
        const double *mu = qsrc + 1, *Q = qsrc + 4;
        double md = mu[0] * dr[0] + mu[1] * dr[1] + mu[2] * dr[2];
        V += md * kernel[1];
        mu_field(mu, dr, kernel, E);
        double qr[3];
        qr[0] = Q[0] * dr[0] + Q[1] * dr[1] + Q[3] * dr[2];
        qr[1] = Q[1] * dr[0] + Q[2] * dr[1] + Q[4] * dr[2];
        qr[2] = Q[3] * dr[0] + Q[4] * dr[1] + Q[5] * dr[2];
        V += 3.0 * (qr[0] * dr[0] + qr[1] * dr[1] + qr[2] * dr[2]) * kernel[2];
        quad_field(Q, dr, kernel, E);
    }
    return V;
}

void oracle_damped_kernel(void *h, int i_mm, int j_mm, int maxder, double *res, double *dr)
{
    damped_kernel((osys *)h, i_mm, j_mm, maxder, res, dr);
}

/* ---------------------------------------------------------------- RHS */
/* elec_prop_M2D direct branch (mod_electrostatics.F90:2579-2732): E[n_ipd][npol][3]
   (D = set 0, P = set 1 for AMOEBA). */
void oracle_m2d(void *h, double *E)
{
    osys *s = (osys *)h;
    int npol = s->npol, nmm = s->nmm, amoeba = s->amoeba;
    memset(E, 0, sizeof(double) * 3 * npol * s->n_ipd);
    #pragma omp parallel
    {
        double *sp = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
        double *sd = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
        char *hp = (char *)calloc(nmm ? nmm : 1, 1), *hd = (char *)calloc(nmm ? nmm : 1, 1);
        #pragma omp for schedule(dynamic, 4)
        for (int j = 0; j < npol; j++) {
            for (int k = s->tS_P_P.ri[j]; k < s->tS_P_P.ri[j + 1]; k++) {
                sp[s->tS_P_P.ci[k]] = s->tS_P_P.val[k]; hp[s->tS_P_P.ci[k]] = 1;
            }
            if (amoeba)
                for (int k = s->tS_P_D.ri[j]; k < s->tS_P_D.ri[j + 1]; k++) {
                    sd[s->tS_P_D.ci[k]] = s->tS_P_D.val[k]; hd[s->tS_P_D.ci[k]] = 1;
                }
            int jm = s->polar_mm[j];
            double EP[3] = {0, 0, 0}, ED[3] = {0, 0, 0};
            for (int i = 0; i < nmm; i++) {
                if (i == jm) continue;
                int to_do_p = 1, to_scale_p = hp[i], to_do_d = 1, to_scale_d = 0;
                double scp = 1.0, scd = 1.0;
                if (to_scale_p) { scp = sp[i]; to_do_p = fabs(scp) > EPS_RP; }
                if (amoeba) {
                    to_scale_d = hd[i];
                    if (to_scale_d) { scd = sd[i]; to_do_d = fabs(scd) > EPS_RP; }
                } else {
                    to_do_d = 0;
                }
                if (!(to_do_p || to_do_d)) continue;
                double k[5], dr[3], tE[3] = {0, 0, 0};
                damped_kernel(s, i, jm, amoeba ? 3 : 1, k, dr);
                q_field(s->q[s->ld_cart * i], dr, k, tE);
                if (amoeba) {
                    mu_field(s->q + 10 * i + 1, dr, k, tE);
                    quad_field(s->q + 10 * i + 4, dr, k, tE);
                }
                if (to_do_p)
                    for (int c = 0; c < 3; c++) EP[c] += to_scale_p ? tE[c] * scp : tE[c];
                if (to_do_d)
                    for (int c = 0; c < 3; c++) ED[c] += to_scale_d ? tE[c] * scd : tE[c];
            }
            if (amoeba) {
                for (int c = 0; c < 3; c++) { E[3 * j + c] = ED[c]; E[3 * npol + 3 * j + c] = EP[c]; }
            } else {
                for (int c = 0; c < 3; c++) E[3 * j + c] = EP[c];
            }
            for (int k = s->tS_P_P.ri[j]; k < s->tS_P_P.ri[j + 1]; k++) hp[s->tS_P_P.ci[k]] = 0;
            if (amoeba)
                for (int k = s->tS_P_D.ri[j]; k < s->tS_P_D.ri[j + 1]; k++) hd[s->tS_P_D.ci[k]] = 0;
        }
        free(sp); free(sd); free(hp); free(hd);
    }
}

/* ------------------------------------------------------------- T matvec */
/* field_extD2D direct branch (mod_electrostatics.F90:2068-2109) + TMatVec_otf /
   TMatVec_diag (mod_polarization.F90:459-504): y = -E_dip(x) [+ x/alpha]. */
void oracle_tmatvec_otf(void *h, const double *x, double *y, int dodiag)
{
    osys *s = (osys *)h;
    int npol = s->npol;
    #pragma omp parallel
    {
        double *sc = (double *)malloc(sizeof(double) * (npol ? npol : 1));
        char *hs = (char *)calloc(npol ? npol : 1, 1);
        #pragma omp for schedule(dynamic, 4)
        for (int j = 0; j < npol; j++) {
            for (int k = s->tP_P.ri[j]; k < s->tP_P.ri[j + 1]; k++) {
                sc[s->tP_P.ci[k]] = s->tP_P.val[k]; hs[s->tP_P.ci[k]] = 1;
            }
            double E[3] = {0, 0, 0};
            for (int i = 0; i < npol; i++) {
                if (i == j) continue;
                int to_scale = hs[i], to_do = 1;
                double f = 1.0;
                if (to_scale) { f = sc[i]; to_do = fabs(f) > EPS_RP; }
                if (!to_do) continue;
                double k[3], dr[3], tE[3] = {0, 0, 0};
                damped_kernel(s, s->polar_mm[i], s->polar_mm[j], 2, k, dr);
                mu_field(x + 3 * i, dr, k, tE);
                for (int c = 0; c < 3; c++) E[c] += to_scale ? tE[c] * f : tE[c];
            }
            for (int c = 0; c < 3; c++) {
                y[3 * j + c] = -E[c];
                if (dodiag) y[3 * j + c] += x[3 * j + c] / s->pol[j];
            }
            for (int k = s->tP_P.ri[j]; k < s->tP_P.ri[j + 1]; k++) hs[s->tP_P.ci[k]] = 0;
        }
        free(sc); free(hs);
    }
}

/* dipole_T + create_tmat (mod_polarization.F90:327-435): dense n x n, row-major
   here (the matrix is symmetric, so layout does not matter). */
void oracle_create_tmat(void *h, double *T)
{
    osys *s = (osys *)h;
    int npol = s->npol;
    long n = 3L * npol;
    memset(T, 0, sizeof(double) * n * n);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npol; i++) {
        for (int j = 0; j <= i; j++) {
            double t[3][3] = {{0}};
            if (i == j) {
                for (int c = 0; c < 3; c++) t[c][c] = 1.0 / s->pol[i];
            } else {
                double sc = screening_rule(s, i, 'P', j, 'P', '-');
                if (fabs(sc) > EPS_RP) {
                    double k[3], dr[3];
                    damped_kernel(s, s->polar_mm[i], s->polar_mm[j], 2, k, dr);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            t[a][b] = (a == b ? k[1] : 0.0) - 3.0 * k[2] * dr[a] * dr[b];
                    if (fabs(sc - 1.0) > EPS_RP)
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++) t[a][b] *= sc;
                }
            }
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    T[(3L * j + a) * n + 3L * i + b] = t[a][b];
                    T[(3L * i + b) * n + 3L * j + a] = t[a][b];
                }
        }
    }
}

/* ------------------------------------------------------------- solvers */
typedef struct {
    osys *s;
    int matv;
    const double *T; /* dense, for INCORE */
} mv_ctx;

static void matvec(mv_ctx *c, const double *x, double *y, int dodiag)
{
    osys *s = c->s;
    long n = 3L * s->npol;
    if (c->matv == MATV_DIRECT) {
        oracle_tmatvec_otf(s, x, y, dodiag);
    } else {
        /* TMatVec_incore: dgemm then remove the diagonal (mod_polarization.F90:437-533) */
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            double acc = 0.0;
            const double *row = c->T + i * n;
            for (long k = 0; k < n; k++) acc += row[k] * x[k];
            y[i] = acc - row[i] * x[i];
            if (dodiag) y[i] += x[i] / s->pol[i / 3];
        }
    }
}

static double dot(long n, const double *a, const double *b)
{
    double acc = 0.0;
    for (long i = 0; i < n; i++) acc += a[i] * b[i];
    return acc;
}

/* Threshold of the |p.Tp| quick return of the CG loop.  The reference uses
   eps_rp; golden-vector generation lowers it to converge far below the
   reference tolerance (oracle_set_small_step_eps). */
static double g_small_step_eps = EPS_RP;
void oracle_set_small_step_eps(double e) { g_small_step_eps = e; }

/* conjugate_gradient_solver (mod_solvers.F90:86-216).  Returns 0 on success,
   1 if not converged (the reference calls fatal_error). */
static int cg_solve(mv_ctx *c, long n, const double *rhs, double *x, double tol, int maxit, int *iters)
{
    osys *s = c->s;
    double *r = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *hh = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double rms = dot(n, x, x), alpha, gnew, gold, gama;
    if (rms < EPS_RP)
        for (long i = 0; i < n; i++) x[i] = s->pol[i / 3] * x[i];
    matvec(c, x, z, 1);
    for (long i = 0; i < n; i++) r[i] = rhs[i] - z[i];
    for (long i = 0; i < n; i++) z[i] = s->pol[i / 3] * r[i];
    memcpy(p, z, sizeof(double) * n);
    gold = dot(n, r, z);
    gama = 0.0;
    int it;
    for (it = 1; it <= maxit; it++) {
        matvec(c, p, hh, 1);
        gama = dot(n, hh, p);
        if (fabs(gama) < g_small_step_eps || gama == 0.0) break;
        alpha = gold / gama;
        for (long i = 0; i < n; i++) x[i] += alpha * p[i];
        for (long i = 0; i < n; i++) r[i] -= alpha * hh[i];
        for (long i = 0; i < n; i++) z[i] = s->pol[i / 3] * r[i];
        gnew = dot(n, r, z);
        rms = sqrt(gnew / (double)n);
        if (rms < tol) break;
        gama = gnew / gold;
        for (long i = 0; i < n; i++) p[i] = gama * p[i] + z[i];
        gold = gnew;
    }
    if (iters) *iters = it > maxit ? maxit : it;
    free(r); free(p); free(hh); free(z);
    if (rms > tol && fabs(gama) > g_small_step_eps) return 1;
    return 0;
}

/* LU with partial pivoting, in place, row-major; returns info (0 ok, k>0 zero pivot at k). */
static int lu_factor(long n, double *A, int *piv)
{
    int info = 0;
    for (long k = 0; k < n; k++) {
        long p = k; double mx = fabs(A[k * n + k]);
        for (long i = k + 1; i < n; i++) if (fabs(A[i * n + k]) > mx) { mx = fabs(A[i * n + k]); p = i; }
        piv[k] = (int)p;
        if (A[p * n + k] == 0.0) { if (!info) info = (int)k + 1; continue; }
        if (p != k)
            for (long j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
        double inv = 1.0 / A[k * n + k];
        #pragma omp parallel for schedule(static) if (n > 256)
        for (long i = k + 1; i < n; i++) {
            double l = A[i * n + k] * inv;
            A[i * n + k] = l;
            for (long j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
        }
    }
    return info;
}

static void lu_solve(long n, const double *A, const int *piv, double *b)
{
    for (long k = 0; k < n; k++) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (long i = 0; i < n; i++) { double acc = b[i]; for (long j = 0; j < i; j++) acc -= A[i * n + j] * b[j]; b[i] = acc; }
    for (long i = n - 1; i >= 0; i--) {
        double acc = b[i];
        for (long j = i + 1; j < n; j++) acc -= A[i * n + j] * b[j];
        b[i] = acc / A[i * n + i];
    }
}

/* diis + makeb (mod_solvers.F90:367-469).  x, e are [ndiis][n]; b is
   (ndiis+1)^2 row-major with 0-based indices (b[0][0] is the Lagrangian corner). */
static void diis_step(long n, int *nmat, int ndiis, double *x, double *e, double *b, double *xnew)
{
    int B = ndiis + 1;
    if (*nmat >= ndiis) {
        for (int j = 2; j <= *nmat - 10; j++)
            for (int k = 2; k <= *nmat - 10; k++)
                b[(j - 1) * B + (k - 1)] = b[(j + 10 - 1) * B + (k + 10 - 1)];
        for (int j = 1; j <= *nmat - 10; j++) {
            memcpy(x + (long)(j - 1) * n, x + (long)(j + 10 - 1) * n, sizeof(double) * n);
            memcpy(e + (long)(j - 1) * n, e + (long)(j + 10 - 1) * n, sizeof(double) * n);
        }
        *nmat -= 10;
    }
    int m = *nmat, nm1 = m + 1;
    /* makeb */
    if (m == 1) {
        b[0] = 0.0; b[1] = 1.0; b[B] = 1.0;
        b[B + 1] = dot(n, e, e);
    } else {
        b[m * B + 0] = 1.0; b[0 * B + m] = 1.0;
        for (int i = 1; i <= m - 1; i++) {
            double bij = dot(n, e + (long)(i - 1) * n, e + (long)(m - 1) * n);
            b[m * B + i] = bij; b[i * B + m] = bij;
        }
        b[m * B + m] = dot(n, e + (long)(m - 1) * n, e + (long)(m - 1) * n);
    }
    double *bloc = (double *)malloc(sizeof(double) * nm1 * nm1);
    double *cex = (double *)calloc(nm1, sizeof(double));
    int *piv = (int *)malloc(sizeof(int) * nm1);
    for (int i = 0; i < nm1; i++) for (int j = 0; j < nm1; j++) bloc[i * nm1 + j] = b[i * B + j];
    cex[0] = 1.0;
    int info = lu_factor(nm1, bloc, piv);
    if (info != 0) {
        *nmat = 1;
    } else {
        lu_solve(nm1, bloc, piv, cex);
        memset(xnew, 0, sizeof(double) * n);
        for (int i = 1; i <= m; i++) {
            const double *xi = x + (long)(i - 1) * n;
            for (long k = 0; k < n; k++) xnew[k] += cex[i] * xi[k];
        }
        *nmat = m + 1;
    }
    free(bloc); free(cex); free(piv);
}

/* jacobi_diis_solver (mod_solvers.F90:218-365) */
static int diis_solve(mv_ctx *c, long n, const double *rhs, double *x, double tol, int maxit, int diis_max, int *iters)
{
    osys *s = c->s;
    double *xnew = (double *)malloc(sizeof(double) * n), *y = (double *)malloc(sizeof(double) * n);
    double *xd = NULL, *ed = NULL, *b = NULL;
    int do_diis = diis_max > 0, nmat = 1;
    if (do_diis) {
        xd = (double *)calloc((size_t)n * diis_max, sizeof(double));
        ed = (double *)calloc((size_t)n * diis_max, sizeof(double));
        b = (double *)calloc((size_t)(diis_max + 1) * (diis_max + 1), sizeof(double));
    }
    if (dot(n, x, x) < EPS_RP)
        for (long i = 0; i < n; i++) x[i] = s->pol[i / 3] * rhs[i];
    double max_norm = 0.0;
    int it;
    for (it = 1; it <= maxit; it++) {
        matvec(c, x, y, 0);
        for (long i = 0; i < n; i++) y[i] = rhs[i] - y[i];
        for (long i = 0; i < n; i++) xnew[i] = s->pol[i / 3] * y[i];
        if (do_diis) {
            memcpy(xd + (long)(nmat - 1) * n, xnew, sizeof(double) * n);
            for (long i = 0; i < n; i++) ed[(long)(nmat - 1) * n + i] = xnew[i] - x[i];
            diis_step(n, &nmat, diis_max, xd, ed, b, xnew);
        }
        max_norm = 0.0;
        for (long i = 0; i < n; i++) { double d = fabs(xnew[i] - x[i]); if (d > max_norm) max_norm = d; }
        memcpy(x, xnew, sizeof(double) * n);
        if (max_norm < tol) break;
    }
    if (iters) *iters = it > maxit ? maxit : it;
    free(xnew); free(y); free(xd); free(ed); free(b);
    return max_norm > tol ? 1 : 0;
}

/*
 * polarization() driver (mod_polarization.F90:49-313).
 *   e    [n_ipd][npol][3] field (D then P)
 *   ipd  [n_ipd][npol][3] in: previous dipoles (used if use_guess); out: dipoles
 *   mask bit k set = solve set k (AMOEBA); ignored for n_ipd == 1
 *   iters[n_ipd] out: iterations per set
 * Returns 0, or 1 when a solver did not converge (reference: fatal_error).
 */
int oracle_polarization(void *h, const double *e, double *ipd, int solver, int matv,
                        int use_guess, int mask, double tol, int maxit, int *iters)
{
    osys *s = (osys *)h;
    if (s->npol == 0) return 0;
    long n = 3L * s->npol;
    int nipd = s->n_ipd, rc = 0;
    int m[2] = {1, 1};
    if (nipd > 1) { m[0] = mask & 1; m[1] = (mask >> 1) & 1; }
    if (solver == SOLVER_NONE) solver = SOLVER_CG;
    if (matv == MATV_NONE) matv = MATV_DIRECT;
    double *x = (double *)calloc((size_t)n * nipd, sizeof(double));
    double *T = NULL;
    if (matv == MATV_INCORE || solver == SOLVER_INVERSION) {
        T = (double *)malloc(sizeof(double) * n * n);
        oracle_create_tmat(s, T);
    }
    mv_ctx c = {s, matv, T};
    if (solver != SOLVER_INVERSION && use_guess)
        for (int k = 0; k < nipd; k++) if (m[k]) memcpy(x + k * n, ipd + k * n, sizeof(double) * n);
    for (int k = 0; k < nipd; k++) if (iters) iters[k] = 0;
    if (solver == SOLVER_CG || solver == SOLVER_DIIS) {
        for (int k = 0; k < nipd; k++) {
            if (!m[k]) continue;
            if (k == 1 && m[0] && !use_guess) memcpy(x + n, x, sizeof(double) * n);
            int it = 0, bad;
            if (solver == SOLVER_CG) bad = cg_solve(&c, n, e + k * n, x + k * n, tol, maxit, &it);
            else bad = diis_solve(&c, n, e + k * n, x + k * n, tol, maxit, 20, &it);
            if (iters) iters[k] = it;
            if (bad) rc = 1;
        }
    } else if (solver == SOLVER_INVERSION) {
        int *piv = (int *)malloc(sizeof(int) * n);
        double *A = (double *)malloc(sizeof(double) * n * n);
        memcpy(A, T, sizeof(double) * n * n);
        if (lu_factor(n, A, piv)) rc = 1;
        for (int k = 0; k < nipd; k++) {
            if (!m[k]) continue;
            memcpy(x + k * n, e + k * n, sizeof(double) * n);
            lu_solve(n, A, piv, x + k * n);
        }
        free(piv); free(A);
    } else {
        rc = 2;
    }
    memcpy(ipd, x, sizeof(double) * n * nipd);
    free(x); free(T);
    return rc;
}

/* energy_MM_pol (mod_electrostatics.F90:943-976): E = -1/2 sum mu_D . E_M2D_P */
double oracle_energy_pol(void *h, const double *ipd, const double *e_m2d)
{
    osys *s = (osys *)h;
    long n = 3L * s->npol;
    double acc = 0.0;
    if (s->amoeba) {
        for (int c = 0; c < 3; c++)
            for (int j = 0; j < s->npol; j++) acc -= ipd[3 * j + c] * e_m2d[n + 3 * j + c];
    } else {
        for (int c = 0; c < 3; c++)
            for (int j = 0; j < s->npol; j++) acc -= ipd[3 * j + c] * e_m2d[3 * j + c];
    }
    return 0.5 * acc;
}

/* Field of point charges at the polarizable sites, undamped (E_n2p,
   mod_qm_helper.F90:376-391): used for the QM-region external fields. */
void oracle_charges_field(void *h, int nq, const double *cq, const double *qq, double *E)
{
    osys *s = (osys *)h;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < s->npol; j++) {
        const double *r = s->cmm + 3 * s->polar_mm[j];
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < nq; a++) {
            double dr[3] = {r[0] - cq[3 * a], r[1] - cq[3 * a + 1], r[2] - cq[3 * a + 2]};
            double k[2];
            if (coulomb_kernel(dr, 1, k)) continue;
            for (int c = 0; c < 3; c++) acc[c] += qq[a] * k[1] * dr[c];
        }
        for (int c = 0; c < 3; c++) E[3 * j + c] = acc[c];
    }
}

/* potential_M2E / potential_D2E / field_M2E / field_D2E
   (mod_electrostatics.F90:3067-3336): potential V[n] and field E[n][3] at n
   external points cext[n][3], undamped and unscreened, ADDED to V and E
   (either may be NULL).  which: 1 = static multipoles (M2E), 2 = induced
   dipoles (D2E), 3 = both.  D2E uses the D set for the potential (:3094-3110)
   and 1/2 (mu_P + mu_D) for the field (:3250) with AMOEBA; ipd is
   [n_ipd][npol][3].  A point on top of a source is fatal (coulomb_kernel,
   :1003-1008). */
void oracle_x2e(void *h, const double *ipd, int n, const double *cext, int which, double *V, double *E)
{
    osys *s = (osys *)h;
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 16) reduction(|:bad)
    for (int j = 0; j < n; j++) {
        const double *p = cext + 3 * j;
        double v = 0.0, e[3] = {0.0, 0.0, 0.0};
        if (which & 1) {
            for (int i = 0; i < s->nmm; i++) {
                const double *c = s->cmm + 3 * i;
                double dr[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
                double k[4];
                if (coulomb_kernel(dr, s->amoeba ? 3 : 1, k)) { bad = 1; continue; }
                double et[3];
                v += oracle_multipole_potential_field(s->q + (long)s->ld_cart * i, s->ld_cart, dr, k, et);
                for (int c3 = 0; c3 < 3; c3++) e[c3] += et[c3];
            }
        }
        if ((which & 2) && s->npol > 0) {
            for (int i = 0; i < s->npol; i++) {
                const double *c = s->cmm + 3 * s->polar_mm[i];
                double dr[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
                double k[3];
                if (coulomb_kernel(dr, 2, k)) { bad = 1; continue; }
                const double *mD = ipd + 3L * i;
                double mu[3];
                for (int c3 = 0; c3 < 3; c3++)
                    mu[c3] = s->amoeba ? 0.5 * (mD[c3] + ipd[3L * s->npol + 3L * i + c3]) : mD[c3];
                v += (mD[0] * dr[0] + mD[1] * dr[1] + mD[2] * dr[2]) * k[1];
                mu_field(mu, dr, k, e);
            }
        }
        if (V) V[j] += v;
        if (E)
            for (int c3 = 0; c3 < 3; c3++) E[3 * j + c3] += e[c3];
    }
    if (bad)
        oracle_fatal(s, "Requesting Coulomb kernel for two atoms placed in the same point");
}

/* ------------------------------------------------- geometric gradients */
/* Field gradient G[3][3] (G_ab = -dE_b/dr_a at the target, the quantity the
   reference stores as Egrd) and field Hessian H[3][3][3] (H_abc = d2E_c /
   dr_a dr_b, stored as EHes) of a charge, a dipole and a quadrupole (stored
   /3), ADDED to G and H (either may be NULL): grdE and HE branches of
   q_elec_prop / mu_elec_prop / quad_elec_prop (mod_electrostatics.F90:1128-
   1174, 1207-1283, 1320-1345), written as full tensors.  k[1..4] = damped
   lambda_3/r^3 .. lambda_9/r^9, d = r_target - r_source. */
static void q_grad(double q, const double d[3], const double *k, double G[3][3])
{
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) G[a][b] += q * (3.0 * k[2] * d[a] * d[b] - (a == b ? k[1] : 0.0));
}
static void mu_grad_hes(const double *mu, const double d[3], const double *k, double G[3][3], double H[3][3][3])
{
    double md = mu[0] * d[0] + mu[1] * d[1] + mu[2] * d[2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double dab = a == b ? 1.0 : 0.0;
            if (G) G[a][b] += 15.0 * k[3] * md * d[a] * d[b] - 3.0 * k[2] * (mu[a] * d[b] + mu[b] * d[a] + md * dab);
            if (H)
                for (int c = 0; c < 3; c++) {
                    double dac = a == c ? 1.0 : 0.0, dbc = b == c ? 1.0 : 0.0;
                    H[a][b][c] += 105.0 * k[4] * md * d[a] * d[b] * d[c]
                                  - 15.0 * k[3] * (mu[a] * d[b] * d[c] + mu[b] * d[a] * d[c] + mu[c] * d[a] * d[b]
                                                   + md * (dab * d[c] + dac * d[b] + dbc * d[a]))
                                  + 3.0 * k[2] * (mu[a] * dbc + mu[b] * dac + mu[c] * dab);
                }
        }
}
static void quad_grad(const double *Qs, const double d[3], const double *k, double G[3][3])
{
    const double Q[3][3] = {{Qs[0], Qs[1], Qs[3]}, {Qs[1], Qs[2], Qs[4]}, {Qs[3], Qs[4], Qs[5]}};
    double qd[3];
    for (int a = 0; a < 3; a++) qd[a] = Q[a][0] * d[0] + Q[a][1] * d[1] + Q[a][2] * d[2];
    double qrr = qd[0] * d[0] + qd[1] * d[1] + qd[2] * d[2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            G[a][b] += 105.0 * k[4] * qrr * d[a] * d[b] - 30.0 * k[3] * (qd[a] * d[b] + qd[b] * d[a])
                       - (a == b ? 15.0 * k[3] * qrr : 0.0) + 6.0 * k[2] * Q[a][b];
}
static void add_scaled_gh(double f, double G[3][3], double H[3][3][3], double Gt[3][3], double Ht[3][3][3])
{
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            Gt[a][b] += f * G[a][b];
            if (H)
                for (int c = 0; c < 3; c++) Ht[a][b][c] += f * H[a][b][c];
        }
}

/* polelec_geomgrad (mod_geomgrad.F90:113-273) without its last call
   (rotation_geomgrad, see oracle_rotation_geomgrad), on the fields that
   prepare_polelec(do_gg) builds (mod_electrostatics.F90:1563-1590):
     Egrd_M2D  field gradient of the static multipoles at the polarizable
               sites, D and P screening as E_M2D (elec_prop_M2D, :2313-2732);
     Egrd_D2D  field gradient of mu_P and of mu_D at the polarizable sites,
               u-screening (elec_prop_D2D, :2113-2311);
     E/Egrd/EHes_D2M  field, gradient and Hessian at every MM site of
               1/2 (mu_P with D screening + mu_D with P screening) for AMOEBA,
               of mu with P screening otherwise (elec_prop_D2M, :2736-3065).
   grad[nmm][3] is OVERWRITTEN with
     - q E_D2M + mu . Egrd_D2M - Q : EHes_D2M       at every MM site,
     + 1/2 mu_D . (Egrd_M2D + Egrd_D2D)(P) + 1/2 mu_P . (...)(D)  (AMOEBA) or
       mu . (Egrd_M2D + Egrd_D2D) (AMBER)            at every polarizable site.
   Emm[nmm][3] and Egrd[nmm][6] (xx, xy, yy, xz, yz, zz) receive E_D2M and
   Egrd_D2M, the inputs of the rotation term.  ipd is [n_ipd][npol][3]. */
void oracle_geomgrad(void *h, const double *ipd, double *grad, double *Emm, double *Egrd)
{
    osys *s = (osys *)h;
    const int npol = s->npol, nmm = s->nmm, amoeba = s->amoeba;
    const long n = 3L * npol;
    const double *muD = ipd, *muP = amoeba ? ipd + n : ipd;
    memset(grad, 0, sizeof(double) * 3 * nmm);
    memset(Emm, 0, sizeof(double) * 3 * nmm);
    memset(Egrd, 0, sizeof(double) * 6 * nmm);
    /* polarizable targets */
    #pragma omp parallel
    {
        double *sp = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
        double *sd = (double *)malloc(sizeof(double) * (nmm ? nmm : 1));
        double *su = (double *)malloc(sizeof(double) * (npol ? npol : 1));
        char *hp = (char *)calloc(nmm ? nmm : 1, 1), *hd = (char *)calloc(nmm ? nmm : 1, 1);
        char *hu = (char *)calloc(npol ? npol : 1, 1);
        #pragma omp for schedule(dynamic, 4)
        for (int j = 0; j < npol; j++) {
            for (int k = s->tS_P_P.ri[j]; k < s->tS_P_P.ri[j + 1]; k++) {
                sp[s->tS_P_P.ci[k]] = s->tS_P_P.val[k]; hp[s->tS_P_P.ci[k]] = 1;
            }
            if (amoeba)
                for (int k = s->tS_P_D.ri[j]; k < s->tS_P_D.ri[j + 1]; k++) {
                    sd[s->tS_P_D.ci[k]] = s->tS_P_D.val[k]; hd[s->tS_P_D.ci[k]] = 1;
                }
            for (int k = s->tP_P.ri[j]; k < s->tP_P.ri[j + 1]; k++) {
                su[s->tP_P.ci[k]] = s->tP_P.val[k]; hu[s->tP_P.ci[k]] = 1;
            }
            const int jm = s->polar_mm[j];
            double GP[3][3] = {{0}}, GD[3][3] = {{0}};
            /* Egrd_M2D */
            for (int i = 0; i < nmm; i++) {
                if (i == jm) continue;
                double scp = hp[i] ? sp[i] : 1.0, scd = 1.0;
                int do_p = !hp[i] || fabs(scp) > EPS_RP, do_d = 0;
                if (amoeba) {
                    scd = hd[i] ? sd[i] : 1.0;
                    do_d = !hd[i] || fabs(scd) > EPS_RP;
                }
                if (!(do_p || do_d)) continue;
                double k[5], dr[3], G[3][3] = {{0}};
                damped_kernel(s, i, jm, amoeba ? 4 : 2, k, dr);
                q_grad(s->q[s->ld_cart * i], dr, k, G);
                if (amoeba) {
                    mu_grad_hes(s->q + 10 * i + 1, dr, k, G, NULL);
                    quad_grad(s->q + 10 * i + 4, dr, k, G);
                }
                if (do_p) add_scaled_gh(scp, G, NULL, GP, NULL);
                if (do_d) add_scaled_gh(scd, G, NULL, GD, NULL);
            }
            /* Egrd_D2D: field of mu_P adds to the P set, of mu_D to the D set */
            for (int i = 0; i < npol; i++) {
                if (i == j) continue;
                double scu = hu[i] ? su[i] : 1.0;
                if (hu[i] && !(fabs(scu) > EPS_RP)) continue;
                double k[5], dr[3], G[3][3] = {{0}};
                damped_kernel(s, s->polar_mm[i], jm, 3, k, dr);
                mu_grad_hes(muP + 3 * i, dr, k, G, NULL);
                add_scaled_gh(scu, G, NULL, GP, NULL);
                if (amoeba) {
                    double G2[3][3] = {{0}};
                    mu_grad_hes(muD + 3 * i, dr, k, G2, NULL);
                    add_scaled_gh(scu, G2, NULL, GD, NULL);
                }
            }
            double g[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    if (amoeba)
                        g[a] += 0.5 * muD[3 * j + b] * GP[a][b] + 0.5 * muP[3 * j + b] * GD[a][b];
                    else
                        g[a] += muP[3 * j + b] * GP[a][b];
                }
            /* each polarizable site belongs to one MM atom: no write conflict */
            for (int a = 0; a < 3; a++) grad[3 * jm + a] += g[a];
            for (int k = s->tS_P_P.ri[j]; k < s->tS_P_P.ri[j + 1]; k++) hp[s->tS_P_P.ci[k]] = 0;
            if (amoeba)
                for (int k = s->tS_P_D.ri[j]; k < s->tS_P_D.ri[j + 1]; k++) hd[s->tS_P_D.ci[k]] = 0;
            for (int k = s->tP_P.ri[j]; k < s->tP_P.ri[j + 1]; k++) hu[s->tP_P.ci[k]] = 0;
        }
        free(sp); free(sd); free(su); free(hp); free(hd); free(hu);
    }
    /* MM targets: fields of the induced dipoles */
    #pragma omp parallel
    {
        double *sp = (double *)malloc(sizeof(double) * (npol ? npol : 1));
        double *sd = (double *)malloc(sizeof(double) * (npol ? npol : 1));
        char *hp = (char *)calloc(npol ? npol : 1, 1), *hd = (char *)calloc(npol ? npol : 1, 1);
        #pragma omp for schedule(dynamic, 4)
        for (int m = 0; m < nmm; m++) {
            for (int k = s->list_S_P_P.ri[m]; k < s->list_S_P_P.ri[m + 1]; k++) {
                sp[s->list_S_P_P.ci[k]] = s->list_S_P_P.val[k]; hp[s->list_S_P_P.ci[k]] = 1;
            }
            if (amoeba)
                for (int k = s->list_S_P_D.ri[m]; k < s->list_S_P_D.ri[m + 1]; k++) {
                    sd[s->list_S_P_D.ci[k]] = s->list_S_P_D.val[k]; hd[s->list_S_P_D.ci[k]] = 1;
                }
            double E[3] = {0, 0, 0}, G[3][3] = {{0}}, H[3][3][3] = {{{0}}};
            for (int i = 0; i < npol; i++) {
                if (s->polar_mm[i] == m) continue;
                double k[5], dr[3];
                damped_kernel(s, s->polar_mm[i], m, amoeba ? 4 : 2, k, dr);
                if (amoeba) {
                    /* 'P' kind: mu_P screened as the D field; 'D' kind: mu_D screened as P */
                    double scd = hd[i] ? sd[i] : 1.0, scp = hp[i] ? sp[i] : 1.0;
                    if (!hd[i] || fabs(scd) > EPS_RP) {
                        double Et[3] = {0, 0, 0}, Gt[3][3] = {{0}}, Ht[3][3][3] = {{{0}}};
                        mu_field(muP + 3 * i, dr, k, Et);
                        mu_grad_hes(muP + 3 * i, dr, k, Gt, Ht);
                        for (int c = 0; c < 3; c++) E[c] += 0.5 * scd * Et[c];
                        add_scaled_gh(0.5 * scd, Gt, Ht, G, H);
                    }
                    if (!hp[i] || fabs(scp) > EPS_RP) {
                        double Et[3] = {0, 0, 0}, Gt[3][3] = {{0}}, Ht[3][3][3] = {{{0}}};
                        mu_field(muD + 3 * i, dr, k, Et);
                        mu_grad_hes(muD + 3 * i, dr, k, Gt, Ht);
                        for (int c = 0; c < 3; c++) E[c] += 0.5 * scp * Et[c];
                        add_scaled_gh(0.5 * scp, Gt, Ht, G, H);
                    }
                } else {
                    double scp = hp[i] ? sp[i] : 1.0;
                    if (hp[i] && !(fabs(scp) > EPS_RP)) continue;
                    double Et[3] = {0, 0, 0};
                    mu_field(muP + 3 * i, dr, k, Et);
                    for (int c = 0; c < 3; c++) E[c] += scp * Et[c];
                }
            }
            double g[3];
            const double *qm = s->q + (long)s->ld_cart * m;
            for (int a = 0; a < 3; a++) g[a] = -qm[0] * E[a];
            if (amoeba) {
                const double *Qs = qm + 4;
                const double Q[3][3] = {{Qs[0], Qs[1], Qs[3]}, {Qs[1], Qs[2], Qs[4]}, {Qs[3], Qs[4], Qs[5]}};
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        g[a] += qm[1 + b] * G[a][b];
                        for (int c = 0; c < 3; c++) g[a] -= Q[b][c] * H[b][c][a];
                    }
                Egrd[6 * m + 0] = G[0][0]; Egrd[6 * m + 1] = G[0][1]; Egrd[6 * m + 2] = G[1][1];
                Egrd[6 * m + 3] = G[0][2]; Egrd[6 * m + 4] = G[1][2]; Egrd[6 * m + 5] = G[2][2];
            }
            for (int a = 0; a < 3; a++) Emm[3 * m + a] = E[a];
            #pragma omp critical
            for (int a = 0; a < 3; a++) grad[3 * m + a] += g[a];
            for (int k = s->list_S_P_P.ri[m]; k < s->list_S_P_P.ri[m + 1]; k++) hp[s->list_S_P_P.ci[k]] = 0;
            if (amoeba)
                for (int k = s->list_S_P_D.ri[m]; k < s->list_S_P_D.ri[m + 1]; k++) hd[s->list_S_P_D.ci[k]] = 0;
        }
        free(sp); free(sd); free(hp); free(hd);
    }
}

/* rotation_matrix (rotate_multipoles.F90:298-463) in complex arithmetic: the
   frame vectors of site c from the z, x and y reference atoms.  Evaluated
   with an imaginary perturbation of one coordinate it returns that
   coordinate's derivative in the imaginary part (complex-step
   differentiation, Squire & Trapp, SIAM Rev. 40 (1998) 110), which is exact
   to rounding -- the reference instead chains hand-written derivatives
   (:464-642).  Returns 1 for an ill-defined frame. */
static int rot_cplx(int frame, const double complex c[3], const double complex cx[3],
                    const double complex cy[3], const double complex cz[3], double complex r[3][3])
{
    double complex xi[3], eta[3], zeta[3], u[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
        xi[k] = cz[k] - c[k]; eta[k] = cx[k] - c[k]; zeta[k] = cy[k] - c[k];
        for (int l = 0; l < 3; l++) r[k][l] = k == l ? 1.0 : 0.0;
    }
    if (frame == 0) return 0;
    double complex xn = csqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
    double complex en = csqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
    double complex zn = csqrt(zeta[0] * zeta[0] + zeta[1] * zeta[1] + zeta[2] * zeta[2]);
    for (int k = 0; k < 3; k++) {
        switch (frame) {
        case 1: u[k] = xi[k]; v[k] = eta[k]; break;
        case 2: u[k] = en * xi[k] + xn * eta[k]; v[k] = eta[k]; break;
        case 3: u[k] = xi[k]; break;
        case 4: u[k] = xi[k]; v[k] = zn * eta[k] + en * zeta[k]; break;
        case 5: u[k] = en * zn * xi[k] + xn * zn * eta[k] + xn * en * zeta[k]; v[k] = eta[k]; break;
        default: return 1;
        }
    }
    if (frame == 3) {
        /* constant v: picked by the real geometry only */
        if (creal(u[2]) / creal(xn) <= 0.866 && fabs(creal(u[1])) > EPS_RP) v[0] = 1.0;
        else v[1] = 1.0;
    }
    double un = sqrt(creal(u[0]) * creal(u[0]) + creal(u[1]) * creal(u[1]) + creal(u[2]) * creal(u[2]));
    double vn = sqrt(creal(v[0]) * creal(v[0]) + creal(v[1]) * creal(v[1]) + creal(v[2]) * creal(v[2]));
    if (un < EPS_RP || vn < EPS_RP) return 1;
    double complex uu = csqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), ez[3], ex[3], ey[3];
    for (int k = 0; k < 3; k++) ez[k] = u[k] / uu;
    double complex vez = v[0] * ez[0] + v[1] * ez[1] + v[2] * ez[2];
    for (int k = 0; k < 3; k++) ex[k] = v[k] - vez * ez[k];
    double complex exn = csqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
    for (int k = 0; k < 3; k++) ex[k] /= exn;
    ey[0] = ez[1] * ex[2] - ez[2] * ex[1];
    ey[1] = ez[2] * ex[0] - ez[0] * ex[2];
    ey[2] = ez[0] * ex[1] - ez[1] * ex[0];
    for (int k = 0; k < 3; k++) { r[k][0] = ex[k]; r[k][1] = ey[k]; r[k][2] = ez[k]; }
    return 0;
}

/* rotation_geomgrad (rotate_multipoles.F90:3-230): gradient of the energy of
   the rotated multipoles in a field E[nmm][3] and field gradient
   Egrd[nmm][6] through the frame atoms of each site, ADDED to grad[nmm][3]:
   for every site j and every coordinate t of j and of its z, x, y frame
   atoms,  dE/dt = -(dR/dt mu0) . E_j + sum_ab dQ_ab Egrd_j,ab  with
   dQ = dR Q0 R^T + R Q0 dR^T.  q0 [nmm][10] unrotated (quadrupoles /3),
   frames [nmm][4] = (convention, z, x, y atom; 0-based, -1 = none).
   The reference's z-bisector branch (:575-587) stores the diagonal of
   d v / d r_y into the u derivative; this restatement differentiates the
   frame exactly, so the two differ for that convention only.  Returns 1 on
   an ill-defined frame. */
int oracle_rotation_geomgrad(int nmm, const double *cmm, const double *q0, const int *frames,
                             const double *E, const double *Egrd, double *grad)
{
    const double hstep = 1e-30;
    for (int j = 0; j < nmm; j++) {
        const int fr = frames[4 * j];
        if (fr == 0) continue;
        int at[4] = {j, frames[4 * j + 2], frames[4 * j + 3], frames[4 * j + 1]}; /* c, x, y, z */
        for (int r = 1; r < 4; r++) if (at[r] < 0) at[r] = j;
        const double *mu0 = q0 + 10 * j + 1, *Qs = q0 + 10 * j + 4;
        const double Q0[3][3] = {{Qs[0], Qs[1], Qs[3]}, {Qs[1], Qs[2], Qs[4]}, {Qs[3], Qs[4], Qs[5]}};
        const double *g6 = Egrd + 6 * j;
        const double G[3][3] = {{g6[0], g6[1], g6[3]}, {g6[1], g6[2], g6[4]}, {g6[3], g6[4], g6[5]}};
        for (int role = 0; role < 4; role++)
            for (int t = 0; t < 3; t++) {
                double complex p[4][3], R[3][3];
                for (int a = 0; a < 4; a++)
                    for (int k = 0; k < 3; k++) p[a][k] = cmm[3 * at[a] + k];
                p[role][t] += I * hstep;
                if (rot_cplx(fr, p[0], p[1], p[2], p[3], R)) return 1;
                double R0[3][3], dR[3][3];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) { R0[a][b] = creal(R[a][b]); dR[a][b] = cimag(R[a][b]) / hstep; }
                double dmu[3], RQ[3][3], dRQ[3][3], acc = 0.0;
                for (int a = 0; a < 3; a++) {
                    dmu[a] = dR[a][0] * mu0[0] + dR[a][1] * mu0[1] + dR[a][2] * mu0[2];
                    for (int b = 0; b < 3; b++) {
                        RQ[a][b] = R0[a][0] * Q0[0][b] + R0[a][1] * Q0[1][b] + R0[a][2] * Q0[2][b];
                        dRQ[a][b] = dR[a][0] * Q0[0][b] + dR[a][1] * Q0[1][b] + dR[a][2] * Q0[2][b];
                    }
                    acc -= dmu[a] * E[3 * j + a];
                }
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        double dq = 0.0;
                        for (int c = 0; c < 3; c++) dq += dRQ[a][c] * R0[b][c] + RQ[a][c] * dR[b][c];
                        acc += dq * G[a][b];
                    }
                grad[3 * at[role] + t] += acc;
            }
    }
    return 0;
}

/* ------------------------------------------------ fixed multipole energy */
/* Potential, field and field gradient (grdE, storage xx, xy, yy, xz, yz, zz) of
   one static multipole at a target, ADDED to V, E, G: the do_V, do_E and
   do_grdE branches of q_elec_prop / mu_elec_prop / quad_elec_prop
   (mod_electrostatics.F90:1119-1141, 1197-1223, 1307-1345), literally.
   k[0..4] = 1/r, 1/r^3, .., 1/r^9 (coulomb_kernel), dr = r_target - r_source. */
static void m2m_props(const double *q, int amoeba, const double dr[3], const double *k, double *V, double E[3],
                      double G[6])
{
    const double qq = q[0];
    *V += k[0] * qq;
    if (!amoeba) return;
    for (int c = 0; c < 3; c++) E[c] += qq * k[1] * dr[c];
    G[0] += qq * 3.0 * k[2] * dr[0] * dr[0] - qq * k[1];
    G[1] += qq * 3.0 * k[2] * dr[0] * dr[1];
    G[2] += qq * 3.0 * k[2] * dr[1] * dr[1] - qq * k[1];
    G[3] += qq * 3.0 * k[2] * dr[0] * dr[2];
    G[4] += qq * 3.0 * k[2] * dr[1] * dr[2];
    G[5] += qq * 3.0 * k[2] * dr[2] * dr[2] - qq * k[1];
    const double *mu = q + 1;
    const double md = mu[0] * dr[0] + mu[1] * dr[1] + mu[2] * dr[2];
    *V += md * k[1];
    for (int c = 0; c < 3; c++) E[c] += 3.0 * md * dr[c] * k[2] - mu[c] * k[1];
    G[0] += md * k[3] * 15.0 * dr[0] * dr[0] - (md + 2.0 * mu[0] * dr[0]) * 3.0 * k[2];
    G[1] += md * k[3] * 15.0 * dr[0] * dr[1] - (mu[0] * dr[1] + mu[1] * dr[0]) * 3.0 * k[2];
    G[2] += md * k[3] * 15.0 * dr[1] * dr[1] - (md + 2.0 * mu[1] * dr[1]) * 3.0 * k[2];
    G[3] += md * k[3] * 15.0 * dr[0] * dr[2] - (mu[0] * dr[2] + mu[2] * dr[0]) * 3.0 * k[2];
    G[4] += md * k[3] * 15.0 * dr[1] * dr[2] - (mu[1] * dr[2] + mu[2] * dr[1]) * 3.0 * k[2];
    G[5] += md * k[3] * 15.0 * dr[2] * dr[2] - (md + 2.0 * mu[2] * dr[2]) * 3.0 * k[2];
    const double *Q = q + 4;
    double qr[3];
    qr[0] = Q[0] * dr[0] + Q[1] * dr[1] + Q[3] * dr[2];
    qr[1] = Q[1] * dr[0] + Q[2] * dr[1] + Q[4] * dr[2];
    qr[2] = Q[3] * dr[0] + Q[4] * dr[1] + Q[5] * dr[2];
    const double qrr = qr[0] * dr[0] + qr[1] * dr[1] + qr[2] * dr[2];
    *V += 3.0 * qrr * k[2];
    for (int c = 0; c < 3; c++) E[c] += 15.0 * qrr * dr[c] * k[3] - 6.0 * qr[c] * k[2];
    G[0] += 105.0 * k[4] * qrr * dr[0] * dr[0] + 6.0 * k[2] * Q[0] - 15.0 * k[3] * (qrr + 4.0 * qr[0] * dr[0]);
    G[1] += 105.0 * k[4] * qrr * dr[0] * dr[1] + 6.0 * k[2] * Q[1] - 30.0 * k[3] * (qr[0] * dr[1] + qr[1] * dr[0]);
    G[2] += 105.0 * k[4] * qrr * dr[1] * dr[1] + 6.0 * k[2] * Q[2] - 15.0 * k[3] * (qrr + 4.0 * qr[1] * dr[1]);
    G[3] += 105.0 * k[4] * qrr * dr[0] * dr[2] + 6.0 * k[2] * Q[3] - 30.0 * k[3] * (qr[0] * dr[2] + qr[2] * dr[0]);
    G[4] += 105.0 * k[4] * qrr * dr[1] * dr[2] + 6.0 * k[2] * Q[4] - 30.0 * k[3] * (qr[1] * dr[2] + qr[2] * dr[1]);
    G[5] += 105.0 * k[4] * qrr * dr[2] * dr[2] + 6.0 * k[2] * Q[5] - 15.0 * k[3] * (qrr + 4.0 * qr[2] * dr[2]);
}

/* energy_MM_MM (mod_electrostatics.F90:899-941) on the fields of
   prepare_fixedelec -> elec_prop_M2M, direct branch (:1678-1952, AMOEBA
   :1847-1899, AMBER :1900-1949): at every MM site j the potential (AMOEBA: also
   field and field gradient) of every other site i, undamped, a pair of
   list_S_S scaled by its factor and skipped when that is 0; then
   E = 1/2 sum_j [q_j V_j - mu_j . E_j + Theta_j : Egrd_j] (off-diagonal
   gradient entries counted twice).  Per-site sums are kept and added in site
   order, so the result does not depend on the thread count.  Returns 0 and sets
   the fatal flag if two sites coincide. */
double oracle_energy_mm(void *h)
{
    osys *s = (osys *)h;
    const int n = s->nmm, amoeba = s->amoeba, ld = s->ld_cart;
    csr_t tSS;
    csr_transpose(&s->list_S_S, n, &tSS); /* row = target j: sources i with a scale */
    double *ej = (double *)calloc(n ? n : 1, sizeof(double));
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        double *scal = (double *)malloc(sizeof(double) * (n ? n : 1));
        char *has = (char *)calloc(n ? n : 1, 1);
#pragma omp for schedule(dynamic, 16)
        for (int j = 0; j < n; j++) {
            for (int k = tSS.ri[j]; k < tSS.ri[j + 1]; k++) { has[tSS.ci[k]] = 1; scal[tSS.ci[k]] = tSS.val[k]; }
            double V = 0.0, E[3] = {0, 0, 0}, G[6] = {0, 0, 0, 0, 0, 0};
            for (int i = 0; i < n; i++) {
                if (i == j) continue;
                const double scalf = has[i] ? scal[i] : 1.0;
                if (has[i] && !(fabs(scalf) > EPS_RP)) continue; /* todo_S_S false */
                double dr[3], k[5];
                for (int c = 0; c < 3; c++) dr[c] = s->cmm[3 * j + c] - s->cmm[3 * i + c];
                if (coulomb_kernel(dr, amoeba ? 4 : 0, k)) { bad = 1; continue; }
                double tV = 0.0, tE[3] = {0, 0, 0}, tG[6] = {0, 0, 0, 0, 0, 0};
                m2m_props(s->q + (size_t)ld * i, amoeba, dr, k, &tV, tE, tG);
                V += tV * scalf;
                for (int c = 0; c < 3; c++) E[c] += tE[c] * scalf;
                for (int c = 0; c < 6; c++) G[c] += tG[c] * scalf;
            }
            for (int k = tSS.ri[j]; k < tSS.ri[j + 1]; k++) has[tSS.ci[k]] = 0;
            const double *q = s->q + (size_t)ld * j;
            double e = q[0] * V;
            if (amoeba) {
                e -= q[1] * E[0] + q[2] * E[1] + q[3] * E[2];
                /* xx, yy, zz once; xy, xz, yz twice */
                e += q[4] * G[0] + q[6] * G[2] + q[9] * G[5];
                e += 2.0 * (q[5] * G[1] + q[7] * G[3] + q[8] * G[4]);
            }
            ej[j] = e;
        }
        free(scal);
        free(has);
    }
    double tot = 0.0;
    for (int j = 0; j < n; j++) tot += ej[j];
    free(ej);
    csr_free(&tSS);
    if (bad) {
        oracle_fatal(s, "Requesting Coulomb kernel for two atoms placed in the same point");
        return 0.0;
    }
    return 0.5 * tot;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
