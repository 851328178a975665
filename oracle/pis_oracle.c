/*
 * pis_oracle.c -- CPU FP64 restatement of the reference's LJ force/integration path.
 * TEST INFRASTRUCTURE ONLY (the checker); see pis_oracle.h for the parity status of
 * each function.  Compile with -ffp-contract=off: every a*b+c below must round twice,
 * as the Rust reference does.
 */
#include "pis_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define M(a, r, c) ((a)[(c) * 3 + (r)])

/* nalgebra 0.34.1 linalg/inverse.rs 3x3 branch (called from
 * SimulationBox::new, simulation_box.rs:12-15).  Cofactors, determinant by first-row
 * expansion evaluated left to right, then each entry divided by the determinant. */
int orc_box_inverse(const double* h, double* hinv)
{
    double m11 = M(h, 1, 1) * M(h, 2, 2) - M(h, 2, 1) * M(h, 1, 2);
    double m12 = M(h, 1, 0) * M(h, 2, 2) - M(h, 2, 0) * M(h, 1, 2);
    double m13 = M(h, 1, 0) * M(h, 2, 1) - M(h, 2, 0) * M(h, 1, 1);
    double det = M(h, 0, 0) * m11 - M(h, 0, 1) * m12 + M(h, 0, 2) * m13;
    if (det == 0.0) return -1;
    M(hinv, 0, 0) = m11 / det;
    M(hinv, 0, 1) = -(M(h, 0, 1) * M(h, 2, 2) - M(h, 2, 1) * M(h, 0, 2)) / det;
    M(hinv, 0, 2) = (M(h, 0, 1) * M(h, 1, 2) - M(h, 1, 1) * M(h, 0, 2)) / det;
    M(hinv, 1, 0) = -m12 / det;
    M(hinv, 1, 1) = (M(h, 0, 0) * M(h, 2, 2) - M(h, 2, 0) * M(h, 0, 2)) / det;
    M(hinv, 1, 2) = -(M(h, 0, 0) * M(h, 1, 2) - M(h, 1, 0) * M(h, 0, 2)) / det;
    M(hinv, 2, 0) = m13 / det;
    M(hinv, 2, 1) = -(M(h, 0, 0) * M(h, 2, 1) - M(h, 2, 0) * M(h, 0, 1)) / det;
    M(hinv, 2, 2) = (M(h, 0, 0) * M(h, 1, 1) - M(h, 1, 0) * M(h, 0, 1)) / det;
    return 0;
}

/* SimulationBox::from_lammps_data (simulation_box.rs:44-65): columns a=(ax,0,0),
 * b=(xy,ay,0), c=(xz,yz,az); pbc on in every dimension. */
void orc_box_from_lammps(double xlo, double xhi, double ylo, double yhi, double zlo,
                         double zhi, double xy, double xz, double yz, orc_box* out)
{
    memset(out, 0, sizeof(*out));
    M(out->h, 0, 0) = xhi - xlo;
    M(out->h, 0, 1) = xy;
    M(out->h, 1, 1) = yhi - ylo;
    M(out->h, 0, 2) = xz;
    M(out->h, 1, 2) = yz;
    M(out->h, 2, 2) = zhi - zlo;
    orc_box_inverse(out->h, out->hinv);
    out->pbc[0] = out->pbc[1] = out->pbc[2] = 1;
}

/* SimulationBox::volume (simulation_box.rs:67-69): |det h|, nalgebra 3x3 determinant. */
double orc_box_volume(const orc_box* b)
{
    const double* h = b->h;
    double m11 = M(h, 1, 1) * M(h, 2, 2) - M(h, 2, 1) * M(h, 1, 2);
    double m12 = M(h, 1, 0) * M(h, 2, 2) - M(h, 2, 0) * M(h, 1, 2);
    double m13 = M(h, 1, 0) * M(h, 2, 1) - M(h, 2, 0) * M(h, 1, 1);
    return fabs(M(h, 0, 0) * m11 - M(h, 0, 1) * m12 + M(h, 0, 2) * m13);
}

/* nalgebra Matrix3 * Vector3 (gemv, column by column): y = m[:,0]*x0; y = m[:,c]*xc + y. */
void orc_gemv(const double* m, const double* x, double* y)
{
    double t[3];
    for (int r = 0; r < 3; ++r) t[r] = M(m, r, 0) * x[0];
    for (int c = 1; c < 3; ++c)
        for (int r = 0; r < 3; ++r) t[r] = M(m, r, c) * x[c] + t[r];
    y[0] = t[0];
    y[1] = t[1];
    y[2] = t[2];
}

/* apply_boundary_conditions_dis (simulation_box.rs:17-27): s = h^-1 d;
 * s_k -= round(s_k) (half away from zero) on periodic dims; d = h s. */
void orc_min_image(const orc_box* b, double* d)
{
    double s[3];
    orc_gemv(b->hinv, d, s);
    for (int k = 0; k < 3; ++k)
        if (b->pbc[k]) s[k] -= round(s[k]);
    orc_gemv(b->h, s, d);
}

/* apply_boundary_conditions_pos (simulation_box.rs:29-42): s = h^-1 r;
 * s_k -= floor(s_k) on periodic dims; r = h s. */
void orc_wrap(const orc_box* b, double* r)
{
    double s[3];
    orc_gemv(b->hinv, r, s);
    for (int k = 0; k < 3; ++k)
        if (b->pbc[k]) s[k] -= floor(s[k]);
    orc_gemv(b->h, s, r);
}

/* nalgebra norm(): sqrt(0 + ((x*x + y*y) + z*z)) for a 3-vector. */
double orc_norm(const double* v) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }

/* Rust `f as usize`: saturating, NaN -> 0. */
static int64_t sat_usize(double f)
{
    if (!(f > 0.0)) return 0; /* negative, -0, NaN */
    if (f >= 9.2e18) return INT64_MAX;
    return (int64_t)f;
}

/* divide_into_cells (neighbour_list.rs:45-57): n_d = max(1, floor(|h[:,d]| / rcut)). */
void orc_divide_into_cells(const orc_box* b, double rcut, int64_t* ncell)
{
    for (int d = 0; d < 3; ++d) {
        double len = orc_norm(&b->h[d * 3]);
        int64_t n = sat_usize(floor(len / rcut));
        ncell[d] = n == 0 ? 1 : n;
    }
}

/* rcut_cells / cell_index (neighbour_list.rs:17-39,61-63): s = h^-1 r,
 * c_d = floor(s_d * n_d) as usize, c_d %= n_d on periodic dims,
 * index (cz*ny + cy)*nx + cx.  Returns -1 if the index falls outside the grid (the
 * reference would panic on the out-of-range Vec index). */
int64_t orc_cell_of(const orc_box* b, const int64_t* n, const double* r)
{
    double s[3];
    orc_gemv(b->hinv, r, s);
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
        c[d] = sat_usize(floor(s[d] * (double)n[d]));
        if (b->pbc[d]) c[d] %= n[d];
        if (c[d] >= n[d]) return -1;
    }
    return (c[2] * n[1] + c[1]) * n[0] + c[0];
}

/* PairPotentialManager::max_rcut (potential.rs:170-179). */
double orc_max_rcut(const orc_table* t)
{
    double m = 0.0;
    for (int i = 0; i < t->ntypes * t->ntypes; ++i)
        if (t->present[i] && m < t->rcut[i]) m = t->rcut[i];
    return m;
}

/* LennardJones::compute_potential (lennard_jones.rs:32-55), exact op order:
 * r2 = (x^2+y^2)+z^2; inv = 1/r2; a = (sigma*sigma*inv)^3 as x*(x*x); r = a*a;
 * U = (4 eps)(r - a) [- (4 eps)(rc12 - rc6) if shift];
 * F = (((24 eps)(2r - a)) inv) rij  -- the force on j; i receives -F. */
void orc_lj_pair(double eps, double sigma, double rcut, int32_t shift, const double* rij,
                 double* u, double* f)
{
    double rij2 = (rij[0] * rij[0] + rij[1] * rij[1]) + rij[2] * rij[2];
    double inv = 1.0 / rij2;
    double x = (sigma * sigma) * inv;
    double att = x * (x * x);
    double rep = att * att;
    double pe = 4.0 * eps * (rep - att);
    double c = 24.0 * eps * (2.0 * rep - att) * inv;
    f[0] = c * rij[0];
    f[1] = c * rij[1];
    f[2] = c * rij[2];
    if (shift) {
        double ci2 = (sigma / rcut) * (sigma / rcut);
        double ca = ci2 * (ci2 * ci2);
        double cr = ca * ca;
        double ucut = 4.0 * eps * (cr - ca);
        pe -= ucut;
    }
    *u = pe;
}

/* get_potential_ij (potential.rs:181-192): key (min(ti,tj), max(ti,tj)). */
static int pair_params(const orc_sys* s, int64_t i, int64_t j, double* eps, double* sig,
                       double* rc)
{
    int32_t ti = s->type[i], tj = s->type[j];
    int32_t a = ti < tj ? ti : tj, b = ti < tj ? tj : ti;
    int nt = s->table.ntypes;
    if (a < 1 || b > nt) return 0;
    int k = (a - 1) * nt + (b - 1);
    if (!s->table.present[k]) return 0;
    *eps = s->table.eps[k];
    *sig = s->table.sigma[k];
    *rc = s->table.rcut[k];
    return 1;
}

typedef struct {
    int64_t n[3];
    int64_t ncell;
    int64_t* start; /* ncell+1 */
    int64_t* idx;   /* atoms, ascending within each cell */
} cells_t;

/* rcut_cells (neighbour_list.rs:9-42) as CSR; atoms appended in ascending index. */
static int build_cells(const orc_sys* s, double rcut, cells_t* c)
{
    orc_divide_into_cells(&s->box, rcut, c->n);
    c->ncell = c->n[0] * c->n[1] * c->n[2];
    c->start = (int64_t*)calloc((size_t)c->ncell + 1, sizeof(int64_t));
    c->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n > 0 ? s->n : 1));
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n > 0 ? s->n : 1));
    if (!c->start || !c->idx || !cid) return -1;
    for (int64_t i = 0; i < s->n; ++i) {
        cid[i] = orc_cell_of(&s->box, c->n, &s->pos[3 * i]);
        if (cid[i] < 0) {
            free(cid);
            return -2;
        }
        c->start[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < c->ncell; ++k) c->start[k + 1] += c->start[k];
    int64_t* fill = (int64_t*)calloc((size_t)c->ncell, sizeof(int64_t));
    for (int64_t i = 0; i < s->n; ++i) c->idx[c->start[cid[i]] + fill[cid[i]]++] = i;
    free(fill);
    free(cid);
    return 0;
}

static void free_cells(cells_t* c)
{
    free(c->start);
    free(c->idx);
}

static int64_t cell_index(int64_t cx, int64_t cy, int64_t cz, const int64_t* n)
{
    return (cz * n[1] + cy) * n[0] + cx;
}

static int64_t rem_euclid(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* Evaluate one candidate pair exactly as the reference managers do: rij = r_j - r_i,
 * minimum image, table lookup, `norm > rcut` rejects, accumulate f_i -= F, f_j += F. */
static inline void half_pair(orc_sys* s, int64_t i, int64_t j, double* pe)
{
    double rij[3] = {s->pos[3 * j] - s->pos[3 * i], s->pos[3 * j + 1] - s->pos[3 * i + 1],
                     s->pos[3 * j + 2] - s->pos[3 * i + 2]};
    orc_min_image(&s->box, rij);
    double eps, sig, rc;
    if (!pair_params(s, i, j, &eps, &sig, &rc)) return; /* lennard_jones.rs:216-222 */
    if (orc_norm(rij) > rc) return;
    double u, f[3];
    orc_lj_pair(eps, sig, rc, s->table.shift, rij, &u, f);
    *pe += u;
    for (int k = 0; k < 3; ++k) {
        s->force[3 * i + k] -= f[k];
        s->force[3 * j + k] += f[k];
    }
}

/* FORWARD_NEIGHBOUR_OFFSETS (neighbour_list.rs:93-165): self cell plus 13. */
static const int OFF14[14][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 1, 0}, {0, 1, 0}, {1, 1, 0},
                                 {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},
                                 {1, 0, 1},  {-1, 1, 1},  {0, 1, 1},  {1, 1, 1}};

/* LJVOffsetManager::compute_potential (lennard_jones.rs:186-244): cells cx->cy->cz, the
 * 14 forward offsets, i in cell, j in neighbour cell, self cell skips i <= j.  Forces
 * ACCUMULATE into s->force (the caller zeroes them, potential.rs:24). */
double orc_compute_offset(orc_sys* s)
{
    cells_t c;
    if (build_cells(s, orc_max_rcut(&s->table), &c)) return NAN;
    double pe = 0.0;
    for (int64_t cx = 0; cx < c.n[0]; ++cx)
        for (int64_t cy = 0; cy < c.n[1]; ++cy)
            for (int64_t cz = 0; cz < c.n[2]; ++cz) {
                int64_t cur = cell_index(cx, cy, cz, c.n);
                for (int o = 0; o < 14; ++o) {
                    int64_t nb = cell_index(rem_euclid(cx + OFF14[o][0], c.n[0]),
                                            rem_euclid(cy + OFF14[o][1], c.n[1]),
                                            rem_euclid(cz + OFF14[o][2], c.n[2]), c.n);
                    for (int64_t a = c.start[cur]; a < c.start[cur + 1]; ++a)
                        for (int64_t b = c.start[nb]; b < c.start[nb + 1]; ++b) {
                            int64_t i = c.idx[a], j = c.idx[b];
                            if (cur == nb && i <= j) continue;
                            half_pair(s, i, j, &pe);
                        }
                }
            }
    free_cells(&c);
    return pe;
}

/* LJManager::compute_potential (lennard_jones.rs:66-97): all i<j pairs. */
double orc_compute_brute(orc_sys* s)
{
    double pe = 0.0;
    for (int64_t i = 0; i < s->n; ++i)
        for (int64_t j = i + 1; j < s->n; ++j) half_pair(s, i, j, &pe);
    return pe;
}

/* LJVerletManager::compute_potential (lennard_jones.rs:107-176): 27 cells deduplicated
 * per current cell, pairs with i > j. */
double orc_compute_verlet27(orc_sys* s)
{
    cells_t c;
    if (build_cells(s, orc_max_rcut(&s->table), &c)) return NAN;
    double pe = 0.0;
    int64_t seen[27];
    for (int64_t cx = 0; cx < c.n[0]; ++cx)
        for (int64_t cy = 0; cy < c.n[1]; ++cy)
            for (int64_t cz = 0; cz < c.n[2]; ++cz) {
                int64_t cur = cell_index(cx, cy, cz, c.n);
                int nseen = 0;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            int64_t nb = cell_index(rem_euclid(cx + dx, c.n[0]),
                                                    rem_euclid(cy + dy, c.n[1]),
                                                    rem_euclid(cz + dz, c.n[2]), c.n);
                            int dup = 0;
                            for (int q = 0; q < nseen; ++q) dup |= seen[q] == nb;
                            if (dup) continue;
                            seen[nseen++] = nb;
                            for (int64_t a = c.start[cur]; a < c.start[cur + 1]; ++a)
                                for (int64_t b = c.start[nb]; b < c.start[nb + 1]; ++b) {
                                    int64_t i = c.idx[a], j = c.idx[b];
                                    if (i <= j) continue;
                                    half_pair(s, i, j, &pe);
                                }
                        }
            }
    free_cells(&c);
    return pe;
}

double orc_compute(orc_sys* s, int32_t manager)
{
    switch (manager) {
    case ORC_MGR_OFFSET: return orc_compute_offset(s);
    case ORC_MGR_BRUTE: return orc_compute_brute(s);
    case ORC_MGR_VERLET27: return orc_compute_verlet27(s);
    case ORC_MGR_FULLLIST: {
        orc_list l;
        if (orc_build_list(s, 0.0, &l)) return NAN;
        for (int64_t k = 0; k < 3 * s->n; ++k) s->force[k] = 0.0; /* copy_from, :451 */
        double pe = orc_compute_fulllist(s, &l, 0, NULL);
        orc_list_free(&l);
        return pe;
    }
    }
    return NAN;
}

static int list_push(orc_list* l, int64_t* pos, int64_t v)
{
    if (*pos >= l->cap) {
        int64_t nc = l->cap ? l->cap * 2 : 1024;
        int64_t* ni = (int64_t*)realloc(l->idx, sizeof(int64_t) * (size_t)nc);
        if (!ni) return -1;
        l->idx = ni;
        l->cap = nc;
    }
    l->idx[(*pos)++] = v;
    return 0;
}

/* LJVPBuildListManager::build_neighbour_list (lennard_jones.rs:345-415) with every table
 * rcut replaced by rcut + rcut_add (the skin): full list over the 27-cell stencil,
 * i != j, per-pair cutoff, the cell grid from max_rcut + rcut_add.  The reference clears
 * its dedup set inside the dx loop (:410), so with n_x < 3 a cell can be visited twice;
 * that quirk is kept.  Row order: stencil order, then ascending j. */
int orc_build_list(const orc_sys* s, double rcut_add, orc_list* out)
{
    cells_t c;
    memset(out, 0, sizeof(*out));
    out->n = s->n;
    if (build_cells(s, orc_max_rcut(&s->table) + rcut_add, &c)) return -1;
    /* Per-atom rows collected in cell order, then laid out by atom. */
    int64_t** rows = (int64_t**)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(int64_t*));
    int64_t* rlen = (int64_t*)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(int64_t));
    int64_t* rcap = (int64_t*)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(int64_t));
    int64_t seen[27];
    for (int64_t cx = 0; cx < c.n[0]; ++cx)
        for (int64_t cy = 0; cy < c.n[1]; ++cy)
            for (int64_t cz = 0; cz < c.n[2]; ++cz) {
                int64_t cur = cell_index(cx, cy, cz, c.n);
                for (int dx = -1; dx <= 1; ++dx) {
                    int nseen = 0; /* cleared per dx, as at lennard_jones.rs:410 */
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            int64_t nb = cell_index(rem_euclid(cx + dx, c.n[0]),
                                                    rem_euclid(cy + dy, c.n[1]),
                                                    rem_euclid(cz + dz, c.n[2]), c.n);
                            int dup = 0;
                            for (int q = 0; q < nseen; ++q) dup |= seen[q] == nb;
                            if (dup) continue;
                            seen[nseen++] = nb;
                            for (int64_t a = c.start[cur]; a < c.start[cur + 1]; ++a)
                                for (int64_t b = c.start[nb]; b < c.start[nb + 1]; ++b) {
                                    int64_t i = c.idx[a], j = c.idx[b];
                                    if (cur == nb && i == j) continue;
                                    double rij[3] = {s->pos[3 * j] - s->pos[3 * i],
                                                     s->pos[3 * j + 1] - s->pos[3 * i + 1],
                                                     s->pos[3 * j + 2] - s->pos[3 * i + 2]};
                                    orc_min_image(&s->box, rij);
                                    double eps, sig, rc;
                                    if (!pair_params(s, i, j, &eps, &sig, &rc)) continue;
                                    if (orc_norm(rij) > rc + rcut_add) continue;
                                    if (rlen[i] == rcap[i]) {
                                        rcap[i] = rcap[i] ? 2 * rcap[i] : 64;
                                        rows[i] = (int64_t*)realloc(
                                            rows[i], sizeof(int64_t) * (size_t)rcap[i]);
                                    }
                                    rows[i][rlen[i]++] = j;
                                }
                        }
                }
            }
    free_cells(&c);
    out->start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n + 1));
    out->start[0] = 0;
    for (int64_t i = 0; i < s->n; ++i) out->start[i + 1] = out->start[i] + rlen[i];
    out->cap = out->start[s->n] > 0 ? out->start[s->n] : 1;
    out->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)out->cap);
    for (int64_t i = 0; i < s->n; ++i) {
        if (rlen[i]) memcpy(&out->idx[out->start[i]], rows[i], sizeof(int64_t) * (size_t)rlen[i]);
        free(rows[i]);
    }
    free(rows);
    free(rlen);
    free(rcap);
    return 0;
}

/* Brute-force full list with LJManager semantics (lennard_jones.rs:66-85): every j != i,
 * one minimum image per pair, per-pair cutoff rc + rcut_add.  Correct for any grid. */
int orc_brute_list(const orc_sys* s, double rcut_add, orc_list* out)
{
    memset(out, 0, sizeof(*out));
    out->n = s->n;
    out->start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n + 1));
    int64_t pos = 0;
    out->start[0] = 0;
    for (int64_t i = 0; i < s->n; ++i) {
        for (int64_t j = 0; j < s->n; ++j) {
            if (i == j) continue;
            double rij[3] = {s->pos[3 * j] - s->pos[3 * i], s->pos[3 * j + 1] - s->pos[3 * i + 1],
                             s->pos[3 * j + 2] - s->pos[3 * i + 2]};
            orc_min_image(&s->box, rij);
            double eps, sig, rc;
            if (!pair_params(s, i, j, &eps, &sig, &rc)) continue;
            if (orc_norm(rij) > rc + rcut_add) continue;
            if (list_push(out, &pos, j)) return -1;
        }
        out->start[i + 1] = pos;
    }
    return 0;
}

void orc_list_free(orc_list* l)
{
    free(l->start);
    free(l->idx);
    memset(l, 0, sizeof(*l));
}

/* LJVPBuildListManager::compute_potential (lennard_jones.rs:419-455): per-atom gather,
 * tf_i += -F, tpe_i += U, PE = sum(tpe)/2; forces written with `+=` into s->force (the
 * reference overwrites with copy_from; callers that want that zero first).
 * check_rcut=1 adds the `norm > rcut` test, needed when the list carries a skin. */
double orc_compute_fulllist(orc_sys* s, const orc_list* l, int32_t check_rcut, double* pe_atom)
{
    double pe = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        double fi[3] = {0.0, 0.0, 0.0}, tpe = 0.0;
        for (int64_t k = l->start[i]; k < l->start[i + 1]; ++k) {
            int64_t j = l->idx[k];
            double rij[3] = {s->pos[3 * j] - s->pos[3 * i], s->pos[3 * j + 1] - s->pos[3 * i + 1],
                             s->pos[3 * j + 2] - s->pos[3 * i + 2]};
            orc_min_image(&s->box, rij);
            double eps, sig, rc;
            if (!pair_params(s, i, j, &eps, &sig, &rc)) continue;
            if (check_rcut && orc_norm(rij) > rc) continue;
            double u, f[3];
            orc_lj_pair(eps, sig, rc, s->table.shift, rij, &u, f);
            tpe += u;
            for (int d = 0; d < 3; ++d) fi[d] += -f[d];
        }
        for (int d = 0; d < 3; ++d) s->force[3 * i + d] += fi[d];
        if (pe_atom) pe_atom[i] = tpe;
        pe += tpe;
    }
    return pe / 2.0;
}

/* splitmix64's output function (Steele, Lea & Flood 2014; Vigna's constants): the entry
 * hash of orc_list_digest. */
uint64_t orc_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* One atom's row of build_neighbour_list (lennard_jones.rs:377-411) with rcut + rcut_add,
 * visited in the order orc_build_list stores it (dx, dy, dz with the per-dx dedup set,
 * then ascending j); `visit` gets every accepted j with its minimum-image vector. */
#define ORC_ROW_BEGIN(s, c, i, rcut_add)                                                          \
    {                                                                                             \
        const int64_t cur_ = orc_cell_of(&(s)->box, (c)->n, &(s)->pos[3 * (i)]);                   \
        const int64_t cx_ = cur_ % (c)->n[0], cy_ = (cur_ / (c)->n[0]) % (c)->n[1],               \
                      cz_ = cur_ / ((c)->n[0] * (c)->n[1]);                                       \
        int64_t seen_[27];                                                                        \
        for (int dx_ = -1; dx_ <= 1; ++dx_) {                                                     \
            int nseen_ = 0;                                                                       \
            for (int dy_ = -1; dy_ <= 1; ++dy_)                                                   \
                for (int dz_ = -1; dz_ <= 1; ++dz_) {                                             \
                    const int64_t nb_ = cell_index(rem_euclid(cx_ + dx_, (c)->n[0]),              \
                                                   rem_euclid(cy_ + dy_, (c)->n[1]),              \
                                                   rem_euclid(cz_ + dz_, (c)->n[2]), (c)->n);     \
                    int dup_ = 0;                                                                 \
                    for (int q_ = 0; q_ < nseen_; ++q_) dup_ |= seen_[q_] == nb_;                 \
                    if (dup_) continue;                                                           \
                    seen_[nseen_++] = nb_;                                                        \
                    for (int64_t b_ = (c)->start[nb_]; b_ < (c)->start[nb_ + 1]; ++b_) {          \
                        const int64_t j = (c)->idx[b_];                                           \
                        if (cur_ == nb_ && (i) == j) continue;                                    \
                        double rij[3] = {(s)->pos[3 * j] - (s)->pos[3 * (i)],                     \
                                         (s)->pos[3 * j + 1] - (s)->pos[3 * (i) + 1],             \
                                         (s)->pos[3 * j + 2] - (s)->pos[3 * (i) + 2]};            \
                        orc_min_image(&(s)->box, rij);                                            \
                        double eps, sig, rc;                                                      \
                        if (!pair_params((s), (i), j, &eps, &sig, &rc)) continue;                 \
                        if (orc_norm(rij) > rc + (rcut_add)) continue;
#define ORC_ROW_END \
    }               \
    }               \
    }               \
    }

/* Per-atom digest of build_neighbour_list at rcut + rcut_add (the list orc_build_list
 * returns, without storing it): count[i] = row length, hash[i] = sum mod 2^64 of
 * orc_mix64(j) over the row -- independent of the order of the row, so it compares a
 * list built in any order.  Rows are independent: parallel over atoms. */
int orc_list_digest(const orc_sys* s, double rcut_add, uint32_t* count, uint64_t* hash)
{
    cells_t c;
    if (build_cells(s, orc_max_rcut(&s->table) + rcut_add, &c)) return -1;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < s->n; ++i) {
        uint32_t k = 0;
        uint64_t h = 0;
        ORC_ROW_BEGIN(s, &c, i, rcut_add)
        ++k;
        h += orc_mix64((uint64_t)j);
        (void)rij;
        ORC_ROW_END
        count[i] = k;
        hash[i] = h;
    }
    free_cells(&c);
    return 0;
}

/* LJVPBuildListManager::compute_potential (lennard_jones.rs:419-455) for the atoms in
 * idx[0..m) only: f_out[3k..3k+2] = sum over the row of -F_ij and pe_out[k] = sum of U_ij
 * (not halved), in the row order of orc_build_list(s, 0) -- the same bits as
 * orc_compute_fulllist gives those atoms.  pe_out may be NULL. */
int orc_forces_subset(const orc_sys* s, const int64_t* idx, int64_t m, double* f_out, double* pe_out)
{
    cells_t c;
    if (build_cells(s, orc_max_rcut(&s->table), &c)) return -1;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int64_t k = 0; k < m; ++k) {
        const int64_t i = idx[k];
        if (i < 0 || i >= s->n) {
            bad = 1;
            continue;
        }
        double fi[3] = {0.0, 0.0, 0.0}, tpe = 0.0;
        ORC_ROW_BEGIN(s, &c, i, 0.0)
        double u, f[3];
        orc_lj_pair(eps, sig, rc, s->table.shift, rij, &u, f);
        tpe += u;
        for (int d = 0; d < 3; ++d) fi[d] += -f[d];
        ORC_ROW_END
        for (int d = 0; d < 3; ++d) f_out[3 * k + d] = fi[d];
        if (pe_out) pe_out[k] = tpe;
    }
    free_cells(&c);
    return bad ? -2 : 0;
}

/* Periodic pair virial sum_{i<j} r_ij . F_ij at rcut (no reference counterpart: the
 * reference's virial is sum_i r_i . F_i, properties.rs:49-51): per atom the full-list row
 * of orc_build_list(s, 0) (pair function of lennard_jones.rs:32-55; F_ij . r_ij =
 * (24 eps)(2 r12 - r6)), halved, summed over atoms in ascending order. */
double orc_pair_virial(const orc_sys* s)
{
    cells_t c;
    if (build_cells(s, orc_max_rcut(&s->table), &c)) return NAN;
    double* part = (double*)malloc(sizeof(double) * (size_t)(s->n > 0 ? s->n : 1));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < s->n; ++i) {
        double w = 0.0;
        ORC_ROW_BEGIN(s, &c, i, 0.0)
        double u, f[3];
        orc_lj_pair(eps, sig, rc, s->table.shift, rij, &u, f);
        w += (f[0] * rij[0] + f[1] * rij[1]) + f[2] * rij[2];
        ORC_ROW_END
        part[i] = w;
    }
    double sum = 0.0;
    for (int64_t i = 0; i < s->n; ++i) sum += part[i];
    free(part);
    free_cells(&c);
    return 0.5 * sum;
}

/* mass_i (properties.rs:9-13). */
static inline double mass_i(const orc_sys* s, int64_t i) { return s->mass[s->type[i] - 1]; }

/* verlet_step_nve (potential.rs:15-33):
 *   a_t = F/m; x += v dt + (a_t 0.5) dt^2; wrap; F = 0; PE = compute; a = F/m;
 *   v += ((a_t + a) 0.5) dt.  Returns PE(t+dt). */
double orc_verlet_step_nve(orc_sys* s, double dt, int32_t manager)
{
    int64_t n = s->n;
    double* at = (double*)malloc(sizeof(double) * (size_t)(3 * n > 0 ? 3 * n : 1));
    double dt2 = dt * dt;
    for (int64_t i = 0; i < n; ++i) {
        double m = mass_i(s, i);
        for (int d = 0; d < 3; ++d) {
            at[3 * i + d] = s->force[3 * i + d] / m;
            s->pos[3 * i + d] += s->vel[3 * i + d] * dt + at[3 * i + d] * 0.5 * dt2;
        }
    }
    for (int64_t i = 0; i < n; ++i) orc_wrap(&s->box, &s->pos[3 * i]);
    for (int64_t k = 0; k < 3 * n; ++k) s->force[k] = 0.0;
    double pe = orc_compute(s, manager);
    for (int64_t i = 0; i < n; ++i) {
        double m = mass_i(s, i);
        for (int d = 0; d < 3; ++d) {
            double a = s->force[3 * i + d] / m;
            s->vel[3 * i + d] += (at[3 * i + d] + a) * 0.5 * dt;
        }
    }
    free(at);
    return pe;
}

/* kinetic_energy (properties.rs:17-23): serial sum of (0.5 m) |v|^2. */
double orc_kinetic_energy(const orc_sys* s)
{
    double ek = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        const double* v = &s->vel[3 * i];
        ek += 0.5 * mass_i(s, i) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    }
    return ek;
}

/* temerature (properties.rs:28-30,41-43): 2 KE / (3N kB). */
double orc_temperature(const orc_sys* s, double ke, double kb)
{
    return (2.0 * ke) / ((double)(3 * s->n) * kb);
}

/* trace of virial_tensor = positions * forces^T (properties.rs:49-51): sum_i r_i . F_i.
 * The gemm summation order of matrixmultiply is not restated (parity unpinned at the
 * last bits); this sums per component, ascending i. */
double orc_virial_trace(const orc_sys* s)
{
    double w[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d) w[d] += s->pos[3 * i + d] * s->force[3 * i + d];
    return (w[0] + w[1]) + w[2];
}

/* pressure (properties.rs:61-65): (2 KE + tr W) / (3 V). */
double orc_pressure(const orc_sys* s, double ke)
{
    return (2.0 * ke + orc_virial_trace(s)) / (3.0 * orc_box_volume(&s->box));
}

/* Simulation::run for NVE (simulation.rs:8-39, 41-52, 66-86): one initial compute with
 * forces starting from zero, then steps of verlet_step_nve with a thermo record each. */
/* ---- Nose-Hoover chain (ensemble/nvt.rs) ------------------------------------ */

int orc_nhc_new(orc_nhc* c, double t_start, double t_end, double tau, int32_t chain, double kb)
{
    if (chain < 1 || chain > ORC_NHC_MAX) return -1;
    memset(c, 0, sizeof(*c));
    c->chain = chain;
    c->t_start = t_start;
    c->t_end = t_end;
    c->target = t_start; /* new_from_args passes start_temperature as the target */
    c->kb = kb;
    const double q_value = kb * c->target * (tau * tau); /* KB * T * tau.powi(2) */
    double p10 = 1.0;
    for (int i = 0; i < chain; ++i) {
        c->q[i] = q_value / p10; /* 10.0_f64.powi(i): exact powers of ten */
        p10 *= 10.0;
    }
    return 0;
}

void orc_nhc_compute_forces(orc_nhc* c, double ke, int64_t n)
{
    c->g[0] = 2.0 * ke - ((double)(n * 3)) * c->kb * c->target;
    for (int j = 1; j < c->chain; ++j) c->g[j] = c->q[j - 1] * (c->xi[j - 1] * c->xi[j - 1]) - c->kb * c->target;
}

void orc_nhc_propagate_half_step(orc_nhc* c, double dt)
{
    const int j = c->chain - 1;
    c->xi[j] = 0.5 * dt * c->g[j] / c->q[j];
    for (int l = j - 1; l >= 0; --l) c->xi[l] = (0.5 * dt * c->g[l] / c->q[l]) * exp(-0.25 * dt * c->xi[l + 1]);
}

double orc_nhc_kinetic_energy(const orc_nhc* c)
{
    double k = 0.0;
    for (int i = 0; i < c->chain; ++i) k += 0.5 * c->q[i] * (c->xi[i] * c->xi[i]);
    return k;
}

double orc_nhc_potential_energy(const orc_nhc* c, int64_t n)
{
    double p = (double)(n * 3) * c->kb * c->target * c->eta[0];
    for (int i = 1; i < c->chain; ++i) p += c->kb * c->target * c->eta[i];
    return p;
}

void orc_nhc_target(orc_nhc* c, int64_t step, int64_t total)
{
    c->target = c->t_start + ((c->t_end - c->t_start) / (double)total) * (double)step;
}

double orc_verlet_step_nvt_nhc(orc_sys* s, double dt, orc_nhc* c, int32_t manager)
{
    double ke = orc_kinetic_energy(s);
    orc_nhc_compute_forces(c, ke, s->n);
    orc_nhc_propagate_half_step(c, dt);
    const double scale = exp(-0.5 * dt * c->xi[0]);
    for (int64_t k = 0; k < 3 * s->n; ++k) s->vel[k] = s->vel[k] * scale;
    const double pe = orc_verlet_step_nve(s, dt, manager);
    for (int64_t k = 0; k < 3 * s->n; ++k) s->vel[k] = s->vel[k] * scale;
    ke = orc_kinetic_energy(s);
    orc_nhc_compute_forces(c, ke, s->n);
    orc_nhc_propagate_half_step(c, dt);
    return pe;
}

int orc_run_nvt(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb, double t_start,
                double t_end, double tau, int32_t chain, double* thermo)
{
    orc_nhc c;
    if (orc_nhc_new(&c, t_start, t_end, tau, chain, kb) != 0) return -2;
    for (int64_t k = 0; k < 3 * s->n; ++k) s->force[k] = 0.0;
    double pe0 = orc_compute(s, manager);
    if (pe0 != pe0) return -1;
    for (int64_t t = 0; t < nsteps; ++t) {
        const double pe = orc_verlet_step_nvt_nhc(s, dt, &c, manager);
        orc_nhc_target(&c, t, nsteps);
        const double ke = orc_kinetic_energy(s);
        if (thermo) {
            thermo[5 * t + 0] = pe;
            thermo[5 * t + 1] = ke;
            thermo[5 * t + 2] = pe + ke + orc_nhc_kinetic_energy(&c) + orc_nhc_potential_energy(&c, s->n);
            thermo[5 * t + 3] = orc_temperature(s, ke, kb);
            thermo[5 * t + 4] = orc_pressure(s, ke);
        }
    }
    return 0;
}

int orc_run_nve(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb,
                double* thermo)
{
    for (int64_t k = 0; k < 3 * s->n; ++k) s->force[k] = 0.0;
    double pe0 = orc_compute(s, manager);
    if (pe0 != pe0) return -1;
    for (int64_t t = 0; t < nsteps; ++t) {
        double pe = orc_verlet_step_nve(s, dt, manager);
        double ke = orc_kinetic_energy(s);
        if (thermo) {
            thermo[5 * t + 0] = pe;
            thermo[5 * t + 1] = ke;
            thermo[5 * t + 2] = pe + ke;
            thermo[5 * t + 3] = orc_temperature(s, ke, kb);
            thermo[5 * t + 4] = orc_pressure(s, ke);
        }
    }
    return 0;
}

/* ---- MTK barostat (ensemble/npt.rs) and the NPT step --------------------------- */

#define ORC_KB_KJPERMOLEKELVIN 0.0083144621 /* constants.rs:3 */

void orc_mtk_new(orc_mtk* m, double p0, double tau, int64_t n, double t_target)
{
    memset(m, 0, sizeof(*m));
    for (int k = 0; k < 3; ++k) M(m->target, k, k) = p0; /* Matrix3::identity() * p0 */
    m->w = ((double)(3 * n)) * ORC_KB_KJPERMOLEKELVIN * t_target * (tau * tau);
}

/* velocities * velocities^T: (r, c) = sum_i v_i[r] v_i[c], ascending i (the gemm order of
 * matrixmultiply is not restated). */
void orc_kinetic_tensor(const orc_sys* s, double* out)
{
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    for (int64_t i = 0; i < s->n; ++i)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) M(out, r, c) += s->vel[3 * i + r] * s->vel[3 * i + c];
}

/* positions * forces^T: (r, c) = sum_i x_i[r] F_i[c]. */
void orc_virial_tensor(const orc_sys* s, double* out)
{
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    for (int64_t i = 0; i < s->n; ++i)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) M(out, r, c) += s->pos[3 * i + r] * s->force[3 * i + c];
}

void orc_pressure_tensor(const orc_sys* s, double* out)
{
    double kt[9], wt[9];
    orc_kinetic_tensor(s, kt);
    orc_virial_tensor(s, wt);
    const double v = orc_box_volume(&s->box);
    for (int k = 0; k < 9; ++k) out[k] = (kt[k] + wt[k]) / v;
}

/* math.rs:39-41: (a + a^T) * 0.5. */
static void symmetrize(const double* a, double* out)
{
    double t[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) M(t, r, c) = (M(a, r, c) + M(a, c, r)) * 0.5;
    memcpy(out, t, sizeof(t));
}

void orc_mtk_delta_momentum(const orc_mtk* m, const orc_sys* s, double dt, double* out)
{
    double p[9];
    orc_pressure_tensor(s, p);
    const double f = orc_box_volume(&s->box) * 0.5 * dt;
    double d[9];
    for (int k = 0; k < 9; ++k) d[k] = (p[k] - m->target[k]) * f;
    symmetrize(d, out);
}

/* Matrix3 * Matrix3, column by column as gemv: out[:,c] = a * b[:,c]. */
void orc_mat3_mul(const double* a, const double* b, double* out)
{
    double t[9];
    for (int c = 0; c < 3; ++c) orc_gemv(a, &b[3 * c], &t[3 * c]);
    memcpy(out, t, sizeof(t));
}

/* exp(A) by scaling and squaring a Taylor series (Moler & Van Loan, "Nineteen dubious
 * ways to compute the exponential of a matrix", 2003, method 3): A is halved until its
 * 1-norm is at most 1/2, E = sum_{k<=18} A^k/k!, then squared back. */
void orc_mat3_exp(const double* a, double* out)
{
    double norm = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double col = (fabs(M(a, 0, c)) + fabs(M(a, 1, c))) + fabs(M(a, 2, c));
        norm = col > norm ? col : norm;
    }
    int sq = 0;
    double scale = 1.0;
    while (norm * scale > 0.5 && sq < 64) {
        scale *= 0.5;
        ++sq;
    }
    double x[9], e[9], t[9];
    for (int k = 0; k < 9; ++k) {
        x[k] = a[k] * scale;
        e[k] = 0.0;
        t[k] = 0.0;
    }
    for (int k = 0; k < 3; ++k) M(e, k, k) = M(t, k, k) = 1.0;
    for (int j = 1; j <= 18; ++j) {
        orc_mat3_mul(t, x, t);
        for (int k = 0; k < 9; ++k) {
            t[k] = t[k] / (double)j;
            e[k] = e[k] + t[k];
        }
    }
    for (int q = 0; q < sq; ++q) orc_mat3_mul(e, e, e);
    memcpy(out, e, sizeof(e));
}

void orc_mtk_scale(const orc_mtk* m, double dt, int32_t velocity_scaling, double* out)
{
    double eta[9], a[9];
    for (int k = 0; k < 9; ++k) eta[k] = m->momentum[k] / m->w;
    symmetrize(eta, eta);
    const double factor = velocity_scaling ? -0.5 : 1.0;
    for (int k = 0; k < 9; ++k) a[k] = eta[k] * factor * dt;
    orc_mat3_exp(a, out);
}

/* (M M^T).trace() / (2 w) */
double orc_mtk_kinetic_energy(const orc_mtk* m)
{
    double mt[9], p[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) M(mt, r, c) = M(m->momentum, c, r);
    orc_mat3_mul(m->momentum, mt, p);
    return ((M(p, 0, 0) + M(p, 1, 1)) + M(p, 2, 2)) / (2.0 * m->w);
}

/* (P0^T h).trace() */
double orc_mtk_potential_energy(const orc_mtk* m, const double* h)
{
    double tt[9], p[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) M(tt, r, c) = M(m->target, c, r);
    orc_mat3_mul(tt, h, p);
    return (M(p, 0, 0) + M(p, 1, 1)) + M(p, 2, 2);
}

/* s = h^-1 r per atom; h = scale h; h^-1 = inverse(h); r = h s. */
void orc_scale_box(orc_sys* s, const double* scale)
{
    double* sf = (double*)malloc(sizeof(double) * 3 * (size_t)s->n);
    if (!sf) return;
    for (int64_t i = 0; i < s->n; ++i) orc_gemv(s->box.hinv, &s->pos[3 * i], &sf[3 * i]);
    orc_mat3_mul(scale, s->box.h, s->box.h);
    orc_box_inverse(s->box.h, s->box.hinv);
    for (int64_t i = 0; i < s->n; ++i) orc_gemv(s->box.h, &sf[3 * i], &s->pos[3 * i]);
    free(sf);
}

static void scale_velocities(orc_sys* s, const double* scale)
{
    for (int64_t i = 0; i < s->n; ++i) {
        double v[3];
        orc_gemv(scale, &s->vel[3 * i], v);
        memcpy(&s->vel[3 * i], v, sizeof(v));
    }
}

double orc_verlet_step_npt_mtk(orc_sys* s, double dt, orc_mtk* m, orc_nhc* c, int32_t manager)
{
    double d[9], scale[9], scale_h[9];
    orc_mtk_delta_momentum(m, s, dt, d);
    for (int k = 0; k < 9; ++k) m->momentum[k] += d[k];
    orc_mtk_scale(m, dt, 1, scale);
    scale_velocities(s, scale);
    orc_mtk_scale(m, dt, 0, scale_h);
    orc_scale_box(s, scale_h);
    const double pe = orc_verlet_step_nvt_nhc(s, dt, c, manager);
    scale_velocities(s, scale);
    orc_mtk_delta_momentum(m, s, dt, d);
    for (int k = 0; k < 9; ++k) m->momentum[k] += d[k];
    return pe;
}

int orc_run_npt(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb, double t_start,
                double t_end, double tau_t, int32_t chain, double p0, double tau_p, double* thermo,
                double* hbox)
{
    orc_nhc c;
    orc_mtk m;
    if (orc_nhc_new(&c, t_start, t_end, tau_t, chain, kb) != 0) return -2;
    orc_mtk_new(&m, p0, tau_p, s->n, t_start);
    for (int64_t k = 0; k < 3 * s->n; ++k) s->force[k] = 0.0;
    double pe0 = orc_compute(s, manager);
    if (pe0 != pe0) return -1;
    for (int64_t t = 0; t < nsteps; ++t) {
        const double pe = orc_verlet_step_npt_mtk(s, dt, &m, &c, manager);
        orc_nhc_target(&c, t, nsteps);
        const double ke = orc_kinetic_energy(s);
        if (thermo) {
            thermo[5 * t + 0] = pe;
            thermo[5 * t + 1] = ke;
            thermo[5 * t + 2] = pe + ke + orc_nhc_kinetic_energy(&c) + orc_nhc_potential_energy(&c, s->n) +
                                orc_mtk_kinetic_energy(&m) + orc_mtk_potential_energy(&m, s->box.h);
            thermo[5 * t + 3] = orc_temperature(s, ke, kb);
            thermo[5 * t + 4] = orc_pressure(s, ke);
        }
        if (hbox) memcpy(&hbox[9 * t], s->box.h, sizeof(double) * 9);
    }
    return 0;
}
