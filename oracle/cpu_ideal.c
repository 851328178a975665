/*
 * cpu_ideal.c -- idealised multi-core CPU baseline (TEST/BENCH INFRASTRUCTURE ONLY).
 *
 * Not a restatement: the reference's rayon managers are lock-serialised
 * (lennard_jones.rs:278-280,371-372,429-430).  This is what a lock-free version of the
 * reference's full-list structure (lennard_jones.rs:344-456) would do on all host cores:
 * a skin-triggered full list built over a cell grid, a per-atom force gather with no
 * shared writes, and the verlet_step_nve update (potential.rs:15-33).  Used only as the
 * "ideal parallel" CPU figure beside the faithful single-core oracle.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "pis_oracle.h"

typedef struct {
    int64_t n, maxnb;
    int32_t* nnb;
    int32_t* nb;
    double* ref;
} ideal_list;

static double ideal_max_rcut(const orc_table* t) { return orc_max_rcut(t); }

static int ideal_build(const orc_sys* s, double skin, ideal_list* L)
{
    double rl = ideal_max_rcut(&s->table) + skin;
    int64_t n3[3];
    orc_divide_into_cells(&s->box, rl, n3);
    int64_t nc = n3[0] * n3[1] * n3[2];
    int64_t* cnt = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)s->n);
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)s->n);
    for (int64_t i = 0; i < s->n; ++i) {
        cid[i] = orc_cell_of(&s->box, n3, &s->pos[3 * i]);
        if (cid[i] < 0) cid[i] = 0;
        cnt[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < nc; ++k) cnt[k + 1] += cnt[k];
    int64_t* fill = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t i = 0; i < s->n; ++i) idx[cnt[cid[i]] + fill[cid[i]]++] = i;
    free(fill);
    int overflow = 0;
    int nt = s->table.ntypes;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
    for (int64_t i = 0; i < s->n; ++i) {
        int64_t c = cid[i];
        int64_t cx = c % n3[0], cy = (c / n3[0]) % n3[1], cz = c / (n3[0] * n3[1]);
        int64_t seen[27];
        int ns = 0;
        int32_t k = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t x = ((cx + dx) % n3[0] + n3[0]) % n3[0];
                    int64_t y = ((cy + dy) % n3[1] + n3[1]) % n3[1];
                    int64_t z = ((cz + dz) % n3[2] + n3[2]) % n3[2];
                    int64_t nb = (z * n3[1] + y) * n3[0] + x;
                    int dup = 0;
                    for (int q = 0; q < ns; ++q) dup |= seen[q] == nb;
                    if (dup) continue;
                    seen[ns++] = nb;
                    for (int64_t b = cnt[nb]; b < cnt[nb + 1]; ++b) {
                        int64_t j = idx[b];
                        if (j == i) continue;
                        double d[3] = {s->pos[3 * j] - s->pos[3 * i],
                                       s->pos[3 * j + 1] - s->pos[3 * i + 1],
                                       s->pos[3 * j + 2] - s->pos[3 * i + 2]};
                        orc_min_image(&s->box, d);
                        int32_t ti = s->type[i], tj = s->type[j];
                        int32_t a = ti < tj ? ti : tj, bb = ti < tj ? tj : ti;
                        int kk = (a - 1) * nt + (bb - 1);
                        if (!s->table.present[kk]) continue;
                        double lim = s->table.rcut[kk] + skin;
                        if ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] > lim * lim) continue;
                        if (k < L->maxnb)
                            L->nb[i * L->maxnb + k] = (int32_t)j;
                        else
                            overflow = 1;
                        ++k;
                    }
                }
        L->nnb[i] = k;
        for (int d = 0; d < 3; ++d) L->ref[3 * i + d] = s->pos[3 * i + d];
    }
    free(cnt);
    free(cid);
    free(idx);
    return overflow;
}

static double ideal_force(orc_sys* s, const ideal_list* L)
{
    double pe = 0.0;
    int nt = s->table.ntypes;
#pragma omp parallel for schedule(static) reduction(+ : pe)
    for (int64_t i = 0; i < s->n; ++i) {
        double fi[3] = {0, 0, 0}, tpe = 0.0;
        for (int32_t k = 0; k < L->nnb[i]; ++k) {
            int64_t j = L->nb[i * L->maxnb + k];
            double d[3] = {s->pos[3 * j] - s->pos[3 * i], s->pos[3 * j + 1] - s->pos[3 * i + 1],
                           s->pos[3 * j + 2] - s->pos[3 * i + 2]};
            orc_min_image(&s->box, d);
            int32_t ti = s->type[i], tj = s->type[j];
            int32_t a = ti < tj ? ti : tj, b = ti < tj ? tj : ti;
            int kk = (a - 1) * nt + (b - 1);
            double rc = s->table.rcut[kk];
            if (orc_norm(d) > rc) continue;
            double u, f[3];
            orc_lj_pair(s->table.eps[kk], s->table.sigma[kk], rc, s->table.shift, d, &u, f);
            tpe += u;
            for (int q = 0; q < 3; ++q) fi[q] -= f[q];
        }
        for (int q = 0; q < 3; ++q) s->force[3 * i + q] = fi[q];
        pe += tpe;
    }
    return pe / 2.0;
}

int orc_ideal_run_nve(orc_sys* s, double dt, int64_t nsteps, double skin, int32_t nthreads,
                      double* pe_out)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    ideal_list L;
    L.n = s->n;
    double rl = ideal_max_rcut(&s->table) + skin;
    double vol = orc_box_volume(&s->box);
    double dens = (double)s->n / vol;
    L.maxnb = (int64_t)(2.0 * 4.18879 * rl * rl * rl * dens) + 64;
    L.nnb = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->n);
    L.nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->n * L.maxnb));
    L.ref = (double*)malloc(sizeof(double) * (size_t)(3 * s->n));
    double* at = (double*)malloc(sizeof(double) * (size_t)(3 * s->n));
    if (ideal_build(s, skin, &L)) return -1;
    double pe = ideal_force(s, &L);
    double dt2 = dt * dt, half2 = 0.25 * skin * skin;
    for (int64_t t = 0; t < nsteps; ++t) {
        int rebuild = 0;
#pragma omp parallel for schedule(static) reduction(| : rebuild)
        for (int64_t i = 0; i < s->n; ++i) {
            double m = s->mass[s->type[i] - 1];
            for (int d = 0; d < 3; ++d) {
                at[3 * i + d] = s->force[3 * i + d] / m;
                s->pos[3 * i + d] += s->vel[3 * i + d] * dt + at[3 * i + d] * 0.5 * dt2;
            }
            orc_wrap(&s->box, &s->pos[3 * i]);
            double d[3] = {s->pos[3 * i] - L.ref[3 * i], s->pos[3 * i + 1] - L.ref[3 * i + 1],
                           s->pos[3 * i + 2] - L.ref[3 * i + 2]};
            orc_min_image(&s->box, d);
            if ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] > half2) rebuild = 1;
        }
        if (rebuild && ideal_build(s, skin, &L)) return -1;
        pe = ideal_force(s, &L);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < s->n; ++i) {
            double m = s->mass[s->type[i] - 1];
            for (int d = 0; d < 3; ++d)
                s->vel[3 * i + d] += (at[3 * i + d] + s->force[3 * i + d] / m) * 0.5 * dt;
        }
    }
    if (pe_out) *pe_out = pe;
    free(L.nnb);
    free(L.nb);
    free(L.ref);
    free(at);
    return 0;
}
