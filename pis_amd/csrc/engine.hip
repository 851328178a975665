// engine.hip -- host side of the MI355X LJ engine: planning, device memory, the
// device-resident step loop and the C ABI of include/pis_hip.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"

using namespace pis;

// ---------------------------------------------------------------------------------
// errors

static thread_local std::string g_last_error;

pis_status pis::fail(pis_status s, const std::string& msg)
{
    g_last_error = msg;
    return s;
}

const std::string& pis::last_error() { return g_last_error; }

// ---------------------------------------------------------------------------------
// planning (host only)

namespace {

// Largest double r2 with sqrt_rn(r2) <= rc, so that `norm() > rc` (the reference's
// rejection, lennard_jones.rs:224) is exactly `r2 > r2max` for the same r2 bits.
double r2max_for(double rc)
{
    if (!(rc > 0.0) || !std::isfinite(rc)) return -1.0;
    double r2 = rc * rc;
    while (std::sqrt(r2) > rc) r2 = std::nextafter(r2, 0.0);
    for (;;) {
        const double nx = std::nextafter(r2, INFINITY);
        if (std::sqrt(nx) <= rc)
            r2 = nx;
        else
            break;
    }
    return r2;
}

int scap_for(double need)
{
    const int opts[7] = {1024, 1280, 1536, 1792, 2048, 3072, 4096};
    for (int k = 0; k < 7; ++k)
        if (need <= opts[k]) return opts[k];
    return -1;
}

// Global tile override used by the autotuner (0 = heuristic).
int g_tile_override[3] = {0, 0, 0};

// LDS of the force kernel for a staging class: three coordinates per slot (SCAP + 1 slots;
// FP64, or FP32 in mixed precision) plus the halo tables, the pair tables and the block-sum
// scratch (as compiled).
double tile_lds(int scap, bool f32) { return (scap + 1) * (f32 ? 12.0 : 24.0) + 4800.0; }
// Workgroups per CU the force kernel's registers allow (one wave per SIMD each).
double force_wg_cap(bool f32) { return f32 ? 4.0 : 3.0; }

// Tile choice: maximise lane use of the 256-thread workgroup times an occupancy factor
// (workgroups per CU the LDS allows, capped at the 3 the register file allows), with a
// small penalty for halo re-staging.  Staging capacity gets a margin for density
// fluctuations; an overflow at run time is detected and reported, never silent.
void choose_tiles(Plan& p)
{
    const double cellvol = p.cellw[0] * p.cellw[1] * p.cellw[2];
    double best = -1e300;
    int bt[3] = {1, 1, 1};
    const int lim[3] = {std::max(1, p.nc[0] - 2), std::max(1, p.nc[1] - 2), std::max(1, p.nc[2] - 2)};
    // First only tiles planned into classes of at most 3072 slots: a configuration denser
    // than the average (lattice planes on cell faces) then still has the 4096 class to grow
    // into when its halos are measured; the largest class only if nothing else fits.
    for (int pass = 0; pass < 2 && best == -1e300; ++pass)
    for (int tz = 1; tz <= std::min(lim[2], 8); ++tz)
        for (int ty = 1; ty <= std::min(lim[1], 8); ++ty)
            for (int tx = 1; tx <= std::min(lim[0], 8); ++tx) {
                if ((tx + 2) * (ty + 2) * (tz + 2) > kMaxHalo) continue;
                const double own = p.density * cellvol * tx * ty * tz;
                const double stage = p.density * p.halo_scale * cellvol * (tx + 2) * (ty + 2) * (tz + 2);
                const int sc = scap_for(stage * 1.1 + 48);
                if (sc < 0 || (pass == 0 && sc > 3072)) continue;
                const double passes = std::ceil((own * 1.06 + 4) / kTileThreads);
                const double eff = own / (passes * kTileThreads);
                const double cap = force_wg_cap(p.mixed);
                const double wg = std::min(cap, std::floor(163840.0 / tile_lds(sc, p.mixed)));
                const double score = eff * (0.5 + 0.5 * wg / cap) - 0.004 * (stage / std::max(own, 1.0)) -
                                     0.01 * (std::max(tx, std::max(ty, tz)) - std::min(tx, std::min(ty, tz)));
                if (score > best) {
                    best = score;
                    bt[0] = tx;
                    bt[1] = ty;
                    bt[2] = tz;
                }
            }
    const bool req = p.tile_req[0] > 0 && p.tile_req[1] > 0 && p.tile_req[2] > 0;
    if (req || g_tile_override[0] > 0) {
        int ov[3];
        for (int d = 0; d < 3; ++d)
            ov[d] = std::max(1, std::min(req ? p.tile_req[d] : g_tile_override[d], lim[d]));
        if ((ov[0] + 2) * (ov[1] + 2) * (ov[2] + 2) <= kMaxHalo)
            for (int d = 0; d < 3; ++d) bt[d] = ov[d];
    }
    for (int d = 0; d < 3; ++d) p.tile[d] = bt[d];
}

// Rounding bands of every pair threshold.  FP64: r^2 computed from staged (tile-relative
// or raw) coordinates versus the reference's r^2 -- the displacement differs by at most
// ~4 ulp(qmax) in each component plus a few relative ulps, so relative to r^2 at distance
// r the gap is below 8 qmax 2^-52 / r + 11 * 2^-53; the band is 16 times that (twice for a
// general box, whose minimum image takes two 3-term products).
void set_bands(Plan* p)
{
    {
        const double Lmax = p->qmax * (p->tri ? 2.0 : 1.0);
        const double u = std::ldexp(1.0, -52);
        auto band = [&](double r) { return 16.0 * (8.0 * Lmax * u / r + 5.5 * u); };
        for (auto& pp : p->pp) {
            if (pp.r2_force < 0.0) {
                pp.f_lo = pp.f_hi = pp.l_lo = pp.l_hi = -1.0;
                continue;
            }
            const double bf = band(std::sqrt(pp.r2_force)), bl = band(std::sqrt(pp.r2_list));
            pp.f_lo = pp.r2_force * (1.0 - bf);
            pp.f_hi = pp.r2_force * (1.0 + bf);
            pp.l_lo = pp.r2_list * (1.0 - bl);
            pp.l_hi = pp.r2_list * (1.0 + bl);
        }
    }
    // Band for r^2 from the list build's FP32 staging.  Staged coordinates lie within
    // qmax of the origin (the halo half-extent for tile-relative staging, the box for raw
    // positions) and round to FP32 with error 2^-24 qmax, so a displacement component is
    // within e = 6 * 2^-24 * qmax of its exact value (difference and image shift rounded
    // once more), and at distance r, r^2 is within 2 sqrt(3) e r + 3 e^2 + 3 * 2^-24 r^2.
    // The band is 4 times that; the x windows widen by twice the coordinate error.
    {
        const double u32 = std::ldexp(1.0, -24);
        double qmax = 0.0, cwmax = 0.0;
        for (int k = 0; k < 3; ++k) {
            qmax = std::max(qmax, p->mode == MODE_TILE_REL ? 0.5 * (p->tile[k] + 2) * p->cellw[k] : p->L[k]);
            cwmax = std::max(cwmax, p->cellw[k]);
        }
        // a general box with raw staging: positions up to qmax per component, and the FP32
        // image takes two 3-term products (twice the error).  Tile-relative staging of a
        // general box: a halo spans (tile + 2) / (2 nc) of each fractional axis around the
        // tile's centre, atoms drift up to skin/2 out of their cells, and a barostat may
        // grow the box by the factor qmax holds over the starting extent.
        bool img32 = false;
        if (p->tri && p->mode == MODE_TILE_REL) {
            double ext = 0.0, half = 0.0;
            for (int r = 0; r < 3; ++r) {
                double q = 0.0, e = 0.0;
                for (int k = 0; k < 3; ++k) {
                    q += std::fabs(p->h[r + 3 * k]) * (p->tile[k] + 2) / (2.0 * p->nc[k]);
                    e += std::fabs(p->h[r + 3 * k]);
                }
                half = std::max(half, q);
                ext = std::max(ext, e);
            }
            qmax = half * std::max(1.0, p->qmax / ext) + p->skin;
        } else if (p->tri) {
            qmax = p->qmax;
            img32 = true;
        }
        const double e = (img32 ? 12.0 : 6.0) * u32 * qmax;
        p->wslack = 2.0 * u32 * qmax + 1e-9 * cwmax;
        for (auto& q : p->pp) {
            if (q.r2_list < 0.0) {
                q.l_lo32 = q.l_hi32 = -1.0f;
                continue;
            }
            const double R = std::sqrt(q.r2_list);
            const double beta = 4.0 * ((2.0 * std::sqrt(3.0) * e * R + 3.0 * e * e) / q.r2_list + 3.0 * u32);
            const double lo = q.r2_list * (1.0 - beta), hi = q.r2_list * (1.0 + beta);
            float flo = (float)lo, fhi = (float)hi;
            if ((double)flo > lo) flo = std::nextafter(flo, -INFINITY);
            if ((double)fhi < hi) fhi = std::nextafter(fhi, INFINITY);
            q.l_lo32 = std::max(flo, 0.0f);
            q.l_hi32 = fhi;
        }
    }
}

pis_status plan_general_box(Plan* p);
pis_status plan_grid(Plan* p, double margin);

pis_status make_plan_impl(const pis_system_desc* d, Plan* p, bool need_types)
{
    if (!d) return fail(PIS_ERR_INVALID_ARGUMENT, "null system descriptor");
    if (d->n_atoms == 0) return fail(PIS_ERR_NO_ATOMS, "No atoms defined in input file");
    if (d->n_atoms > (uint64_t)1 << 30) return fail(PIS_ERR_UNSUPPORTED, "more than 2^30 atoms");
    if (d->flags != 0) return fail(PIS_ERR_INVALID_ARGUMENT, "pis_system_desc.flags must be 0");
    if (d->n_types == 0 || d->n_types > (uint32_t)kMaxTypes)
        return fail(PIS_ERR_INVALID_ATOM_TYPE, "n_types must be 1.." + std::to_string(kMaxTypes));
    if ((need_types && !d->type_ids) || !d->masses) return fail(PIS_ERR_INVALID_ARGUMENT, "type_ids/masses missing");
    if (d->n_pairs == 0 || !d->pairs)
        return fail(PIS_ERR_MISSING_POTENTIAL,
                    "Potential manager not initialized - missing pair_style or pair_coeff commands");
    if (d->precision > 1) return fail(PIS_ERR_INVALID_ARGUMENT, "unknown precision");
    if (!d->pbc[0] || !d->pbc[1] || !d->pbc[2])
        return fail(PIS_ERR_UNSUPPORTED, "only fully periodic boxes are supported");
    p->tri = false;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            if (r != c && (d->h[c * 3 + r] != 0.0 || d->h_inv[c * 3 + r] != 0.0)) p->tri = true;
    for (int k = 0; k < 9; ++k) {
        p->h[k] = d->h[k];
        p->hinv[k] = d->h_inv[k];
        if (!std::isfinite(d->h[k]) || !std::isfinite(d->h_inv[k]))
            return fail(PIS_ERR_INVALID_ARGUMENT, "box matrix must be finite");
    }
    if (!(d->skin >= 0.0) || !std::isfinite(d->skin)) return fail(PIS_ERR_INVALID_ARGUMENT, "bad skin");
    if (!(d->kb > 0.0)) return fail(PIS_ERR_INVALID_ARGUMENT, "kb must be > 0");
    const int nt = (int)d->n_types;
    p->n = (int64_t)d->n_atoms;
    for (int k = 0; k < 3; ++k) p->tile_req[k] = d->tile[k];
    p->ntypes = nt;
    p->skin = d->skin;
    p->kb = d->kb;
    p->mixed = d->precision == PIS_PRECISION_MIXED;
    if (p->tri) {
        PIS_TRY(plan_general_box(p));
    } else {
        for (int k = 0; k < 3; ++k) {
            p->L[k] = d->h[k * 3 + k];
            p->Linv[k] = d->h_inv[k * 3 + k];
            if (!(p->L[k] > 0.0) || !(p->Linv[k] > 0.0) || !std::isfinite(p->L[k]))
                return fail(PIS_ERR_INVALID_ARGUMENT, "box lengths must be positive");
        }
        p->volume = std::fabs(p->L[0] * p->L[1] * p->L[2]);
        p->qmax = std::max(p->L[0], std::max(p->L[1], p->L[2]));
    }
    p->mass.resize(nt);
    for (int t = 0; t < nt; ++t) {
        if (!(d->masses[t] > 0.0)) return fail(PIS_ERR_INVALID_ARGUMENT, "masses must be > 0");
        p->mass[t] = d->masses[t];
    }
    p->type0.resize(need_types ? p->n : 0);
    for (int64_t i = 0; need_types && i < p->n; ++i) {
        const uint32_t t = d->type_ids[i];
        if (t < 1 || t > (uint32_t)nt)
            return fail(PIS_ERR_INVALID_ATOM_TYPE, "Atom type " + std::to_string(t) + " out of range");
        p->type0[i] = (uint8_t)(t - 1);
    }
    // pair table: (min,max) keyed like the reference; absent pairs never interact
    p->pp.assign((size_t)nt * nt, PairParam{});
    for (auto& q : p->pp) {
        q.r2_force = -1.0;
        q.r2_list = -1.0;
        q.f_lo = q.f_hi = q.l_lo = q.l_hi = -1.0;
    }
    p->rcmax = 0.0;
    for (uint32_t k = 0; k < d->n_pairs; ++k) {
        const pis_pair_coeff& c = d->pairs[k];
        if (c.type_i < 1 || c.type_j < 1 || c.type_i > (uint32_t)nt || c.type_j > (uint32_t)nt)
            return fail(PIS_ERR_INVALID_ATOM_TYPE, "pair_coeff type out of range");
        if (!(c.rcut > 0.0) || !(c.sigma > 0.0) || !std::isfinite(c.epsilon))
            return fail(PIS_ERR_INVALID_ARGUMENT, "pair_coeff needs sigma > 0, rcut > 0");
        const int a = (int)std::min(c.type_i, c.type_j) - 1, b = (int)std::max(c.type_i, c.type_j) - 1;
        PairParam q;
        q.r2_force = r2max_for(c.rcut);
        q.r2_list = r2max_for(c.rcut + d->skin);
        q.sig2 = c.sigma * c.sigma;
        q.c4 = 4.0 * c.epsilon;
        q.c24 = 24.0 * c.epsilon;
        q.c48 = 48.0 * c.epsilon;
        q.ucut = 0.0;
        if (d->shift) {  // lennard_jones.rs:44-52
            const double ci2 = (c.sigma / c.rcut) * (c.sigma / c.rcut);
            const double ca = ci2 * (ci2 * ci2);
            const double cr = ca * ca;
            q.ucut = 4.0 * c.epsilon * (cr - ca);
        }
        q.sig2f = (float)q.sig2;
        q.c4f = (float)q.c4;
        q.c24f = (float)q.c24;
        q.c48f = (float)q.c48;
        q.ucutf = (float)q.ucut;
        q.r2f = (float)(c.rcut * c.rcut);
        q.isig = 1.0 / c.sigma;
        q.r2s32 = (float)((c.rcut / c.sigma) * (c.rcut / c.sigma));
        p->pp[(size_t)a * nt + b] = q;
        p->pp[(size_t)b * nt + a] = q;
        p->rcmax = std::max(p->rcmax, c.rcut);
    }
    p->rcl = p->rcmax + d->skin;
    p->nbcap_req = d->neighbor_capacity;
    return plan_grid(p, 1e-12);
}

// Box geometry of a general box from p->h: SimulationBox::volume (simulation_box.rs:67-69),
// |det h|, and the perpendicular widths V / |a_i x a_j| the cells follow; qmax bounds the
// components of a wrapped position.
pis_status plan_general_box(Plan* p)
{
    const double* h = p->h;
    const double m11 = h[4] * h[8] - h[5] * h[7];
    const double m12 = h[1] * h[8] - h[2] * h[7];
    const double m13 = h[1] * h[5] - h[2] * h[4];
    p->volume = std::fabs(h[0] * m11 - h[3] * m12 + h[6] * m13);
    if (!(p->volume > 0.0)) return fail(PIS_ERR_INVALID_ARGUMENT, "box matrix is singular");
    for (int k = 0; k < 3; ++k) {
        const double* a = h + 3 * ((k + 1) % 3);
        const double* b = h + 3 * ((k + 2) % 3);
        const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
        p->L[k] = p->volume / std::sqrt(cx * cx + cy * cy + cz * cz);
        p->Linv[k] = 1.0 / p->L[k];
    }
    p->qmax = 0.0;
    for (int r = 0; r < 3; ++r)
        p->qmax = std::max(p->qmax, std::fabs(h[r]) + std::fabs(h[3 + r]) + std::fabs(h[6 + r]));
    return PIS_OK;
}

// Cell grid, tiles, staging mode and class, rounding bands and list capacity of a plan
// whose box geometry and pair table are set; cells are at least rcl (1 + margin) wide.
pis_status plan_grid(Plan* p, double margin)
{
    p->grid_margin = margin;
    p->density = (double)p->n / p->volume;
    // Cell grid: edge strictly larger than rcut + skin (relative margin far above
    // rounding) so every pair within the list cutoff is in adjacent cells.
    bool thin = false;
    for (int k = 0; k < 3; ++k) {
        const double nf = std::floor(p->L[k] / (p->rcl * (1.0 + margin)));
        p->nc[k] = (int)std::max(1.0, std::min(nf, 1024.0));
        p->cellw[k] = p->L[k] / p->nc[k];
        if (p->nc[k] < 3) thin = true;
    }
    p->ncell = p->nc[0] * p->nc[1] * p->nc[2];
    const double vsphere = 4.0 / 3.0 * M_PI * p->rcl * p->rcl * p->rcl;
    if (thin) {
        p->mode = MODE_WHOLE;
        for (int k = 0; k < 3; ++k) {
            p->tile[k] = p->nc[k];
            p->nt[k] = 1;
        }
        p->scap = scap_for((double)p->n + 1);
        if (p->scap < 0)
            return fail(PIS_ERR_UNSUPPORTED,
                        "box thinner than 3 cells of rcut+skin holds more than 4095 atoms");
    } else {
        choose_tiles(*p);
        const double cellvol = p->cellw[0] * p->cellw[1] * p->cellw[2];
        p->mode = MODE_TILE_REL;
        for (int k = 0; k < 3; ++k) {
            p->nt[k] = (p->nc[k] + p->tile[k] - 1) / p->tile[k];
            // tile-relative staging needs half the tile + skin/2 + rcut below L/2
            if (!(p->tile[k] * p->cellw[k] + p->skin + 2.0 * p->rcmax < p->L[k] * (1.0 - 1e-9)))
                p->mode = MODE_RAW_MI;
        }
        // a general box stages tile-relative positions when two cells separate a tile from
        // its own periodic image on every axis: then the condition above holds for any box
        // whose cells stay wider than rcut + skin, as a barostat's boxes must
        if (p->tri)
            for (int k = 0; k < 3; ++k)
                if (p->nc[k] - p->tile[k] < 2) p->mode = MODE_RAW_MI;
        const double stage =
            p->density * p->halo_scale * cellvol * (p->tile[0] + 2) * (p->tile[1] + 2) * (p->tile[2] + 2);
        p->scap = scap_for(stage * 1.1 + 48);
        if (p->scap < 0) return fail(PIS_ERR_UNSUPPORTED, "tile staging exceeds LDS capacity");
    }
    p->ntile = p->nt[0] * p->nt[1] * p->nt[2];
    set_bands(p);
    p->npad = (int)((p->n + 63) / 64 * 64);
    int maxnb;
    if (p->nbcap_req)
        maxnb = (int)p->nbcap_req;
    else
        maxnb = (int)std::ceil(1.6 * p->density * vsphere) + 48;
    if (maxnb > kMaxList) {
        if (p->nbcap_req)
            return fail(PIS_ERR_UNSUPPORTED, "neighbor_capacity above " + std::to_string(kMaxList));
        maxnb = kMaxList;
    }
    if (p->mode == MODE_WHOLE) maxnb = std::min(maxnb, (int)p->n + 8);
    p->maxnb = std::max(8, (maxnb + 7) / 8 * 8);
    return PIS_OK;
}

}  // namespace

size_t pis::align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---------------------------------------------------------------------------------
// context

// Carving of one system's allocation (offsets of every array, 256-byte aligned).
Layout pis::layout_of(const Plan& p)
{
    const size_t np = (size_t)p.npad, nc = (size_t)p.ncell;
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align256(bytes);
        return o;
    };
    for (int s = 0; s < 2; ++s) {
        for (int k = 0; k < 9; ++k) L.st[s][k] = take(np * 8);
        L.st[s][9] = take(np);
        L.st[s][10] = take(np * 4);
    }
    L.r = take(np * 8 * 3);
    L.ctl = take(64 * 4);
    L.cc = take(nc * 4);
    L.cs = take((nc + 1) * 4);
    L.n2s = take(nc * 4);
    L.tc = take(((size_t)p.ntile + 1) * 4);
    L.ac = take(np * 4);
    L.ss = take(np * 4);
    L.sk = take(np * 4);
    L.sx = take(np * 8);
    L.nbr = take((size_t)(p.maxnb / 8) * np * 16);
    L.nnb = take(np * 2);
    L.part = take((size_t)p.ntile * 4 * 8);
    L.pp = take(p.pp.size() * sizeof(PairParam));
    L.m = take(p.mass.size() * 8);
    L.nhc = take(NHC_COUNT * 8);
    L.box = take(BOX_COUNT * 8);
    L.npt = take(NPT_COUNT * 8);
    L.tens = take((size_t)((p.npad + 255) / 256) * TENS_COUNT * 8);
    L.total = off;
    return L;
}


pis_status pis::make_plan(const pis_system_desc* d, Plan* p, bool need_types) { return make_plan_impl(d, p, need_types); }
int pis::stage_class_for(double need) { return scap_for(need); }

size_t pis::system_bytes(const Plan& p, int thermo_records)
{
    return layout_of(p).total + align256(sizeof(ThermoDev) * (size_t)std::max(thermo_records, 0));
}

// Points a system descriptor into memory carved as layout_of(p) lays it out, and fills its
// scalar fields from the plan.
void pis::carve_system(const Plan& p, const Layout& Lo, char* b, SysDev& d)
{
    const size_t np = (size_t)p.npad;
    for (int s = 0; s < 2; ++s) {
        double** f[9] = {&d.st[s].x, &d.st[s].y, &d.st[s].z, &d.st[s].vx, &d.st[s].vy,
                         &d.st[s].vz, &d.st[s].fx, &d.st[s].fy, &d.st[s].fz};
        for (int k = 0; k < 9; ++k) *f[k] = (double*)(b + Lo.st[s][k]);
        d.st[s].type = (uint8_t*)(b + Lo.st[s][9]);
        d.st[s].orig = (uint32_t*)(b + Lo.st[s][10]);
    }
    d.rx = (double*)(b + Lo.r);
    d.ry = d.rx + np;
    d.rz = d.ry + np;
    d.ctl = (int32_t*)(b + Lo.ctl);
    d.cell_count = (int32_t*)(b + Lo.cc);
    d.cell_start = (int32_t*)(b + Lo.cs);
    d.nat2sorted = (int32_t*)(b + Lo.n2s);
    d.tile_cell0 = (int32_t*)(b + Lo.tc);
    d.atom_cell = (int32_t*)(b + Lo.ac);
    d.scat_src = (int32_t*)(b + Lo.ss);
    d.scat_key = (uint32_t*)(b + Lo.sk);
    d.scat_x = (double*)(b + Lo.sx);
    d.nbr = (uint4*)(b + Lo.nbr);
    d.nnb = (uint16_t*)(b + Lo.nnb);
    d.partial = (double*)(b + Lo.part);
    d.pp = (const PairParam*)(b + Lo.pp);
    d.mass = (const double*)(b + Lo.m);
    d.nhc = (double*)(b + Lo.nhc);
    d.nhc_on = 0;
    d.box = (double*)(b + Lo.box);
    d.npt = (double*)(b + Lo.npt);
    d.tens = (double*)(b + Lo.tens);
    d.npt_on = 0;
    d.tri = p.tri ? 1 : 0;
    d.nblk = (p.npad + 255) / 256;
    d.thermo = nullptr;
    d.thermo_cap = 0;
    for (int k = 0; k < 3; ++k) {
        d.L[k] = p.L[k];
        d.Linv[k] = p.Linv[k];
        d.cellw[k] = p.cellw[k];
        d.nc[k] = p.nc[k];
        d.tile[k] = p.tile[k];
        d.nt[k] = p.nt[k];
    }
    d.half_skin2 = 0.25 * p.skin * p.skin;
    d.wslack = p.wslack;
    d.dt = 0.0;
    d.dt2 = 0.0;
    d.kb = p.kb;
    d.volume = p.volume;
    d.n = (int32_t)p.n;
    d.npad = p.npad;
    d.ntypes = p.ntypes;
    d.maxnb = p.maxnb;
    d.ncell = p.ncell;
    d.ntile = p.ntile;
    d.mode = p.mode;
    d.list_shift = plan_list_shift(p);
}

// The box of a plan as the device keeps it: h, h^-1 (the caller's bits) and the volume.
void pis::plan_box(const Plan& p, double* bx)
{
    for (int k = 0; k < BOX_COUNT; ++k) bx[k] = 0.0;
    for (int k = 0; k < 9; ++k) {
        bx[BOX_H + k] = p.h[k];
        bx[BOX_HINV + k] = p.hinv[k];
    }
    bx[BOX_VOL] = p.volume;
}

pis_status pis::alloc_system(SysHost& h)
{
    const Plan& p = h.plan;
    const Layout Lo = layout_of(p);
    void* mem = nullptr;
    PIS_HIP(hipMalloc(&mem, Lo.total));
    h.mem = mem;
    h.mem_bytes = Lo.total;
    SysDev& d = h.dev;
    carve_system(p, Lo, (char*)mem, d);
    // zero everything once (outside any captured sequence), then tables
    PIS_HIP(hipMemset(mem, 0, Lo.total));
    // natural -> sorted (tile-major) cell ids and the first sorted cell of every tile
    std::vector<int32_t> n2s((size_t)p.ncell), tc((size_t)p.ntile + 1);
    for (int z = 0; z < p.nc[2]; ++z)
        for (int y = 0; y < p.nc[1]; ++y)
            for (int x = 0; x < p.nc[0]; ++x)
                n2s[((size_t)z * p.nc[1] + y) * p.nc[0] + x] = sorted_cell_id(p.nc, p.tile, x, y, z);
    for (int t = 0; t < p.ntile; ++t)
        tc[t] = tile_first_cell(p.nc, p.tile, t % p.nt[0], (t / p.nt[0]) % p.nt[1], t / (p.nt[0] * p.nt[1]));
    tc[p.ntile] = p.ncell;
    PIS_HIP(hipMemcpy(d.nat2sorted, n2s.data(), n2s.size() * 4, hipMemcpyHostToDevice));
    PIS_HIP(hipMemcpy(d.tile_cell0, tc.data(), tc.size() * 4, hipMemcpyHostToDevice));
    PIS_HIP(hipMemcpy((void*)d.pp, p.pp.data(), p.pp.size() * sizeof(PairParam), hipMemcpyHostToDevice));
    PIS_HIP(hipMemcpy((void*)d.mass, p.mass.data(), p.mass.size() * 8, hipMemcpyHostToDevice));
    double bx[BOX_COUNT];
    plan_box(p, bx);
    PIS_HIP(hipMemcpy(d.box, bx, sizeof(bx), hipMemcpyHostToDevice));
    return PIS_OK;
}

pis_status pis::alloc_system_at(SysHost& h, void* mem, int thermo_records, hipStream_t st)
{
    const Plan& p = h.plan;
    const Layout Lo = layout_of(p);
    h.mem = mem;
    h.mem_bytes = system_bytes(p, thermo_records);
    h.pooled = true;
    SysDev& d = h.dev;
    carve_system(p, Lo, (char*)mem, d);
    h.thermo = (ThermoDev*)((char*)mem + Lo.total);
    h.thermo_cap = thermo_records;
    d.thermo = h.thermo;
    d.thermo_cap = thermo_records;
    PIS_HIP(hipMemsetAsync(mem, 0, Lo.total, st));
    std::vector<int32_t> n2s((size_t)p.ncell), tc((size_t)p.ntile + 1);
    for (int z = 0; z < p.nc[2]; ++z)
        for (int y = 0; y < p.nc[1]; ++y)
            for (int x = 0; x < p.nc[0]; ++x)
                n2s[((size_t)z * p.nc[1] + y) * p.nc[0] + x] = sorted_cell_id(p.nc, p.tile, x, y, z);
    for (int t = 0; t < p.ntile; ++t)
        tc[t] = tile_first_cell(p.nc, p.tile, t % p.nt[0], (t / p.nt[0]) % p.nt[1], t / (p.nt[0] * p.nt[1]));
    tc[p.ntile] = p.ncell;
    double bx[BOX_COUNT];
    plan_box(p, bx);
    PIS_HIP(hipMemcpyAsync(d.nat2sorted, n2s.data(), n2s.size() * 4, hipMemcpyHostToDevice, st));
    PIS_HIP(hipMemcpyAsync(d.tile_cell0, tc.data(), tc.size() * 4, hipMemcpyHostToDevice, st));
    PIS_HIP(hipMemcpyAsync((void*)d.pp, p.pp.data(), p.pp.size() * sizeof(PairParam), hipMemcpyHostToDevice, st));
    PIS_HIP(hipMemcpyAsync((void*)d.mass, p.mass.data(), p.mass.size() * 8, hipMemcpyHostToDevice, st));
    PIS_HIP(hipMemcpyAsync(d.box, bx, sizeof(bx), hipMemcpyHostToDevice, st));
    // the host tables go out of scope here
    PIS_HIP(hipStreamSynchronize(st));
    return PIS_OK;
}

pis_status pis::upload_sys(pis_ctx* c)
{
    std::vector<SysDev> v(c->sys.size());
    for (size_t k = 0; k < c->sys.size(); ++k) v[k] = c->sys[k].dev;
    PIS_HIP(hipMemcpyAsync(c->d_sys, v.data(), v.size() * sizeof(SysDev), hipMemcpyHostToDevice, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

void pis::free_ctx(pis_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& s : c->sys) {
        if (s.pooled) continue;
        if (s.mem) (void)hipFree(s.mem);
        if (s.thermo) (void)hipFree(s.thermo);
    }
    if (c->d_sys) (void)hipFree(c->d_sys);
    if (c->d_atom_work) (void)hipFree(c->d_atom_work);
    if (c->d_tile_work) (void)hipFree(c->d_tile_work);
    if (c->d_xfer) (void)hipFree(c->d_xfer);
    if (c->snap_stream) {
        (void)hipStreamSynchronize(c->snap_stream);
        (void)hipStreamDestroy(c->snap_stream);
    }
    if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->d_snap) (void)hipFree(c->d_snap);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    for (auto g : c->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto x : c->side) (void)hipStreamDestroy(x);
    for (auto x : c->join) (void)hipEventDestroy(x);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

pis_status pis::reserve_lanes(pis_ctx* c, int nlanes)
{
    if (!c->fork && nlanes > 1) PIS_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    while ((int)c->side.size() + 1 < nlanes) {
        hipStream_t x = nullptr;
        hipEvent_t j = nullptr;
        PIS_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        c->side.push_back(x);
        PIS_HIP(hipEventCreateWithFlags(&j, hipEventDisableTiming));
        c->join.push_back(j);
    }
    return PIS_OK;
}

// Groups every tile of every system by its LDS staging class (one launch per class) and
// uploads the work table.  Captured step graphs refer to the old table and are dropped.
pis_status pis::group_tiles(pis_ctx* c)
{
    // classes keyed by (staging capacity, kind): one build and one force launch each
    std::map<std::pair<int, int>, std::vector<WorkItem>> by_class;
    c->scap = 1024;
    for (size_t k = 0; k < c->sys.size(); ++k) {
        const Plan& p = c->sys[k].plan;
        c->scap = std::max(c->scap, p.scap);
        auto& v = by_class[{p.scap, plan_kind(p)}];
        for (int t = 0; t < p.ntile; ++t) v.push_back(WorkItem{(int32_t)k, t});
    }
    std::vector<WorkItem> tw;
    c->classes.clear();
    for (auto& kv : by_class) {
        c->classes.push_back({kv.first.first, kv.first.second, (int)tw.size(), (int)kv.second.size(), 0});
        tw.insert(tw.end(), kv.second.begin(), kv.second.end());
    }
    // Classes share up to max_lanes streams (the device serves a few hardware queues per
    // process; more streams would only time-slice): largest estimated work first, each
    // onto the least loaded stream.
    {
        std::vector<int> order(c->classes.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
        auto work = [&](int k) { return (double)c->classes[k].count * c->classes[k].scap; };
        std::sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
        const int nl = (int)std::min<size_t>((size_t)std::max(1, c->max_lanes), c->classes.size());
        std::vector<double> load(nl, 0.0);
        for (int k : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            c->classes[k].lane = l;
            load[l] += work(k);
        }
        c->nlanes = nl;
    }
    for (auto& stale : c->graph_stale) stale = true;
    PIS_TRY(reserve_lanes(c, c->nlanes));
    c->n_tile_work = (int)tw.size();
    if (c->tile_work_cap < tw.size()) {
        // grow only (a free would synchronise the device): twice the need
        if (c->d_tile_work) PIS_HIP(hipFree(c->d_tile_work));
        c->d_tile_work = nullptr;
        c->tile_work_cap = 0;
        const size_t cap = 2 * tw.size();
        PIS_HIP(hipMalloc(&c->d_tile_work, sizeof(WorkItem) * cap));
        c->tile_work_cap = cap;
    }
    if (!tw.empty()) {
        PIS_HIP(hipMemcpyAsync(c->d_tile_work, tw.data(), sizeof(WorkItem) * tw.size(), hipMemcpyHostToDevice,
                               c->stream));
        PIS_HIP(hipStreamSynchronize(c->stream));
    }
    return PIS_OK;
}

// Work tables and the device descriptor array after a membership change: the atom work
// list, the tile classes (captured graphs dropped) and every SysDev.
static pis_status relayout(pis_ctx* c)
{
    std::vector<WorkItem> aw;
    c->multi = false;
    for (size_t k = 0; k < c->sys.size(); ++k) {
        const Plan& p = c->sys[k].plan;
        c->multi |= p.ntypes > 1;
        for (int64_t a = 0; a < p.n; a += kAtomThreads) aw.push_back(WorkItem{(int32_t)k, (int32_t)a});
    }
    c->n_atom_work = (int)aw.size();
    if (c->atom_work_cap < aw.size()) {
        // grow only (a free would synchronise the device): twice the need
        if (c->d_atom_work) PIS_HIP(hipFree(c->d_atom_work));
        c->d_atom_work = nullptr;
        c->atom_work_cap = 0;
        PIS_HIP(hipMalloc(&c->d_atom_work, sizeof(WorkItem) * 2 * aw.size()));
        c->atom_work_cap = 2 * aw.size();
    }
    if (!aw.empty()) {
        PIS_HIP(hipMemcpyAsync(c->d_atom_work, aw.data(), sizeof(WorkItem) * aw.size(), hipMemcpyHostToDevice,
                               c->stream));
        PIS_HIP(hipStreamSynchronize(c->stream));
    }
    if (c->sys_cap < c->sys.size()) {
        if (c->d_sys) PIS_HIP(hipFree(c->d_sys));
        c->d_sys = nullptr;
        const size_t cap = std::max<size_t>(c->sys.size(), 2 * c->sys_cap);
        PIS_HIP(hipMalloc(&c->d_sys, sizeof(SysDev) * cap));
        c->sys_cap = cap;
    }
    PIS_TRY(group_tiles(c));
    if (!c->sys.empty()) PIS_TRY(upload_sys(c));
    return PIS_OK;
}

pis_status pis::ctx_create_empty(bool mixed, pis_ctx** out)
{
    if (!out) return fail(PIS_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PIS_ERR_NO_DEVICE, "no HIP device available");
    pis_ctx* c = new pis_ctx();
    c->device = 0;
    c->mixed = mixed;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(PIS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    }
    *out = c;
    return PIS_OK;
}

pis_status pis::ctx_add(pis_ctx* c, Plan&& plan)
{
    if (plan.mixed != c->mixed)
        return fail(PIS_ERR_INVALID_ARGUMENT, "all systems of a context need the same precision");
    SysHost h;
    h.plan = std::move(plan);
    pis_status s = alloc_system(h);
    if (s != PIS_OK) {
        if (h.mem) (void)hipFree(h.mem);
        return s;
    }
    c->sys.push_back(std::move(h));
    return relayout(c);
}

pis_status pis::ctx_remove(pis_ctx* c, std::vector<size_t> idx)
{
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    PIS_HIP(hipStreamSynchronize(c->stream));
    for (size_t j = idx.size(); j-- > 0;) {
        const size_t k = idx[j];
        if (k >= c->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
        SysHost& h = c->sys[k];
        if (!h.pooled) {
            if (h.mem) PIS_HIP(hipFree(h.mem));
            if (h.thermo) PIS_HIP(hipFree(h.thermo));
        }
        c->sys.erase(c->sys.begin() + (ptrdiff_t)k);
    }
    return relayout(c);
}

pis_status pis::ctx_take(pis_ctx* c, std::vector<size_t> idx, std::vector<SysHost>* out)
{
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    for (size_t k : idx)
        if (k >= c->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    PIS_HIP(hipStreamSynchronize(c->stream));
    for (size_t k : idx) out->push_back(std::move(c->sys[k]));
    for (size_t j = idx.size(); j-- > 0;) c->sys.erase(c->sys.begin() + (ptrdiff_t)idx[j]);
    return relayout(c);
}

pis_status pis::ctx_put(pis_ctx* c, std::vector<SysHost>&& add)
{
    if (add.empty()) return PIS_OK;
    for (auto& h : add) {
        if (h.plan.mixed != c->mixed)
            return fail(PIS_ERR_INVALID_ARGUMENT, "all systems of a context need the same precision");
        c->sys.push_back(std::move(h));
    }
    add.clear();
    return relayout(c);
}

static pis_status create_ctx(const pis_system_desc* descs, uint32_t nsys, pis_ctx** out)
{
    if (!out) return fail(PIS_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (!descs || nsys == 0) return fail(PIS_ERR_INVALID_ARGUMENT, "no systems");
    std::vector<Plan> plans(nsys);
    for (uint32_t k = 0; k < nsys; ++k) {
        pis_status s = make_plan(&descs[k], &plans[k]);
        if (s != PIS_OK) return fail(s, "system " + std::to_string(k) + ": " + g_last_error);
        if (plans[k].mixed != plans[0].mixed)
            return fail(PIS_ERR_INVALID_ARGUMENT, "all systems of a batch need the same precision");
    }
    pis_ctx* c = nullptr;
    PIS_TRY(ctx_create_empty(plans[0].mixed, &c));
    c->sys.resize(nsys);
    for (uint32_t k = 0; k < nsys; ++k) {
        c->sys[k].plan = std::move(plans[k]);
        pis_status s = alloc_system(c->sys[k]);
        if (s != PIS_OK) {
            free_ctx(c);
            return s;
        }
    }
    pis_status s = relayout(c);
    if (s != PIS_OK) {
        free_ctx(c);
        return s;
    }
    *out = c;
    return PIS_OK;
}

pis_status pis::ensure_xfer(pis_ctx* c, size_t bytes)
{
    if (c->xfer_bytes >= bytes) return PIS_OK;
    // at least twice the old size, so a stream of growing requests reallocates (and
    // synchronises the device in hipFree) only a few times
    const size_t grow = std::max(bytes, 2 * c->xfer_bytes);
    if (c->d_xfer) PIS_HIP(hipFree(c->d_xfer));
    c->d_xfer = nullptr;
    c->xfer_bytes = 0;
    PIS_HIP(hipMalloc(&c->d_xfer, grow));
    c->xfer_bytes = grow;
    return PIS_OK;
}

pis_status pis::read_ctl(pis_ctx* c, uint32_t k, int32_t* ctl)
{
    PIS_HIP(hipMemcpyAsync(ctl, c->sys[k].dev.ctl, CTL_COUNT * 4, hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

pis_status pis::write_ctl(pis_ctx* c, uint32_t k, const int32_t* ctl)
{
    PIS_HIP(hipMemcpyAsync(c->sys[k].dev.ctl, ctl, CTL_COUNT * 4, hipMemcpyHostToDevice, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

// The per-class part of a step: list build (if `build`) then forces (if `force >= 0`,
// with that integrate flag).  Classes touch disjoint tiles and atoms, so the classes of
// lane l > 0 run on side[l-1], forked after everything already on `stream` and joined
// before anything enqueued later; captured into a graph, the fork becomes parallel
// branches.
pis_status pis::enqueue_classes(pis_ctx* c, bool build, int force)
{
    const int nl = c->nlanes;
    if (nl > 1) {
        PIS_HIP(hipEventRecord(c->fork, c->stream));
        for (int l = 1; l < nl; ++l) PIS_HIP(hipStreamWaitEvent(c->side[l - 1], c->fork, 0));
    }
    for (const auto& q : c->classes) {
        hipStream_t st = q.lane == 0 ? c->stream : c->side[q.lane - 1];
        if (build) launch_build_list(st, c->d_sys, c->d_tile_work + q.first, q.count, q.scap, q.kind);
        if (force >= 0) launch_force(st, c->d_sys, c->d_tile_work + q.first, q.count, q.scap, c->mixed, force, q.kind);
    }
    for (int l = 1; l < nl; ++l) {
        PIS_HIP(hipEventRecord(c->join[l - 1], c->side[l - 1]));
        PIS_HIP(hipStreamWaitEvent(c->stream, c->join[l - 1], 0));
    }
    return PIS_OK;
}

pis_status pis::enqueue_rebuild(pis_ctx* c)
{
    launch_rebuild_sort(c->stream, c->d_sys, (int)c->sys.size(), c->d_atom_work, c->n_atom_work);
    return enqueue_classes(c, true, -1);
}

pis_status pis::run_rebuild(pis_ctx* c)
{
    PIS_TRY(enqueue_rebuild(c));
    launch_flip(c->stream, c->d_sys, (int)c->sys.size());
    // the other buffer set still has the previous order's types and original indices, and
    // the force kernel rewrites them there only on rebuild steps
    launch_sync_labels(c->stream, c->d_sys, c->d_atom_work, c->n_atom_work);
    PIS_HIP(hipGetLastError());
    return PIS_OK;
}

pis_status pis::check_flags(pis_ctx* c, int only)
{
    for (uint32_t k = 0; k < c->sys.size(); ++k) {
        if (only >= 0 && (int)k != only) continue;
        int32_t ctl[CTL_COUNT];
        PIS_TRY(read_ctl(c, k, ctl));
        if (ctl[CTL_OVF_STAGE])
            return fail(PIS_ERR_CAPACITY, "system " + std::to_string(k) + ": LDS staging capacity exceeded");
        if (ctl[CTL_BOX])
            return fail(PIS_ERR_CAPACITY, "system " + std::to_string(k) +
                                              ": the box deformed below the planned cell grid (rcut + skin)");
        if (ctl[CTL_OVF_NB])
            return fail(PIS_ERR_CAPACITY, "system " + std::to_string(k) + ": neighbour capacity " +
                                              std::to_string(c->sys[k].plan.maxnb) + " exceeded (" +
                                              std::to_string(ctl[CTL_MAXNB]) + " needed)");
        if (ctl[CTL_NAN])
            return fail(PIS_ERR_NAN_ENERGY, "system " + std::to_string(k) + ": non-finite energy or position");
    }
    return PIS_OK;
}

// A configuration much denser than its average in places (a lattice filled only in part,
// a cluster) can fill a tile's halo beyond the largest staging class: its tiles are planned
// again for the measured peak, and the system moves into memory laid out for them.
static pis_status replan_tiles_for_halo(pis_ctx* c, uint32_t k, int staged)
{
    SysHost& h = c->sys[k];
    if (h.pooled)  // its memory belongs to a caller's pool, sized for the first plan
        return fail(PIS_ERR_CAPACITY, "system " + std::to_string(k) + ": " + std::to_string(staged) +
                                          " atoms in one tile's halo exceed the LDS staging limit");
    SysHost nh;
    nh.plan = h.plan;
    Plan& p = nh.plan;
    const double cellvol = p.cellw[0] * p.cellw[1] * p.cellw[2];
    const double planned = p.density * p.halo_scale * cellvol * (p.tile[0] + 2) * (p.tile[1] + 2) * (p.tile[2] + 2);
    p.halo_scale *= 1.05 * staged / std::max(planned, 1.0);
    PIS_TRY(plan_grid(&p, p.grid_margin));
    pis_status st = alloc_system(nh);
    if (st != PIS_OK) {
        if (nh.mem) (void)hipFree(nh.mem);
        return st;
    }
    nh.dev.dt = h.dev.dt;
    nh.dev.dt2 = h.dev.dt2;
    nh.thermo = h.thermo;
    nh.thermo_cap = h.thermo_cap;
    nh.dev.thermo = h.thermo;
    nh.dev.thermo_cap = h.thermo_cap;
    // a thermostat or barostat set up before the state keeps its device state
    PIS_HIP(hipMemcpyAsync(nh.dev.nhc, h.dev.nhc, NHC_COUNT * 8, hipMemcpyDeviceToDevice, c->stream));
    PIS_HIP(hipMemcpyAsync(nh.dev.npt, h.dev.npt, NPT_COUNT * 8, hipMemcpyDeviceToDevice, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(h.mem);
    h = std::move(nh);
    return relayout(c);
}

pis_status pis::sys_set_state(pis_ctx* c, uint32_t k, const double* pos, const double* vel,
                                const double* force)
{
    if (!c) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (k >= c->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    if (!pos) return fail(PIS_ERR_INVALID_ARGUMENT, "positions are required");
    SysHost& h = c->sys[k];
    const size_t n = (size_t)h.plan.n;
    for (size_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i])) return fail(PIS_ERR_INVALID_ARGUMENT, "non-finite position");
    PIS_TRY(ensure_xfer(c, 3 * 3 * n * 8 + n + 256));
    double* dp = (double*)c->d_xfer;
    double* dv = dp + 3 * n;
    double* df = dv + 3 * n;
    uint8_t* dt = (uint8_t*)(df + 3 * n);
    PIS_HIP(hipMemcpyAsync(dp, pos, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
    if (vel) PIS_HIP(hipMemcpyAsync(dv, vel, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
    if (force) PIS_HIP(hipMemcpyAsync(df, force, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
    PIS_HIP(hipMemcpyAsync(dt, h.plan.type0.data(), n, hipMemcpyHostToDevice, c->stream));
    int32_t ctl[CTL_COUNT] = {0};
    ctl[CTL_CUR] = 0;
    ctl[CTL_REBUILD] = 1;
    PIS_TRY(write_ctl(c, k, ctl));
    launch_load_aos(c->stream, c->d_sys, (int)k, (int)n, dp, vel ? dv : nullptr, force ? df : nullptr, dt);
    PIS_TRY(run_rebuild(c));
    PIS_HIP(hipStreamSynchronize(c->stream));
    // The planned staging capacity assumes a uniform density.  A configuration with denser
    // spots (a lattice whose planes sit on cell faces, a cluster) that fills a tile's halo
    // to within 2 % + 16 atoms of the class gets a larger class, sized from the halo count
    // just measured with the headroom the plan leaves; the rebuild is then repeated.  A
    // state's order after sorting depends only on its positions, so sorting again is
    // harmless.
    PIS_TRY(read_ctl(c, k, ctl));
    Plan& p = h.plan;
    const int staged = ctl[CTL_MAXSTAGE];
    if (p.mode != MODE_WHOLE && staged * 1.02 + 16 > p.scap) {
        const int grown = scap_for(staged * 1.1 + 48);
        if (grown < 0) {
            if (p.halo_scale > 1.0)  // planned for a measured peak already: give up
                return fail(PIS_ERR_CAPACITY, "system " + std::to_string(k) + ": " + std::to_string(staged) +
                                                  " atoms in one tile's halo exceed the LDS staging limit");
            PIS_TRY(replan_tiles_for_halo(c, k, staged));
            return sys_set_state(c, k, pos, vel, force);
        }
        p.scap = grown;
        PIS_TRY(group_tiles(c));
        int32_t again[CTL_COUNT] = {0};
        again[CTL_CUR] = ctl[CTL_CUR];
        again[CTL_REBUILD] = 1;
        PIS_TRY(write_ctl(c, k, again));
        PIS_TRY(run_rebuild(c));
        PIS_HIP(hipStreamSynchronize(c->stream));
    }
    return check_flags(c, (int)k);
}

pis_status pis::sys_get_state(pis_ctx* c, uint32_t k, double* pos, double* vel, double* force)
{
    if (!c) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (k >= c->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    const size_t n = (size_t)c->sys[k].plan.n;
    PIS_TRY(ensure_xfer(c, 3 * 3 * n * 8 + n + 256));
    double* dp = (double*)c->d_xfer;
    double* dv = dp + 3 * n;
    double* df = dv + 3 * n;
    launch_store_aos(c->stream, c->d_sys, (int)k, (int)n, dp, dv, df);
    PIS_HIP(hipGetLastError());
    if (pos) PIS_HIP(hipMemcpyAsync(pos, dp, 3 * n * 8, hipMemcpyDeviceToHost, c->stream));
    if (vel) PIS_HIP(hipMemcpyAsync(vel, dv, 3 * n * 8, hipMemcpyDeviceToHost, c->stream));
    if (force) PIS_HIP(hipMemcpyAsync(force, df, 3 * n * 8, hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

pis_status pis::ensure_thermo(pis_ctx* c, int cap)
{
    bool changed = false;
    for (auto& h : c->sys) {
        if (h.thermo_cap >= cap) continue;
        if (h.thermo) PIS_HIP(hipFree(h.thermo));
        h.thermo = nullptr;
        PIS_HIP(hipMalloc(&h.thermo, sizeof(ThermoDev) * (size_t)cap));
        h.thermo_cap = cap;
        h.dev.thermo = h.thermo;
        h.dev.thermo_cap = cap;
        changed = true;
    }
    if (changed) PIS_TRY(upload_sys(c));
    return PIS_OK;
}

// Force launches bracketed by HIP events on the context's stream while timing is on; the
// events come from a pool that is reused between calls.
static pis_status launch_force_timed(pis_ctx* c, int integrate)
{
    if (c->timing) {
        while (c->ev.size() < c->ev_used + 2) {
            hipEvent_t e = nullptr;
            PIS_HIP(hipEventCreate(&e));
            c->ev.push_back(e);
        }
        PIS_HIP(hipEventRecord(c->ev[c->ev_used], c->stream));
    }
    PIS_TRY(enqueue_classes(c, false, integrate));
    if (c->timing) {
        PIS_HIP(hipEventRecord(c->ev[c->ev_used + 1], c->stream));
        c->ev_used += 2;
    }
    c->force_launches++;
    return PIS_OK;
}

static pis_status collect_timing(pis_ctx* c)
{
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.0f;
        PIS_HIP(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
        c->force_ms += ms;
        c->timed_launches++;
    }
    c->ev_used = 0;
    return PIS_OK;
}

static pis_status read_thermo(pis_ctx* c, uint64_t nrec, pis_thermo* out)
{
    const size_t ns = c->sys.size();
    std::vector<ThermoDev> buf(nrec);
    for (size_t k = 0; k < ns; ++k) {
        PIS_HIP(hipMemcpyAsync(buf.data(), c->sys[k].thermo, sizeof(ThermoDev) * nrec,
                               hipMemcpyDeviceToHost, c->stream));
        PIS_HIP(hipStreamSynchronize(c->stream));
        for (uint64_t t = 0; t < nrec; ++t) {
            pis_thermo& o = out[t * ns + k];
            o.pe = buf[t].pe;
            o.ke = buf[t].ke;
            o.etot = buf[t].etot;
            o.temperature = buf[t].temperature;
            o.pressure = buf[t].pressure;
            o.virial_abs = buf[t].virial_abs;
            o.virial_pair = buf[t].virial_pair;
        }
    }
    return PIS_OK;
}

static pis_status batch_compute(pis_ctx* c, pis_thermo* out /* [nsys] or null */)
{
    if (!c) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    PIS_TRY(ensure_thermo(c, 1));
    // a pending rebuild (the last step's skin check fired, or the staging changed) first:
    // the list must hold every pair within the cutoff at these positions
    PIS_TRY(run_rebuild(c));
    launch_set_ctl(c->stream, c->d_sys, (int)c->sys.size(), CTL_STEP, 0);
    PIS_TRY(launch_force_timed(c, 0));
    launch_reduce(c->stream, c->d_sys, (int)c->sys.size(), 1, 1, 0);
    PIS_HIP(hipGetLastError());
    PIS_HIP(hipStreamSynchronize(c->stream));
    PIS_TRY(collect_timing(c));
    PIS_TRY(check_flags(c));
    std::vector<pis_thermo> t(c->sys.size());
    PIS_TRY(read_thermo(c, 1, t.data()));
    if (out)
        for (size_t k = 0; k < t.size(); ++k) out[k] = t[k];
    return PIS_OK;
}

// One step's launches; `first`: the position-update kernel runs (a call's first step,
// and every NVT step, whose velocity scaling needs the chain update in between).
// One step.  nvt: the Nose-Hoover chain wraps it; the chain's start-of-step update runs
// as its own launch on a call's first step and on NPT steps (whose barostat rescaled the
// velocities since), otherwise at the end of the previous step (k_reduce).
static pis_status enqueue_one(pis_ctx* c, bool first, bool record, bool nvt, bool timed, bool npt = false)
{
    const int ns = (int)c->sys.size();
    if (nvt && (first || npt)) launch_nhc_pre(c->stream, c->d_sys, ns);
    if (first || nvt) launch_integrate_pos(c->stream, c->d_sys, c->d_atom_work, c->n_atom_work);
    launch_rebuild_sort(c->stream, c->d_sys, ns, c->d_atom_work, c->n_atom_work);
    if (timed) {
        // timed: every list build first, then the force launches between events
        PIS_TRY(enqueue_classes(c, true, -1));
        PIS_TRY(launch_force_timed(c, 1));
    } else {
        PIS_TRY(enqueue_classes(c, true, 1));
        c->force_launches++;
    }
    launch_reduce(c->stream, c->d_sys, ns, 1, record ? 1 : 0, 1);
    return PIS_OK;
}

pis_status pis::enqueue_steps(pis_ctx* c, uint64_t nsteps, bool record, bool nvt)
{
    const int ns = (int)c->sys.size();
    launch_set_ctl(c->stream, c->d_sys, ns, CTL_REMAIN, (int)std::min<uint64_t>(nsteps, (uint64_t)INT32_MAX));
    if (c->use_graph && !c->timing) {
        for (int steady = 0; steady < 2; ++steady) {
            const int gi = (record ? 1 : 0) + (nvt ? 2 : 0) + 4 * steady;
            if ((c->graph[gi] && !c->graph_stale[gi]) || (steady && nsteps < 2)) continue;
            const auto tg = std::chrono::steady_clock::now();
            hipGraph_t g = nullptr;
            PIS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const uint64_t launches_before = c->force_launches;
            pis_status cs = enqueue_one(c, !steady, record, nvt, false);
            c->force_launches = launches_before;
            hipError_t ce = hipStreamEndCapture(c->stream, &g);
            if (cs != PIS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return cs;
            }
            if (ce != hipSuccess) return fail(PIS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
            c->graph_stale[gi] = false;
            if (c->graph[gi]) PIS_HIP(hipGraphExecDestroy(c->graph[gi]));
            c->graph[gi] = nullptr;
            hipError_t e = hipGraphInstantiate(&c->graph[gi], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(PIS_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            c->graph_instantiations++;
            c->graph_build_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tg).count();
        }
        const int g0 = (record ? 1 : 0) + (nvt ? 2 : 0);
        for (uint64_t t = 0; t < nsteps; ++t) PIS_HIP(hipGraphLaunch(c->graph[g0 + (t ? 4 : 0)], c->stream));
        c->force_launches += nsteps;
    } else {
        for (uint64_t t = 0; t < nsteps; ++t) {
            PIS_TRY(enqueue_one(c, t == 0, record, nvt, c->timing));
            if ((t & 255) == 0) PIS_HIP(hipGetLastError());
        }
    }
    return PIS_OK;
}

// NPT on a single-system context: the barostat would have made the box thinner than its
// cell grid (CTL_BOX), and the system is frozen at the start of that step with its state,
// chain and barostat as the previous step left them.  The reference bins cells afresh in
// every force call (lennard_jones.rs:188-190, neighbour_list.rs:45-57); here the grid is
// planned again for the current box with a 5 % margin, the system moves into memory laid
// out for it, and the list is rebuilt.
static pis_status replan_for_box(pis_ctx* c, uint32_t k)
{
    SysHost& h = c->sys[k];
    if (h.pooled) return fail(PIS_ERR_CAPACITY, "the box deformed below the planned cell grid (rcut + skin)");
    const size_t n = (size_t)h.plan.n;
    double box[BOX_COUNT], npt[NPT_COUNT], chain[NHC_COUNT];
    PIS_HIP(hipMemcpyAsync(box, h.dev.box, sizeof(box), hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipMemcpyAsync(npt, h.dev.npt, sizeof(npt), hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipMemcpyAsync(chain, h.dev.nhc, sizeof(chain), hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> pos(3 * n), vel(3 * n), frc(3 * n);
    PIS_TRY(sys_get_state(c, k, pos.data(), vel.data(), frc.data()));
    SysHost nh;
    nh.plan = h.plan;
    Plan& p = nh.plan;
    for (int q = 0; q < 9; ++q) {
        p.h[q] = box[BOX_H + q];
        p.hinv[q] = box[BOX_HINV + q];
    }
    p.tri = true;
    PIS_TRY(plan_general_box(&p));
    double ext = 0.0;
    for (int r = 0; r < 3; ++r) ext = std::max(ext, std::fabs(p.h[r]) + std::fabs(p.h[3 + r]) + std::fabs(p.h[6 + r]));
    p.qmax = std::max(p.qmax, 1.25 * ext);  // the box may grow again (see npt_init)
    PIS_TRY(plan_grid(&p, 0.05));
    pis_status st = alloc_system(nh);
    if (st != PIS_OK) {
        if (nh.mem) (void)hipFree(nh.mem);
        return st;
    }
    // the running system: step settings, the thermo ring, the box bits, chain and barostat
    // (whose list reference is void until the list is rebuilt in the new grid)
    nh.dev.dt = h.dev.dt;
    nh.dev.dt2 = h.dev.dt2;
    nh.dev.nhc_on = h.dev.nhc_on;
    nh.dev.npt_on = h.dev.npt_on;
    nh.thermo = h.thermo;
    nh.thermo_cap = h.thermo_cap;
    nh.dev.thermo = h.thermo;
    nh.dev.thermo_cap = h.thermo_cap;
    npt[NPT_LIST_OK] = 0.0;
    PIS_HIP(hipMemcpy(nh.dev.box, box, sizeof(box), hipMemcpyHostToDevice));
    PIS_HIP(hipMemcpy(nh.dev.npt, npt, sizeof(npt), hipMemcpyHostToDevice));
    PIS_HIP(hipMemcpy(nh.dev.nhc, chain, sizeof(chain), hipMemcpyHostToDevice));
    (void)hipFree(h.mem);
    h = std::move(nh);
    PIS_TRY(relayout(c));
    return sys_set_state(c, k, pos.data(), vel.data(), frc.data());
}

// nsteps of verlet_step_nve (nvt = false) or verlet_step_nvt_nhc (nvt = true, chain set
// by pis_nvt_init) for every system, state resident on the device.
static pis_status batch_run(pis_ctx* c, const double* dts, uint64_t nsteps, pis_thermo* out, bool nvt = false,
                            uint64_t ramp_step0 = 0, uint64_t ramp_total = 0, bool npt = false)
{
    if (!c) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (!dts) return fail(PIS_ERR_INVALID_ARGUMENT, "null timestep array");
    if (nsteps == 0) return PIS_OK;
    if (nsteps > (uint64_t)1 << 30) return fail(PIS_ERR_INVALID_ARGUMENT, "too many steps");
    if (nvt && !c->nhc_ready) return fail(PIS_ERR_INVALID_ARGUMENT, "NVT needs pis_nvt_init first");
    if (npt && !c->npt_ready) return fail(PIS_ERR_INVALID_ARGUMENT, "NPT needs pis_npt_init first");
    for (size_t k = 0; k < c->sys.size(); ++k) {
        if (!(dts[k] > 0.0) || !std::isfinite(dts[k])) return fail(PIS_ERR_INVALID_ARGUMENT, "dt must be > 0");
        c->sys[k].dev.dt = dts[k];
        c->sys[k].dev.dt2 = dts[k] * dts[k];  // dt.powi(2)
        c->sys[k].dev.nhc_on = nvt ? 1 : 0;
        c->sys[k].dev.npt_on = npt ? 1 : 0;
    }
    // The thermo ring holds at least kThermoFloor records, so calls of different lengths
    // reuse one allocation (a reallocation would synchronise the device inside the call);
    // a call that records more than kThermoChunk steps runs in chunks of that many, each
    // downloaded before the next starts.
    // NPT runs in chunks of kNptChunk steps: a system whose box would become thinner than
    // its cell grid freezes on the device, and the host re-plans it between chunks.
    constexpr uint64_t kThermoFloor = 1024, kThermoChunk = 8192, kNptChunk = 64;
    uint64_t chunk = out ? std::min(nsteps, kThermoChunk) : nsteps;
    if (npt) chunk = std::min(chunk, kNptChunk);
    PIS_TRY(ensure_thermo(c, out ? (int)std::max(chunk, kThermoFloor) : 1));
    PIS_TRY(upload_sys(c));
    const int ns = (int)c->sys.size();
    if (nvt) {
        if (ramp_total == 0) ramp_total = nsteps;
        if (ramp_step0 + nsteps > ramp_total)
            return fail(PIS_ERR_INVALID_ARGUMENT, "the steps run past the end of the temperature ramp");
        launch_nhc_run(c->stream, c->d_sys, ns, ramp_step0, ramp_total);
    }
    bool resumed = false;  // the chunk starts on a re-planned system
    for (uint64_t done = 0; done < nsteps;) {
        const uint64_t k = std::min(chunk, nsteps - done);
        launch_set_ctl(c->stream, c->d_sys, ns, CTL_LEFT, -1);
        launch_set_ctl(c->stream, c->d_sys, ns, CTL_STEP, 0);
        if (npt) {
            // verlet_step_npt_mtk (potential.rs:112-135): barostat and box update, the NVT
            // step on the scaled state, the second scaling and momentum update; direct
            // launches, step by step.  The tensor sums a step starts from are those the
            // previous step ended with (nothing moves in between), so only a call's first
            // step, and the first after a re-plan, computes them.
            const int rec = out != nullptr ? 1 : 0;
            for (uint64_t t = 0; t < k; ++t) {
                if (done + t == 0 || (resumed && t == 0))
                    launch_npt(c->stream, c->d_sys, ns, c->d_atom_work, c->n_atom_work, 0, rec);
                launch_npt(c->stream, c->d_sys, ns, c->d_atom_work, c->n_atom_work, 1, rec);
                PIS_TRY(enqueue_one(c, t == 0, out != nullptr, true, false, true));
                launch_npt(c->stream, c->d_sys, ns, c->d_atom_work, c->n_atom_work, 2, rec);
                if ((t & 63) == 0) PIS_HIP(hipGetLastError());
            }
        } else {
            PIS_TRY(enqueue_steps(c, k, out != nullptr, nvt));
        }
        PIS_HIP(hipGetLastError());
        PIS_HIP(hipStreamSynchronize(c->stream));
        PIS_TRY(collect_timing(c));
        if (npt) {
            int32_t ctl[CTL_COUNT];
            PIS_TRY(read_ctl(c, 0, ctl));
            if (ns == 1 && ctl[CTL_BOX]) {
                // the steps before the freeze are done; the one that froze is run again
                double chain[NHC_COUNT];
                PIS_HIP(hipMemcpyAsync(chain, c->sys[0].dev.nhc, sizeof(chain), hipMemcpyDeviceToHost, c->stream));
                PIS_HIP(hipStreamSynchronize(c->stream));
                const uint64_t completed = (uint64_t)chain[NHC_STEP] - (ramp_step0 + done);
                if (completed == 0 && resumed)
                    return fail(PIS_ERR_CAPACITY, "the box shrinks faster than its cell grid can be re-planned");
                if (out && completed) PIS_TRY(read_thermo(c, completed, out + done * (uint64_t)ns));
                c->steps += completed;
                done += completed;
                PIS_TRY(replan_for_box(c, 0));
                resumed = true;
                continue;
            }
        }
        c->steps += k;
        PIS_TRY(check_flags(c));
        if (out) PIS_TRY(read_thermo(c, k, out + done * (uint64_t)ns));
        done += k;
        resumed = false;
    }
    return PIS_OK;
}

// NHThermostatChain::new_from_args (nvt.rs:22-57, 95-108) for every system of the
// context: xi = eta = G = 0, q_i = kB T_start tau^2 / 10^i, target = T_start.
static pis_status nvt_init(pis_ctx* c, const pis_nhc_params* p)
{
    if (!c || !p) return fail(PIS_ERR_INVALID_ARGUMENT, "null argument");
    if (p->chain < 1 || p->chain > NHC_MAX_CHAIN)
        return fail(PIS_ERR_INVALID_ARGUMENT, "chain length must be 1.." + std::to_string(NHC_MAX_CHAIN));
    if (!(p->tau > 0.0) || !std::isfinite(p->t_start) || !std::isfinite(p->t_end) || !(p->t_start > 0.0))
        return fail(PIS_ERR_INVALID_ARGUMENT, "need tau > 0 and a positive start temperature");
    for (auto& h : c->sys) {
        double st[NHC_COUNT] = {0.0};
        const double kb = h.plan.kb;
        const double q_value = kb * p->t_start * (p->tau * p->tau);
        double p10 = 1.0;
        for (int i = 0; i < p->chain; ++i) {
            st[NHC_Q + i] = q_value / p10;
            p10 *= 10.0;
        }
        st[NHC_T] = p->t_start;
        st[NHC_T0] = p->t_start;
        st[NHC_T1] = p->t_end;
        st[NHC_KB] = kb;
        st[NHC_SCALE] = 1.0;
        st[NHC_TOTAL] = 1.0;
        st[NHC_CHAIN] = (double)p->chain;
        PIS_HIP(hipMemcpyAsync(h.dev.nhc, st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    }
    PIS_HIP(hipStreamSynchronize(c->stream));
    c->nhc_ready = true;
    return PIS_OK;
}

// MTKBarostat::new via new_from_args (npt.rs:24-42, 66-87) for every system, together with
// the chain: target = p0 I, momentum 0, w = 3N KB T_start tau^2 with the constant
// KB = 0.0083144621 (npt.rs:32; not the system's kb).
static pis_status npt_init(pis_ctx* c, const pis_nhc_params* chain, const pis_mtk_params* mtk)
{
    if (!c || !chain || !mtk) return fail(PIS_ERR_INVALID_ARGUMENT, "null argument");
    if (!(mtk->tau > 0.0) || !std::isfinite(mtk->tau) || !std::isfinite(mtk->p_start))
        return fail(PIS_ERR_INVALID_ARGUMENT, "need a finite target pressure and tau > 0");
    if (mtk->reserved[0] != 0.0 || mtk->reserved[1] != 0.0)
        return fail(PIS_ERR_INVALID_ARGUMENT, "pis_mtk_params.reserved must be 0");
    PIS_TRY(nvt_init(c, chain));
    bool relist = false;
    for (auto& h : c->sys) {
        // the box becomes general with the first step: raw staging with the general-box
        // image, bands for positions up to 1.25x the current extent (the box may grow)
        Plan& p = h.plan;
        double ext = 0.0;
        for (int r = 0; r < 3; ++r) ext = std::max(ext, std::fabs(p.h[r]) + std::fabs(p.h[3 + r]) + std::fabs(p.h[6 + r]));
        p.tri = true;
        p.qmax = std::max(p.qmax, 1.25 * ext);
        for (int k = 0; k < 3; ++k)  // tile-relative staging of a general box (see make_plan)
            if (p.mode == MODE_TILE_REL && p.nc[k] - p.tile[k] < 2) p.mode = MODE_RAW_MI;
        set_bands(&p);
        h.dev.tri = 1;
        h.dev.mode = p.mode;
        if (h.dev.list_shift != plan_list_shift(p)) relist = true;
        h.dev.list_shift = plan_list_shift(p);
        h.dev.wslack = p.wslack;
        PIS_HIP(hipMemcpyAsync((void*)h.dev.pp, p.pp.data(), p.pp.size() * sizeof(PairParam), hipMemcpyHostToDevice,
                               c->stream));
        double st[NPT_COUNT] = {0.0};
        for (int k = 0; k < 3; ++k) st[NPT_TARGET + 4 * k] = mtk->p_start;
        st[NPT_W] = ((double)(3 * h.plan.n)) * 0.0083144621 * chain->t_start * (mtk->tau * mtk->tau);
        st[NPT_RCL] = h.plan.rcl * (1.0 + 1e-12);
        st[NPT_RCMAX] = h.plan.rcmax;  // NPT_LIST_OK = 0: the first step rebuilds
        PIS_HIP(hipMemcpyAsync(h.dev.npt, st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    }
    PIS_HIP(hipStreamSynchronize(c->stream));
    // the systems' kinds changed (general boxes): regroup the tiles, which also drops the
    // captured step graphs holding the orthorhombic kernel instantiations
    PIS_TRY(group_tiles(c));
    PIS_TRY(upload_sys(c));
    if (relist) {
        // the lists hold slot offsets for the previous staging: the next step or force
        // call rebuilds them
        launch_set_ctl(c->stream, c->d_sys, (int)c->sys.size(), CTL_REBUILD, 1);
        PIS_HIP(hipStreamSynchronize(c->stream));
    }
    c->npt_ready = true;
    return PIS_OK;
}

// ---------------------------------------------------------------------------------
// C ABI

extern "C" {

int pis_abi_version(void) { return PIS_ABI_VERSION; }

const char* pis_last_error(void) { return g_last_error.c_str(); }

const char* pis_status_string(pis_status s)
{
    switch (s) {
    case PIS_OK: return "ok";
    case PIS_ERR_NO_ATOMS: return "no atoms defined";
    case PIS_ERR_MISSING_POTENTIAL: return "missing potential";
    case PIS_ERR_NAN_ENERGY: return "NaN energy";
    case PIS_ERR_INVALID_ATOM_TYPE: return "invalid atom type";
    case PIS_ERR_INVALID_ARGUMENT: return "invalid argument";
    case PIS_ERR_UNSUPPORTED: return "unsupported";
    case PIS_ERR_HIP: return "HIP error";
    case PIS_ERR_OUT_OF_MEMORY: return "out of device memory";
    case PIS_ERR_NO_DEVICE: return "no device";
    case PIS_ERR_CAPACITY: return "capacity exceeded";
    case PIS_ERR_NOT_FOUND: return "not found";
    case PIS_ERR_TIMEOUT: return "timed out";
    case PIS_ERR_IO: return "trajectory write failed";
    }
    return "unknown status";
}

pis_status pis_device_count(int32_t* count)
{
    if (!count) return fail(PIS_ERR_INVALID_ARGUMENT, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return PIS_OK;
}

pis_status pis_plan_system(const pis_system_desc* desc, pis_plan_info* out)
{
    if (!out) return fail(PIS_ERR_INVALID_ARGUMENT, "null output");
    Plan p;
    PIS_TRY(make_plan(desc, &p));
    for (int k = 0; k < 3; ++k) {
        out->ncell[k] = p.nc[k];
        out->tile[k] = p.tile[k];
        out->ntile[k] = p.nt[k];
    }
    out->mode = p.mode;
    out->neighbor_capacity = p.maxnb;
    out->stage_capacity = p.scap;
    out->r2_list_max = r2max_for(p.rcl);
    out->density = p.density;
    return PIS_OK;
}

pis_status pis_set_tile_override(const int32_t* tile)
{
    for (int d = 0; d < 3; ++d) g_tile_override[d] = tile ? std::max(0, tile[d]) : 0;
    return PIS_OK;
}

pis_status pis_ctx_create(const pis_system_desc* desc, pis_ctx** out) { return create_ctx(desc, 1, out); }

pis_status pis_batch_create(const pis_system_desc* descs, uint32_t n_systems, pis_ctx** out)
{
    return create_ctx(descs, n_systems, out);
}

uint32_t pis_batch_size(const pis_ctx* ctx) { return ctx ? (uint32_t)ctx->sys.size() : 0u; }

void pis_ctx_destroy(pis_ctx* ctx) { free_ctx(ctx); }

pis_status pis_set_state(pis_ctx* ctx, const double* pos, const double* vel, const double* force)
{
    return sys_set_state(ctx, 0, pos, vel, force);
}

pis_status pis_get_state(pis_ctx* ctx, double* pos, double* vel, double* force)
{
    return sys_get_state(ctx, 0, pos, vel, force);
}

pis_status pis_sys_set_state(pis_ctx* ctx, uint32_t sys, const double* pos, const double* vel,
                             const double* force)
{
    return sys_set_state(ctx, sys, pos, vel, force);
}

pis_status pis_sys_get_state(pis_ctx* ctx, uint32_t sys, double* pos, double* vel, double* force)
{
    return sys_get_state(ctx, sys, pos, vel, force);
}

pis_status pis_host_alloc(size_t bytes, void** out)
{
    if (!out) return fail(PIS_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    PIS_HIP(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return PIS_OK;
}

pis_status pis_host_free(void* p)
{
    if (p) PIS_HIP(hipHostFree(p));
    return PIS_OK;
}

pis_status pis_snapshot_wait(pis_ctx* ctx)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (!ctx->snap_pending) return PIS_OK;
    ctx->snap_pending = false;
    PIS_HIP(hipEventSynchronize(ctx->snap_done));
    return PIS_OK;
}

pis_status pis_snapshot_positions(pis_ctx* ctx, uint32_t sys, double* host_pos)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (sys >= ctx->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    if (!host_pos) return fail(PIS_ERR_INVALID_ARGUMENT, "null destination");
    PIS_TRY(pis_snapshot_wait(ctx));  // the device buffer is free again
    if (!ctx->snap_stream) {
        PIS_HIP(hipStreamCreateWithFlags(&ctx->snap_stream, hipStreamNonBlocking));
        PIS_HIP(hipEventCreateWithFlags(&ctx->snap_ready, hipEventDisableTiming));
        PIS_HIP(hipEventCreateWithFlags(&ctx->snap_done, hipEventDisableTiming));
    }
    const size_t n = (size_t)ctx->sys[sys].plan.n, bytes = 3 * n * sizeof(double);
    if (ctx->snap_bytes < bytes) {
        if (ctx->d_snap) PIS_HIP(hipFree(ctx->d_snap));
        ctx->d_snap = nullptr;
        ctx->snap_bytes = 0;
        PIS_HIP(hipMalloc(&ctx->d_snap, bytes));
        ctx->snap_bytes = bytes;
    }
    // behind every step already queued on the context's stream
    launch_store_aos(ctx->stream, ctx->d_sys, (int)sys, (int)n, ctx->d_snap, nullptr, nullptr);
    PIS_HIP(hipGetLastError());
    PIS_HIP(hipEventRecord(ctx->snap_ready, ctx->stream));
    PIS_HIP(hipStreamWaitEvent(ctx->snap_stream, ctx->snap_ready, 0));
    PIS_HIP(hipMemcpyAsync(host_pos, ctx->d_snap, bytes, hipMemcpyDeviceToHost, ctx->snap_stream));
    PIS_HIP(hipEventRecord(ctx->snap_done, ctx->snap_stream));
    ctx->snap_pending = true;
    return PIS_OK;
}

pis_status pis_compute_potential(pis_ctx* ctx, double* forces_inout, double* pe)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->sys.size() != 1) return fail(PIS_ERR_INVALID_ARGUMENT, "use pis_batch_compute_potential");
    pis_thermo t;
    PIS_TRY(batch_compute(ctx, &t));
    if (pe) *pe = t.pe;
    if (forces_inout) {
        const size_t n = (size_t)ctx->sys[0].plan.n;
        std::vector<double> f(3 * n);
        PIS_TRY(sys_get_state(ctx, 0, nullptr, nullptr, f.data()));
        for (size_t i = 0; i < 3 * n; ++i) forces_inout[i] += f[i];
    }
    return PIS_OK;
}

pis_status pis_batch_compute_potential(pis_ctx* ctx, double* pe_out)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    std::vector<pis_thermo> t(ctx->sys.size());
    PIS_TRY(batch_compute(ctx, t.data()));
    if (pe_out)
        for (size_t k = 0; k < t.size(); ++k) pe_out[k] = t[k].pe;
    return PIS_OK;
}

pis_status pis_verlet_step_nve(pis_ctx* ctx, double dt, uint64_t nsteps, pis_thermo* per_step_out)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->sys.size() != 1) return fail(PIS_ERR_INVALID_ARGUMENT, "use pis_batch_verlet_step_nve");
    return batch_run(ctx, &dt, nsteps, per_step_out);
}

pis_status pis_batch_verlet_step_nve(pis_ctx* ctx, const double* dts, uint64_t nsteps, pis_thermo* thermo_out)
{
    return batch_run(ctx, dts, nsteps, thermo_out);
}

pis_status pis_nvt_init(pis_ctx* ctx, const pis_nhc_params* params) { return nvt_init(ctx, params); }

pis_status pis_nvt_get_chain(pis_ctx* ctx, uint32_t sys, double* xi, double* target)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (sys >= ctx->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    if (!ctx->nhc_ready) return fail(PIS_ERR_INVALID_ARGUMENT, "NVT needs pis_nvt_init first");
    double st[NHC_COUNT];
    PIS_HIP(hipMemcpyAsync(st, ctx->sys[sys].dev.nhc, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    PIS_HIP(hipStreamSynchronize(ctx->stream));
    if (xi)
        for (int i = 0; i < (int)st[NHC_CHAIN]; ++i) xi[i] = st[NHC_XI + i];
    if (target) *target = st[NHC_T];
    return PIS_OK;
}

pis_status pis_verlet_step_nvt(pis_ctx* ctx, double dt, uint64_t nsteps, pis_thermo* per_step_out)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->sys.size() != 1) return fail(PIS_ERR_INVALID_ARGUMENT, "use pis_batch_verlet_step_nvt");
    return batch_run(ctx, &dt, nsteps, per_step_out, true);
}

pis_status pis_batch_verlet_step_nvt(pis_ctx* ctx, const double* dts, uint64_t nsteps, pis_thermo* thermo_out)
{
    return batch_run(ctx, dts, nsteps, thermo_out, true);
}

pis_status pis_batch_verlet_step_nvt_ramp(pis_ctx* ctx, const double* dts, uint64_t nsteps, uint64_t ramp_step0,
                                          uint64_t ramp_total, pis_thermo* thermo_out)
{
    if (ramp_total == 0) return fail(PIS_ERR_INVALID_ARGUMENT, "ramp_total must be > 0");
    return batch_run(ctx, dts, nsteps, thermo_out, true, ramp_step0, ramp_total);
}

pis_status pis_npt_init(pis_ctx* ctx, const pis_nhc_params* chain, const pis_mtk_params* barostat)
{
    return npt_init(ctx, chain, barostat);
}

pis_status pis_batch_verlet_step_npt(pis_ctx* ctx, const double* dts, uint64_t nsteps, uint64_t ramp_step0,
                                     uint64_t ramp_total, pis_thermo* thermo_out)
{
    if (ramp_total == 0) ramp_total = nsteps;
    return batch_run(ctx, dts, nsteps, thermo_out, true, ramp_step0, ramp_total, true);
}

pis_status pis_npt_get_box(pis_ctx* ctx, uint32_t sys, double* h_out, double* momentum_out)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (sys >= ctx->sys.size()) return fail(PIS_ERR_INVALID_ARGUMENT, "system index out of range");
    double bx[BOX_COUNT], st[NPT_COUNT];
    PIS_HIP(hipMemcpyAsync(bx, ctx->sys[sys].dev.box, sizeof(bx), hipMemcpyDeviceToHost, ctx->stream));
    PIS_HIP(hipMemcpyAsync(st, ctx->sys[sys].dev.npt, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    PIS_HIP(hipStreamSynchronize(ctx->stream));
    if (h_out)
        for (int k = 0; k < 9; ++k) h_out[k] = bx[BOX_H + k];
    if (momentum_out)
        for (int k = 0; k < 9; ++k) momentum_out[k] = st[NPT_MOM + k];
    return PIS_OK;
}

pis_status pis_get_neighbor_list(pis_ctx* ctx, uint64_t* offsets, uint32_t* idx, uint64_t cap, uint64_t* total)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->sys.size() != 1) return fail(PIS_ERR_INVALID_ARGUMENT, "single-system contexts only");
    if (!offsets || !total) return fail(PIS_ERR_INVALID_ARGUMENT, "null output");
    SysHost& h = ctx->sys[0];
    int32_t ctl[CTL_COUNT];
    PIS_TRY(read_ctl(ctx, 0, ctl));
    ctl[CTL_REBUILD] = 1;
    PIS_TRY(write_ctl(ctx, 0, ctl));
    PIS_TRY(run_rebuild(ctx));
    PIS_HIP(hipStreamSynchronize(ctx->stream));
    PIS_TRY(check_flags(ctx));
    const size_t n = (size_t)h.plan.n, mnb = (size_t)h.plan.maxnb;
    uint32_t* d_out = nullptr;
    PIS_HIP(hipMalloc(&d_out, (n * mnb + n) * 4));
    for (const auto& k : ctx->classes)
        launch_list_to_orig(ctx->stream, ctx->d_sys, ctx->d_tile_work + k.first, k.count, k.scap, d_out,
                            d_out + n * mnb);
    std::vector<uint32_t> buf(n * mnb + n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(buf.data(), d_out, buf.size() * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(PIS_ERR_HIP, std::string("list download: ") + hipGetErrorString(e));
    const uint32_t* cnt = buf.data() + n * mnb;
    offsets[0] = 0;
    for (size_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt[i];
    *total = offsets[n];
    if (idx)
        for (size_t i = 0; i < n; ++i)
            for (uint32_t k = 0; k < cnt[i]; ++k) {
                const uint64_t o = offsets[i] + k;
                if (o < cap) idx[o] = buf[i * mnb + k];
            }
    return PIS_OK;
}

pis_status pis_get_list_digest(pis_ctx* ctx, uint32_t* counts, uint64_t* hashes)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->sys.size() != 1) return fail(PIS_ERR_INVALID_ARGUMENT, "single-system contexts only");
    if (!counts || !hashes) return fail(PIS_ERR_INVALID_ARGUMENT, "null output");
    int32_t ctl[CTL_COUNT];
    PIS_TRY(read_ctl(ctx, 0, ctl));
    ctl[CTL_REBUILD] = 1;
    PIS_TRY(write_ctl(ctx, 0, ctl));
    PIS_TRY(run_rebuild(ctx));
    PIS_HIP(hipStreamSynchronize(ctx->stream));
    PIS_TRY(check_flags(ctx));
    const size_t n = (size_t)ctx->sys[0].plan.n;
    void* d_out = nullptr;
    PIS_HIP(hipMalloc(&d_out, n * 12));
    uint64_t* d_hash = (uint64_t*)d_out;
    uint32_t* d_cnt = (uint32_t*)(d_hash + n);
    for (const auto& k : ctx->classes)
        launch_list_digest(ctx->stream, ctx->d_sys, ctx->d_tile_work + k.first, k.count, k.scap, d_cnt, d_hash);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hashes, d_hash, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, d_cnt, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(PIS_ERR_HIP, std::string("list digest: ") + hipGetErrorString(e));
    return PIS_OK;
}

pis_status pis_get_stats(pis_ctx* ctx, pis_stats* out)
{
    if (!ctx || !out) return fail(PIS_ERR_INVALID_ARGUMENT, "null argument");
    int32_t ctl[CTL_COUNT];
    uint64_t rebuilds = 0;
    int32_t mx = 0;
    for (uint32_t k = 0; k < ctx->sys.size(); ++k) {
        PIS_TRY(read_ctl(ctx, k, ctl));
        rebuilds += (uint64_t)ctl[CTL_NREBUILD];
        mx = std::max(mx, ctl[CTL_MAXNB]);
    }
    out->steps = ctx->steps;
    out->rebuilds = rebuilds;
    out->force_launches = ctx->force_launches;
    out->force_kernel_ms = ctx->force_ms;
    out->timed_force_launches = ctx->timed_launches;
    out->neighbor_capacity = ctx->sys[0].plan.maxnb;
    out->max_neighbors = mx;
    return PIS_OK;
}

pis_status pis_copy_calibration(uint64_t bytes, int32_t reps, double* ms_per_copy)
{
    if (bytes < 32 || reps < 1) return fail(PIS_ERR_INVALID_ARGUMENT, "bytes >= 32 and reps >= 1 required");
    bytes &= ~(uint64_t)31;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PIS_ERR_NO_DEVICE, "no HIP device available");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        launch_copy_calibration(st, a, b, bytes);  // warm-up
        e = hipEventRecord(e0, st);
    }
    for (int r = 0; e == hipSuccess && r < reps; ++r) launch_copy_calibration(st, a, b, bytes);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(PIS_ERR_HIP, std::string("copy calibration: ") + hipGetErrorString(e));
    return PIS_OK;
}

pis_status pis_set_graph_mode(pis_ctx* ctx, int32_t enable)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    ctx->use_graph = enable != 0;
    return PIS_OK;
}

pis_status pis_set_kernel_timing(pis_ctx* ctx, int32_t enable)
{
    if (!ctx) return fail(PIS_ERR_INVALID_ARGUMENT, "null context");
    ctx->timing = enable != 0;
    ctx->force_ms = 0.0;
    ctx->timed_launches = 0;
    return PIS_OK;
}

}  // extern "C"
