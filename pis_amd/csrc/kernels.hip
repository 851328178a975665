// kernels.hip -- gfx950 kernels of the LJ hot path.
//
// Compiled with -ffp-contract=off: every a*b+c that must round twice (the reference's
// minimum image, wrap, binning, list predicate and integrator, which Rust never
// contracts) does; fused multiply-adds appear only where written as fma().
//
// Per NVE step (one launch each, all systems of a batch together):
//   k_integrate_pos   x += v dt + (F/m)/2 dt^2, wrap, skin check      (potential.rs:15-22)
//                     (first step of a call, and every NVT step; otherwise the previous
//                     step's force kernel has already done it)
//   k_bin             cell of every atom, integer histogram           (neighbour_list.rs:9-42)
//   k_scan            exclusive scan of the histogram per system
//   k_scatter         atoms into cell slots (atomic slot, then...)
//   k_rank_permute    ...deterministic rank by x (then original index) inside the cell, and the
//                     permutation of all per-atom state into the other buffer set
//   k_build_list      full neighbour list at rcut+skin, reference predicate bit-for-bit
//                     (lennard_jones.rs:344-416)
//   k_force           LDS-staged pair force + fused second half of the Verlet update
//                     (lennard_jones.rs:32-55, 419-455; potential.rs:26-30), then the first
//                     half of the next step when the call runs one (positions into the
//                     other buffer set, wrap, skin check)
//   k_reduce          fixed-order per-system PE/KE/virial sums      (properties.rs:17-65)
// The five rebuild kernels return at once unless the system's rebuild flag is set.
#include <math.h>

#include "kernels.h"

// Tuning knobs fixed at build time: candidates of the force kernel whose LDS loads are
// issued together (4 or 8); the minimum waves per SIMD its register allocation must allow
// (launch-bounds occupancy target) per precision; the grids (workgroups that loop over the
// work items) of the per-atom sort kernels and of the list build.
#ifndef PIS_FORCE_BATCH
#define PIS_FORCE_BATCH 8
#endif
#ifndef PIS_FORCE_MIN_WAVES_MIXED
#define PIS_FORCE_MIN_WAVES_MIXED 3
#endif
#ifndef PIS_FORCE_MIN_WAVES_FP64
#define PIS_FORCE_MIN_WAVES_FP64 1
#endif
#ifndef PIS_BUILD_UNROLL
#define PIS_BUILD_UNROLL 8
#endif
#ifndef PIS_BUILD_SHIFTIN
#define PIS_BUILD_SHIFTIN 1
#endif
#ifndef PIS_BUILD_WINTAB  // 1: row windows from the sub-cell table; 0: binary search
#define PIS_BUILD_WINTAB 1
#endif
#ifndef PIS_BUILD_MIN_WAVES  // waves per SIMD the list build's registers must allow
#define PIS_BUILD_MIN_WAVES 1
#endif
#ifndef PIS_SORT_GRID_CAP
#define PIS_SORT_GRID_CAP 1024
#endif
#ifndef PIS_BUILD_GRID_CAP
#define PIS_BUILD_GRID_CAP (1 << 30)
#endif

namespace pis {

// ---------------------------------------------------------------------------------
// helpers

// Global-address-space view of a device pointer.  Pointers read out of the system
// descriptors are generic to the compiler, which then emits FLAT memory instructions; a
// FLAT access also counts against the LDS/scalar-load wait counter, so the next LDS or
// scalar wait would wait for it to reach memory.  The cast (the address is unchanged) lets
// the compiler emit GLOBAL instructions instead.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range
// of tiles so neighbouring tiles (which share halo atoms) hit the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk)
{
    const int X = 8;
    const int per = nblk / X, rem = nblk % X;
    const int xcd = bid % X, loc = bid / X;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// Minimum image of one orthorhombic component, any tie rule (for staging and forces,
// where a tie means |d| = L/2 > cutoff and the pair is rejected either way).
__device__ __forceinline__ double mi_fast(double d, double L, double Linv)
{
    const double k = rint(d * Linv);
    return fma(-k, L, d);
}

__device__ __forceinline__ float mi_fast32(float d, float L, float Linv)
{
    const float k = rintf(d * Linv);
    return fmaf(-k, L, d);
}

// apply_boundary_conditions_dis (simulation_box.rs:17-27) for an orthorhombic box:
// s = h^-1 d, s -= round(s) (half away from zero), d = h s.  The off-diagonal terms of the
// reference's gemv are signed zeros and do not change a nonzero result.
__device__ __forceinline__ double mi_ref(double d, double L, double Linv)
{
    double s = d * Linv;
    s = s - round(s);
    return L * s;
}

// nalgebra Matrix3 * Vector3 (gemv, column by column): y = m[:,0] x0, then y = m[:,c] xc + y,
// each product and sum rounded (the reference never fuses).  m is column-major.
__device__ __forceinline__ void gemv3(const double* m, double x0, double x1, double x2, double& y0, double& y1,
                                      double& y2)
{
    double t0 = m[0] * x0, t1 = m[1] * x0, t2 = m[2] * x0;
    t0 = m[3] * x1 + t0;
    t1 = m[4] * x1 + t1;
    t2 = m[5] * x1 + t2;
    t0 = m[6] * x2 + t0;
    t1 = m[7] * x2 + t1;
    t2 = m[8] * x2 + t2;
    y0 = t0;
    y1 = t1;
    y2 = t2;
}

// apply_boundary_conditions_dis (simulation_box.rs:17-27) for a general box, bit for bit:
// s = h^-1 d, s -= round(s), d = h s.  box: h and h^-1 (BOX_H, BOX_HINV).
__device__ __forceinline__ void mi_tri(const double* box, double& dx, double& dy, double& dz)
{
    double s0, s1, s2;
    gemv3(box + BOX_HINV, dx, dy, dz, s0, s1, s2);
    s0 = s0 - round(s0);
    s1 = s1 - round(s1);
    s2 = s2 - round(s2);
    gemv3(box + BOX_H, s0, s1, s2, dx, dy, dz);
}

// apply_boundary_conditions_pos (simulation_box.rs:29-42) for a general box:
// s = h^-1 r, s -= floor(s), r = h s.
__device__ __forceinline__ void wrap_tri(const double* box, double& x, double& y, double& z)
{
    double s0, s1, s2;
    gemv3(box + BOX_HINV, x, y, z, s0, s1, s2);
    s0 = s0 - floor(s0);
    s1 = s1 - floor(s1);
    s2 = s2 - floor(s2);
    gemv3(box + BOX_H, s0, s1, s2, x, y, z);
}

// FP32 minimum image of a general box for the list build's staged coordinates (any
// rounding: the list band covers it).  h32 / hi32: h and h^-1, column-major.
__device__ __forceinline__ void mi_tri32(const float* h32, const float* hi32, float& dx, float& dy, float& dz)
{
    float s0 = fmaf(hi32[6], dz, fmaf(hi32[3], dy, hi32[0] * dx));
    float s1 = fmaf(hi32[7], dz, fmaf(hi32[4], dy, hi32[1] * dx));
    float s2 = fmaf(hi32[8], dz, fmaf(hi32[5], dy, hi32[2] * dx));
    s0 -= rintf(s0);
    s1 -= rintf(s1);
    s2 -= rintf(s2);
    dx = fmaf(h32[6], s2, fmaf(h32[3], s1, h32[0] * s0));
    dy = fmaf(h32[7], s2, fmaf(h32[4], s1, h32[1] * s0));
    dz = fmaf(h32[8], s2, fmaf(h32[5], s1, h32[2] * s0));
}

// The same in FP64 for the force kernel's staged raw positions (any rounding: outside
// the rounding band the cutoff decision does not depend on it, inside the band the
// reference predicate is evaluated exactly).
__device__ __forceinline__ void mi_tri_fast(const double* box, double& dx, double& dy, double& dz)
{
    const double* hi = box + BOX_HINV;
    const double* h = box + BOX_H;
    double s0 = fma(hi[6], dz, fma(hi[3], dy, hi[0] * dx));
    double s1 = fma(hi[7], dz, fma(hi[4], dy, hi[1] * dx));
    double s2 = fma(hi[8], dz, fma(hi[5], dy, hi[2] * dx));
    s0 -= rint(s0);
    s1 -= rint(s1);
    s2 -= rint(s2);
    dx = fma(h[6], s2, fma(h[3], s1, h[0] * s0));
    dy = fma(h[7], s2, fma(h[4], s1, h[1] * s0));
    dz = fma(h[8], s2, fma(h[5], s1, h[2] * s0));
}

// Exclusive scan over the block (blockDim.x a multiple of 64, <= 1024).
__device__ int block_excl_scan(int v, int* tmp, int* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    int add = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const int t = tmp[k];
        add += (k < wid) ? t : 0;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return add + x - v;
}

// Runs of equal, adjacent keys among the active lanes of a wave (the active lanes form a
// prefix of the wave): this lane's run head (a lane id) and the run's length.  Atoms arrive
// in the previous cell order, so the lanes of a wave span a few cells and one atomic per
// run replaces one per atom.
__device__ __forceinline__ void wave_runs(int key, bool active, int* head, int* len)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const uint64_t heads = __ballot(active && (lane == 0 || prev != key));
    const uint64_t act = __ballot(active);
    const uint64_t upto = (2ull << lane) - 1;  // lanes 0..lane (all ones for lane 63)
    const int h = 63 - __clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const int end = after ? __ffsll((unsigned long long)after) - 1 : __popcll(act);
    *head = h;
    *len = end - h;
}

// Whether any work item this block will visit (items blockIdx.x + k gridDim.x, through
// xcd_remap when REMAP) belongs to a system with its rebuild flag set.  One parallel probe:
// on steps without a rebuild a block returns after two dependent loads instead of walking
// its items one by one.
template <bool REMAP>
__device__ __forceinline__ bool block_rebuild_pending(const SysDev* __restrict__ sys,
                                                      const WorkItem* __restrict__ work, int nwork)
{
    int flag = 0;
    for (int b = blockIdx.x + (int)threadIdx.x * (int)gridDim.x; b < nwork; b += (int)(blockDim.x * gridDim.x))
        flag |= sys[work[REMAP ? xcd_remap(b, nwork) : b].sys].ctl[CTL_REBUILD];
    return __syncthreads_or(flag) != 0;
}

// Deterministic sum of 4 doubles over the block (<= 16 waves); result valid in thread 0.
__device__ void block_sum4(double v[4], double* tmp /* [16 waves * 4] */)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    if (lane == 0)
        for (int k = 0; k < 4; ++k) tmp[wid * 4 + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += tmp[w * 4 + k];
            v[k] = s;
        }
}

struct TileGeom {
    int x0, y0, z0, sx, sy, sz, ne, c0, c1;
};

// Cartesian centre of a tile: the cell-index centre times the cell widths (orthorhombic),
// or the box applied to the fractional centre (general boxes, whose cells are slices of
// the fractional coordinates).
__device__ __forceinline__ void tile_center(const SysDev& s, const TileGeom& g, double& cx, double& cy,
                                            double& cz)
{
    if (s.tri) {
        gemv3(s.box + BOX_H, (g.x0 + 0.5 * g.sx) / s.nc[0], (g.y0 + 0.5 * g.sy) / s.nc[1],
              (g.z0 + 0.5 * g.sz) / s.nc[2], cx, cy, cz);
    } else {
        cx = (g.x0 + 0.5 * g.sx) * s.cellw[0];
        cy = (g.y0 + 0.5 * g.sy) * s.cellw[1];
        cz = (g.z0 + 0.5 * g.sz) * s.cellw[2];
    }
}

// Position relative to a tile centre at the image nearest to it (tile-relative staging).
__device__ __forceinline__ void rel_to_center(const SysDev& s, double cx, double cy, double cz, double& x,
                                              double& y, double& z)
{
    if (s.tri) {
        x -= cx;
        y -= cy;
        z -= cz;
        mi_tri_fast(s.box, x, y, z);
    } else {
        x = mi_fast(x - cx, s.L[0], s.Linv[0]);
        y = mi_fast(y - cy, s.L[1], s.Linv[1]);
        z = mi_fast(z - cz, s.L[2], s.Linv[2]);
    }
}

__device__ __forceinline__ TileGeom tile_geom(const SysDev& s, int t)
{
    TileGeom g;
    const int tx = t % s.nt[0], ty = (t / s.nt[0]) % s.nt[1], tz = t / (s.nt[0] * s.nt[1]);
    g.x0 = tx * s.tile[0];
    g.y0 = ty * s.tile[1];
    g.z0 = tz * s.tile[2];
    g.sx = min(s.tile[0], s.nc[0] - g.x0);
    g.sy = min(s.tile[1], s.nc[1] - g.y0);
    g.sz = min(s.tile[2], s.nc[2] - g.z0);
    g.ne = s.mode == MODE_WHOLE ? 1 : (g.sx + 2) * (g.sy + 2) * (g.sz + 2);
    g.c0 = s.tile_cell0[t];
    g.c1 = s.tile_cell0[t + 1];
    return g;
}

__device__ __forceinline__ int wrap_cell(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

enum { STAGE_RAW = 0, STAGE_REL = 1, STAGE_INDEX = 2 };
constexpr double kPadFar = 1e30;  // squared: 1e60, finite in FP64, +inf in FP32

// Stage the atoms of a tile's halo box (or of the whole system) into LDS, cell by cell
// in halo order.  pref[e] is the first slot of halo cell e.  Slot SCAP is the padding
// that list entries past an atom's count point to: with tile-relative positions it lies
// kPadFar away, beyond any cutoff, so those entries fail the cutoff test by themselves;
// with raw positions (a per-pair minimum image would fold it back) it sits at the origin
// and the force loop masks those entries by count.
// Returns the number of staged atoms (uniform over the block); > SCAP means overflow
// and nothing was staged.  smap (SCAP + 1 bytes, may be null) receives the halo cell of
// every slot, each cell's thread filling its run, so a slot finds its atom with one LDS
// read instead of a binary search over the halo's cell offsets.  Staged coordinates are
// multiplied by `scale` (1, or 1/sigma for the single-species mixed loop).
template <int SCAP, int WHAT, typename T>
__device__ int stage_tile(const SysDev& s, const State& st, const TileGeom& g, double cx, double cy,
                          double cz, T* Px, T* Py, T* Pz, uint8_t* Pt, int* pref, int* ecell, int* stmp,
                          uint8_t* smap = nullptr, double scale = 1.0)
{
    const int tid = threadIdx.x;
    int cnt = 0;
    const int ne = g.ne;
    if (s.mode == MODE_WHOLE) {
        if (tid == 0) {
            ecell[0] = 0;
            cnt = s.n;
        }
    } else if (tid < ne) {
        const int hx = g.sx + 2, hy = g.sy + 2;
        const int ex = tid % hx, ey = (tid / hx) % hy, ez = tid / (hx * hy);
        const int cxn = wrap_cell(g.x0 - 1 + ex, s.nc[0]);
        const int cyn = wrap_cell(g.y0 - 1 + ey, s.nc[1]);
        const int czn = wrap_cell(g.z0 - 1 + ez, s.nc[2]);
        const int sc = gbl(s.nat2sorted)[(czn * s.nc[1] + cyn) * s.nc[0] + cxn];
        const int b = gbl(s.cell_start)[sc];
        ecell[tid] = b;
        cnt = gbl(s.cell_start)[sc + 1] - b;
    }
    int total;
    const int ex = block_excl_scan(cnt, stmp, &total);
    if (tid < ne) pref[tid] = ex;
    if (tid == 0) pref[ne] = total;
    __syncthreads();
    if (total > SCAP) return total;
    const bool mapped = smap != nullptr && ne > 1;
    if (mapped) {
        if (tid < ne)
            for (int sl = ex; sl < ex + cnt; ++sl) smap[sl] = (uint8_t)tid;
        __syncthreads();
    }
    for (int sl = tid; sl < total; sl += blockDim.x) {
        int lo = 0;
        if (mapped) {
            lo = smap[sl];
        } else {
            int hi = ne - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pref[mid] <= sl)
                    lo = mid;
                else
                    hi = mid - 1;
            }
        }
        const int a = ecell[lo] + (sl - pref[lo]);
        if (WHAT == STAGE_INDEX) {
            reinterpret_cast<int*>(Px)[sl] = a;
        } else {
            double x = gbl(st.x)[a], y = gbl(st.y)[a], z = gbl(st.z)[a];
            if (WHAT == STAGE_REL) rel_to_center(s, cx, cy, cz, x, y, z);
            Px[sl] = (T)(x * scale);
            Py[sl] = (T)(y * scale);
            Pz[sl] = (T)(z * scale);
            if (Pt) Pt[sl] = gbl(st.type)[a];  // the force kernel reads types from the list
        }
    }
    if (tid == 0 && WHAT != STAGE_INDEX) {
        const T pad = WHAT == STAGE_REL ? (T)kPadFar : (T)0;
        Px[SCAP] = pad;
        Py[SCAP] = pad;
        Pz[SCAP] = pad;
        if (Pt) Pt[SCAP] = 0;
    }
    __syncthreads();
    return total;
}

// Tile-local halo entry and LDS slot of own atom `a` (modes TILE_REL / RAW_MI); ownf[k] is
// the first atom of the tile's own cell k (tile order, x fastest), staged in LDS.
__device__ __forceinline__ void own_slot(const TileGeom& g, const int* pref, const int* ownf, int a, int* e0,
                                         int* slot, int* ly_out = nullptr, int* lz_out = nullptr)
{
    int lo = 0, hi = g.sx * g.sy * g.sz - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ownf[mid] <= a)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int k = lo;
    const int lx = k % g.sx, ly = (k / g.sx) % g.sy, lz = k / (g.sx * g.sy);
    *e0 = ((lz + 1) * (g.sy + 2) + (ly + 1)) * (g.sx + 2) + (lx + 1);
    *slot = pref[*e0] + (a - ownf[k]);
    if (ly_out) *ly_out = ly;
    if (lz_out) *lz_out = lz;
}

// First index in [lo, hi) of the ascending LDS run P with P[i] >= v (or > v for UPPER).
template <bool UPPER>
__device__ __forceinline__ int search_sorted(const float* P, int lo, int hi, double v)
{
    int n = hi - lo;
    while (n > 0) {
        const int h = n >> 1;
        const bool right = UPPER ? P[lo + h] <= v : P[lo + h] < v;
        lo = right ? lo + h + 1 : lo;
        n = right ? n - h - 1 : h;
    }
    return lo;
}

// ---------------------------------------------------------------------------------
// state transfer

__global__ void k_load_aos(const SysDev* __restrict__ sys, int k, const double* __restrict__ pos,
                           const double* __restrict__ vel, const double* __restrict__ force,
                           const uint8_t* __restrict__ type)
{
    const SysDev& s = sys[k];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    const State st = s.st[0];
    st.x[i] = pos[3 * i];
    st.y[i] = pos[3 * i + 1];
    st.z[i] = pos[3 * i + 2];
    st.vx[i] = vel ? vel[3 * i] : 0.0;
    st.vy[i] = vel ? vel[3 * i + 1] : 0.0;
    st.vz[i] = vel ? vel[3 * i + 2] : 0.0;
    st.fx[i] = force ? force[3 * i] : 0.0;
    st.fy[i] = force ? force[3 * i + 1] : 0.0;
    st.fz[i] = force ? force[3 * i + 2] : 0.0;
    st.type[i] = type[i];
    st.orig[i] = (uint32_t)i;
}

__global__ void k_store_aos(const SysDev* __restrict__ sys, int k, double* __restrict__ pos,
                            double* __restrict__ vel, double* __restrict__ force)
{
    const SysDev& s = sys[k];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= s.n) return;
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    const uint32_t o = st.orig[p];
    if (pos) {
        pos[3 * o] = st.x[p];
        pos[3 * o + 1] = st.y[p];
        pos[3 * o + 2] = st.z[p];
    }
    if (vel) {
        vel[3 * o] = st.vx[p];
        vel[3 * o + 1] = st.vy[p];
        vel[3 * o + 2] = st.vz[p];
    }
    if (force) {
        force[3 * o] = st.fx[p];
        force[3 * o + 1] = st.fy[p];
        force[3 * o + 2] = st.fz[p];
    }
}

__global__ void k_set_ctl(const SysDev* __restrict__ sys, int nsys, int slot, int value)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsys) sys[k].ctl[slot] = value;
}

// ---------------------------------------------------------------------------------
// integrator, first half (potential.rs:16-22) + skin trigger

__global__ __launch_bounds__(kAtomThreads) void k_integrate_pos(const SysDev* __restrict__ sys,
                                                                 const WorkItem* __restrict__ work)
{
    const WorkItem w = work[blockIdx.x];
    const SysDev& s = sys[w.sys];
    const int a = w.idx + threadIdx.x;
    if (a >= s.n || s.ctl[CTL_LEFT] == 0) return;
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    const double m = gbl(s.mass)[gbl(st.type)[a]];
    const double dt = s.dt, dt2 = s.dt2;
    double vx = gbl(st.vx)[a], vy = gbl(st.vy)[a], vz = gbl(st.vz)[a];
    if (s.nhc_on) {
        // NVT: v *= exp(-dt xi0 / 2) before the NVE step (potential.rs:44-45)
        const double sc = s.nhc[NHC_SCALE];
        vx = vx * sc;
        vy = vy * sc;
        vz = vz * sc;
        gbl(st.vx)[a] = vx;
        gbl(st.vy)[a] = vy;
        gbl(st.vz)[a] = vz;
    }
    // a_t = F/m (current_acceleration, properties.rs:32-39); x += v dt + (a 0.5) dt^2
    double x = gbl(st.x)[a] + (vx * dt + (gbl(st.fx)[a] / m) * 0.5 * dt2);
    double y = gbl(st.y)[a] + (vy * dt + (gbl(st.fy)[a] / m) * 0.5 * dt2);
    double z = gbl(st.z)[a] + (vz * dt + (gbl(st.fz)[a] / m) * 0.5 * dt2);
    // apply_boundary_conditions_pos (simulation_box.rs:29-42): s = h^-1 r; s -= floor(s); r = h s
    double dx, dy, dz;
    double lim = s.half_skin2;
    if (s.tri) {
        wrap_tri(s.box, x, y, z);
        double ux = gbl(s.rx)[a], uy = gbl(s.ry)[a], uz = gbl(s.rz)[a];
        if (s.npt_on) {
            // displacement from the build-time position carried along by the box
            // deformation since the build (the periodic images move with the box)
            const __attribute__((address_space(1))) double* npt = gbl(s.npt);
            double D[9];
            for (int k = 0; k < 9; ++k) D[k] = npt[NPT_DEF + k];
            gemv3(D, ux, uy, uz, ux, uy, uz);
            lim = npt[NPT_SKIN2];
        }
        dx = x - ux;
        dy = y - uy;
        dz = z - uz;
        mi_tri_fast(s.box, dx, dy, dz);
    } else {
        double sx = x * s.Linv[0], sy = y * s.Linv[1], sz = z * s.Linv[2];
        sx = sx - floor(sx);
        sy = sy - floor(sy);
        sz = sz - floor(sz);
        x = s.L[0] * sx;
        y = s.L[1] * sy;
        z = s.L[2] * sz;
        dx = mi_fast(x - gbl(s.rx)[a], s.L[0], s.Linv[0]);
        dy = mi_fast(y - gbl(s.ry)[a], s.L[1], s.Linv[1]);
        dz = mi_fast(z - gbl(s.rz)[a], s.L[2], s.Linv[2]);
    }
    gbl(st.x)[a] = x;
    gbl(st.y)[a] = y;
    gbl(st.z)[a] = z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 <= lim)) {  // also catches NaN
        atomicOr(&s.ctl[CTL_REBUILD], 1);
        if (d2 != d2) atomicOr(&s.ctl[CTL_NAN], 1);
    }
}

// ---------------------------------------------------------------------------------
// rebuild: counting sort by cell, deterministic within a cell

__global__ __launch_bounds__(kAtomThreads) void k_bin(const SysDev* __restrict__ sys,
                                                       const WorkItem* __restrict__ work, int nwork)
{
    if (!block_rebuild_pending<false>(sys, work, nwork)) return;
    for (int b = blockIdx.x; b < nwork; b += gridDim.x) {
        const WorkItem w = work[b];
        const SysDev& s = sys[w.sys];
        if (!s.ctl[CTL_REBUILD]) continue;  // uniform over the block: one system per item
        const int a = w.idx + threadIdx.x;
        const bool act = a < s.n;
        int sc = 0;
        if (act) {
            const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
            // rcut_cells (neighbour_list.rs:17-34): c = floor(s n) as usize (saturating), % n
            double r[3] = {gbl(st.x)[a], gbl(st.y)[a], gbl(st.z)[a]};
            if (s.tri) {  // fractional coordinates s = h^-1 r
                double f0, f1, f2;
                gemv3(s.box + BOX_HINV, r[0], r[1], r[2], f0, f1, f2);
                r[0] = f0;
                r[1] = f1;
                r[2] = f2;
            } else {
                r[0] = r[0] * s.Linv[0];
                r[1] = r[1] * s.Linv[1];
                r[2] = r[2] * s.Linv[2];
            }
            int c[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double f = floor(r[d] * (double)s.nc[d]);
                const int v = f > 0.0 ? (f < 2.0e9 ? (int)f : 2000000000) : 0;
                c[d] = v % s.nc[d];
            }
            sc = gbl(s.nat2sorted)[(c[2] * s.nc[1] + c[1]) * s.nc[0] + c[0]];
            gbl(s.atom_cell)[a] = sc;
        }
        int head, len;
        wave_runs(sc, act, &head, &len);
        if (act && (int)(threadIdx.x & 63) == head) atomicAdd(&s.cell_count[sc], len);
    }
}

// Exclusive scan of the cell histogram of one system by one workgroup: each thread owns a
// contiguous range of cells (a multiple of 4, read as int4), one block-wide scan of the
// range sums, then the ranges are written.  The histogram itself returns to zero in
// k_scatter.
__global__ __launch_bounds__(1024) void k_scan(const SysDev* __restrict__ sys)
{
    __shared__ int tmp[16];
    const SysDev& s = sys[blockIdx.x];
    if (!s.ctl[CTL_REBUILD]) return;
    const int n = s.ncell;
    const int per = ((n + 4 * (int)blockDim.x - 1) / (4 * (int)blockDim.x)) * 4;
    const int i0 = (int)threadIdx.x * per, i1 = min(i0 + per, n);
    int sum = 0;
    int i = i0;
    // four int4 loads in flight per thread (the loop is latency-bound otherwise)
    for (; i + 16 <= i1; i += 16) {
        int4 c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = *reinterpret_cast<const int4*>(s.cell_count + i + 4 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) sum += (c[u].x + c[u].y) + (c[u].z + c[u].w);
    }
    for (; i < i1; i += 4) {
        if (i + 4 <= i1) {
            const int4 c = *reinterpret_cast<const int4*>(s.cell_count + i);
            sum += (c.x + c.y) + (c.z + c.w);
        } else {
            for (int k = i; k < i1; ++k) sum += s.cell_count[k];
        }
    }
    int total;
    int run = block_excl_scan(sum, tmp, &total);
    const int* __restrict__ cc = s.cell_count;
    int* __restrict__ cs = s.cell_start;
    i = i0;
    // again four int4 loads issued before their stores (the arrays do not overlap)
    for (; i + 16 <= i1; i += 16) {
        int4 c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = *reinterpret_cast<const int4*>(cc + i + 4 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int4 o;
            o.x = run;
            o.y = o.x + c[u].x;
            o.z = o.y + c[u].y;
            o.w = o.z + c[u].z;
            run = o.w + c[u].w;
            *reinterpret_cast<int4*>(cs + i + 4 * u) = o;
        }
    }
    for (; i < i1; i += 4) {
        if (i + 4 <= i1) {
            const int4 c = *reinterpret_cast<const int4*>(cc + i);
            int4 o;
            o.x = run;
            o.y = o.x + c.x;
            o.z = o.y + c.y;
            o.w = o.z + c.z;
            run = o.w + c.w;
            *reinterpret_cast<int4*>(cs + i) = o;
        } else {
            for (int k = i; k < i1; ++k) {
                s.cell_start[k] = run;
                run += s.cell_count[k];
            }
        }
    }
    if (threadIdx.x == 0) s.cell_start[n] = total;
}

// Each run of same-cell lanes takes a block of slots with one atomicSub; the decrements
// bring the histogram back to zero for the next rebuild.
__global__ __launch_bounds__(kAtomThreads) void k_scatter(const SysDev* __restrict__ sys,
                                                           const WorkItem* __restrict__ work, int nwork)
{
    if (!block_rebuild_pending<false>(sys, work, nwork)) return;
    const int lane = threadIdx.x & 63;
    for (int b = blockIdx.x; b < nwork; b += gridDim.x) {
        const WorkItem w = work[b];
        const SysDev& s = sys[w.sys];
        if (!s.ctl[CTL_REBUILD]) continue;
        const int a = w.idx + threadIdx.x;
        const bool act = a < s.n;
        const int sc = act ? gbl(s.atom_cell)[a] : 0;
        int head, len;
        wave_runs(sc, act, &head, &len);
        int old = 0;
        if (act && lane == head) old = atomicSub(&s.cell_count[sc], len);
        old = __shfl(old, head, 64);
        if (act) {
            const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
            const int p = gbl(s.cell_start)[sc] + old - len + (lane - head);
            gbl(s.scat_src)[p] = a;
            gbl(s.scat_key)[p] = gbl(st.orig)[a];
            gbl(s.scat_x)[p] = gbl(st.x)[a];
        }
    }
}

// Inside a cell, atoms end up in ascending x (ties by original index), whatever order
// the atomics placed them in: the order is a function of the positions alone, and a row
// of cells along x is then sorted by x, which the list build's x windows rely on.  The
// sort keys of the cells a block's 256 slots fall in are staged in LDS when they fit.
constexpr int kRankStage = 2048;

__global__ __launch_bounds__(kAtomThreads) void k_rank_permute(const SysDev* __restrict__ sys,
                                                                const WorkItem* __restrict__ work,
                                                                int nwork)
{
    __shared__ double kx[kRankStage];
    __shared__ uint32_t kk[kRankStage];
    __shared__ int range[2];
    if (!block_rebuild_pending<false>(sys, work, nwork)) return;
    for (int b = blockIdx.x; b < nwork; b += gridDim.x) {
        const WorkItem w = work[b];
        const SysDev& s = sys[w.sys];
        if (!s.ctl[CTL_REBUILD]) continue;  // uniform over the block
        const int p = w.idx + threadIdx.x;
        const int last = min(w.idx + (int)blockDim.x, s.n) - 1;
        const bool act = p <= last;
        int a = 0, sc = 0;
        if (act) {
            a = gbl(s.scat_src)[p];
            sc = gbl(s.atom_cell)[a];
        }
        const int cur = __builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR]);
        const State src = s.st[cur], dst = s.st[cur ^ 1];
        // the atom's state is read before the cell's keys are staged, so that its loads are
        // in flight behind the staging's
        double x = 0.0, y = 0.0, z = 0.0, v[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
        uint8_t ty = 0;
        if (act) {
            x = gbl(src.x)[a];
            y = gbl(src.y)[a];
            z = gbl(src.z)[a];
            v[0] = gbl(src.vx)[a];
            v[1] = gbl(src.vy)[a];
            v[2] = gbl(src.vz)[a];
            f[0] = gbl(src.fx)[a];
            f[1] = gbl(src.fy)[a];
            f[2] = gbl(src.fz)[a];
            ty = gbl(src.type)[a];
        }
        // slots are grouped by cell, so the block's cells are those of its first and last slot
        if (threadIdx.x == 0) range[0] = gbl(s.cell_start)[sc];
        if (p == last) range[1] = gbl(s.cell_start)[sc + 1];
        __syncthreads();
        const int base = range[0], cnt = range[1] - range[0];
        const bool staged = cnt <= kRankStage;
        if (staged)
            for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
                kx[q] = gbl(s.scat_x)[base + q];
                kk[q] = gbl(s.scat_key)[base + q];
            }
        __syncthreads();
        if (act) {
            const int c0 = gbl(s.cell_start)[sc], c1 = gbl(s.cell_start)[sc + 1];
            const uint32_t key = gbl(s.scat_key)[p];
            int rank = 0;
            if (staged) {
                for (int q = c0 - base; q < c1 - base; ++q) {
                    const double xq = kx[q];
                    const uint32_t kq = kk[q];
                    rank += (xq < x || (xq == x && kq < key)) ? 1 : 0;
                }
            } else {
                for (int q = c0; q < c1; ++q) {
                    const double xq = gbl(s.scat_x)[q];
                    const uint32_t kq = gbl(s.scat_key)[q];
                    rank += (xq < x || (xq == x && kq < key)) ? 1 : 0;
                }
            }
            const int np = c0 + rank;
            gbl(dst.x)[np] = x;
            gbl(dst.y)[np] = y;
            gbl(dst.z)[np] = z;
            gbl(s.rx)[np] = x;
            gbl(s.ry)[np] = y;
            gbl(s.rz)[np] = z;
            gbl(dst.vx)[np] = v[0];
            gbl(dst.vy)[np] = v[1];
            gbl(dst.vz)[np] = v[2];
            gbl(dst.fx)[np] = f[0];
            gbl(dst.fy)[np] = f[1];
            gbl(dst.fz)[np] = f[2];
            gbl(dst.type)[np] = ty;
            gbl(dst.orig)[np] = key;
        }
        __syncthreads();  // the next item restages the keys
    }
}

// Atom index of LDS slot `sl` (the staging is cell by cell in halo order).
__device__ __forceinline__ int slot_atom(const int* pref, const int* ecell, int ne, int sl)
{
    int lo = 0, hi = ne - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= sl)
            lo = mid;
        else
            hi = mid - 1;
    }
    return ecell[lo] + (sl - pref[lo]);
}

// The reference's predicate for one pair, bit for bit: rij = r_j - r_i from the raw
// positions, apply_boundary_conditions_dis with round(), r2 = (x^2+y^2)+z^2 without
// fusion, accepted when r2 <= r2max (<=> norm() <= rcut).
template <bool TRI>
__device__ __forceinline__ bool exact_within(const SysDev& s, const State& st, int a, double xi,
                                             double yi, double zi, double r2max)
{
    double dx = gbl(st.x)[a] - xi, dy = gbl(st.y)[a] - yi, dz = gbl(st.z)[a] - zi;
    if (TRI) {
        mi_tri(s.box, dx, dy, dz);
    } else {
        dx = mi_ref(dx, s.L[0], s.Linv[0]);
        dy = mi_ref(dy, s.L[1], s.Linv[1]);
        dz = mi_ref(dz, s.L[2], s.Linv[2]);
    }
    return (dx * dx + dy * dy) + dz * dz <= r2max;
}

// m = 2 m + (r2 <= b): the compare lands in an SGPR pair that the add takes as its carry-in,
// two VALU instructions per candidate bit (VALU only; v_cmp_ge(b, r2) is false for NaN).
__device__ __forceinline__ void shift_in_le(uint32_t& m, float r2, float b)
{
    uint64_t c;
    asm("v_cmp_ge_f32_e64 %1, %3, %2\n\t"
        "v_addc_co_u32_e64 %0, %1, %0, %0, %1"
        : "+v"(m), "=&s"(c)
        : "v"(r2), "v"(b));
}

// Maximum of v over the wave (all lanes active), uniform: DPP within each row of 16 lanes
// (quad swaps, half-row and row mirrors), then the four row maxima through scalar
// registers.  No LDS round trips, unlike a shuffle reduction.
__device__ __forceinline__ int wave_max(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return max(max(r0, r1), max(r2, r3));
}

// Window table of an orthorhombic tile-relative staging: each halo cell is cut into 8
// equal slices along x, and wtab[8 e + k] is the first slot of cell e whose staged x lies
// in slice k or above (wtab[8 e] = the cell's first slot; entry 8 ne = the staging's end).
// Cells are x-sorted, so the 3-cell row starting at cell c reads its window bounds at
// wtab[8 c + G] for G in [0, 24] without a search.
constexpr int kWinSlices = 8;
__device__ __forceinline__ int win_slice(float x, float face, float inv_h)
{
    return min(max((int)floorf((x - face) * inv_h), 0), kWinSlices - 1);
}

// Descending sort of the 9 row keys (window length << 16 | first slot): a 25-comparator
// network (three sorted triples merged), every index static so the keys stay in registers.
__device__ __forceinline__ void sort9_desc(uint32_t* k)
{
    constexpr int net[25][2] = {{0, 1}, {3, 4}, {6, 7}, {1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7},
                                {0, 3}, {3, 6}, {0, 3}, {1, 4}, {4, 7}, {1, 4}, {2, 5}, {5, 8}, {2, 5},
                                {1, 3}, {5, 7}, {2, 6}, {4, 6}, {2, 4}, {2, 3}, {5, 6}};
#pragma unroll
    for (int q = 0; q < 25; ++q) {
        const uint32_t a = k[net[q][0]], b = k[net[q][1]];
        k[net[q][0]] = max(a, b);
        k[net[q][1]] = min(a, b);
    }
}

typedef __attribute__((address_space(3))) const float lds_cf32;

// An LDS address the compiler cannot relate to the other staging arrays, so the reads at
// small constant offsets from it share one base register through the instruction's offset
// field instead of taking an address add each.
__device__ __forceinline__ lds_cf32* lds_opaque(const float* p)
{
    lds_cf32* q = (lds_cf32*)p;
    asm volatile("" : "+v"(q));
    return q;
}


// Full neighbour list at rcut_ij + skin, accepted by the predicate of build_neighbour_list
// (lennard_jones.rs:398-405) decided exactly; entries are LDS byte offsets.
// Neighbour list of the own atom held by this lane (lane per atom).  The lane's stencil is
// 9 contiguous (dy,dz) rows of the staging; each row is first cut to the x window that can
// hold a partner, and the lane takes its rows longest first, so that the wave's lanes
// sweep windows of similar length together (the trip count of a row is the wave's longest
// window).  A row is swept in chunks of up to 64 candidates with a trip count uniform over
// the wave, one bit per accepted candidate; the set bits are then appended to a 128-bit
// shift register that leaves as one uint4 list group per 8 entries.  r2 from the staged
// coordinates decides every pair outside a +-band around the cutoff; inside the band the
// bit-exact reference predicate (exact_within, raw positions) decides.
template <int SCAP, bool PAIRMI, int NT, bool TRI>
__device__ __forceinline__ int build_atom(const SysDev& s, const State& st, const TileGeom& g,
                                          const float* Px, const float* Py, const float* Pz,
                                          const uint8_t* Pt, const int* pref, const int* ecell,
                                          const int* ownf, const PairParam* ppl, const uint16_t* wtab,
                                          int total, bool active, int a, double cx, double cy, double cz)
{
    // NT: 1 (one species), 2 (a binary mixture) or kMaxTypes (any count up to that)
    constexpr bool MULTI = NT > 1;
    const int nt = s.ntypes;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    int ti = 0, e0 = 0, self = -1, ly = 0, lz = 0;
    if (active) {
        xi = gbl(st.x)[a];
        yi = gbl(st.y)[a];
        zi = gbl(st.z)[a];
        ti = gbl(st.type)[a];
        if (s.mode == MODE_WHOLE)
            self = a;
        else
            own_slot(g, pref, ownf, a, &e0, &self, &ly, &lz);
    }
    double qx = xi, qy = yi, qz = zi;
    if (!PAIRMI) rel_to_center(s, cx, cy, cz, qx, qy, qz);
    const float qx32 = (float)qx, qy32 = (float)qy, qz32 = (float)qz;
    const float L32[3] = {(float)s.L[0], (float)s.L[1], (float)s.L[2]};
    const float Li32[3] = {(float)s.Linv[0], (float)s.Linv[1], (float)s.Linv[2]};
    float H32[9], HI32[9];  // general box (TRI): h and h^-1 in FP32
    if (TRI) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            H32[q] = (float)s.box[BOX_H + q];
            HI32[q] = (float)s.box[BOX_HINV + q];
        }
    }
    // FP32 list-band edges of this lane's type against each type, held in registers
    float tlo[kMaxTypes], thi[kMaxTypes];
    double reach2 = -1.0;  // no pair beyond this r^2 can enter the list
#pragma unroll
    for (int t = 0; t < kMaxTypes; ++t) {
        const PairParam& p = ppl[MULTI ? ti * nt + min(t, nt - 1) : 0];
        tlo[t] = p.l_lo32;
        thi[t] = p.l_hi32;
        reach2 = fmax(reach2, p.l_hi);
    }
    // Row windows.  Tile-relative staging of an orthorhombic box keeps each 3-cell row
    // sorted by x, so a row shrinks to the x window that can hold a partner: half-width
    // sqrt(reach^2 - dy^2 - dz^2), with dy, dz the atom's distances to the row's cells (0 in
    // its own y or z layer).  Slack of 1e-9 cell widths absorbs binning and rounding at the
    // cell faces; the window widens by twice the FP32 error of the staged x.
    const int nrows = s.mode == MODE_WHOLE ? 1 : 9;
    uint32_t key[9];  // (window length << 16) | first slot; 0 for an empty row
    {
        const double slack = 1e-9 * fmax(s.cellw[0], fmax(s.cellw[1], s.cellw[2]));
        const double ylo = (ly - 0.5 * g.sy) * s.cellw[1], zlo = (lz - 0.5 * g.sz) * s.cellw[2];
        const double gym = fmax(0.0, (qy - ylo) - slack), gyp = fmax(0.0, (ylo + s.cellw[1] - qy) - slack);
        const double gzm = fmax(0.0, (qz - zlo) - slack), gzp = fmax(0.0, (zlo + s.cellw[2] - qz) - slack);
        const int hx = g.sx + 2, hxy = (g.sx + 2) * (g.sy + 2);
        // window table: face of the first cell of the lane's rows, slices per unit x
        const float cw32 = (float)s.cellw[0], inv_h = (float)kWinSlices / cw32;
        const float f0 = ((float)(e0 % hx) - 2.0f - 0.5f * (float)g.sx) * cw32;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            int rb = 0, re = 0;
            if (active && r < nrows) {
                if (s.mode == MODE_WHOLE) {
                    re = total;
                } else {
                    const int oy = (r % 3) - 1, oz = (r / 3) - 1;
                    const int e = e0 + oy * hx + oz * hxy;
                    rb = pref[e - 1];
                    re = pref[e + 2];
                    if (!PAIRMI && !TRI) {  // a tilted box's rows overlap in x: no window
                        const double gy = oy < 0 ? gym : (oy > 0 ? gyp : 0.0);
                        const double gz = oz < 0 ? gzm : (oz > 0 ? gzp : 0.0);
                        const double w2 = reach2 - gy * gy - gz * gz;
                        if (w2 < 0.0) {
                            re = rb;
                        } else {
                            const double w = sqrt(w2) * (1.0 + 1e-9) + s.wslack;
                            if (PIS_BUILD_WINTAB) {
                                // slices of the row's first cell (face f0): the bounds move out
                                // by 1 % of a slice, far beyond the FP32 rounding of x, w, f0
                                const float w32 = (float)w * (1.0f + 1e-6f);
                                const int gs = (int)floorf((qx32 - w32 - f0) * inv_h - 0.01f);
                                const int ge = (int)floorf((qx32 + w32 - f0) * inv_h + 0.01f) + 1;
                                const int t0 = (e - 1) * kWinSlices;
                                rb = wtab[t0 + min(max(gs, 0), 3 * kWinSlices)];
                                re = wtab[t0 + min(max(ge, 0), 3 * kWinSlices)];
                            } else {
                                const int wb = search_sorted<false>(Px, rb, re, qx - w);
                                re = search_sorted<true>(Px, wb, re, qx + w);
                                rb = wb;
                            }
                        }
                    }
                }
            }
            key[r] = re > rb ? ((uint32_t)(re - rb) << 16) | (uint32_t)rb : 0u;
        }
    }
    if (nrows > 1) sort9_desc(key);
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // 8 list entries of 16 bits, w0 oldest
    const uint32_t lsh = (uint32_t)s.list_shift;
    int cnt = 0;
    const int maxg = s.maxnb >> 3;
    const size_t npad = s.npad;
    __attribute__((address_space(1))) u32x4* out = gbl(reinterpret_cast<u32x4*>(s.nbr + a));
    // One list entry: the LDS byte offset of slot sl's coordinate in the force kernel's
    // staging (slot << list_shift: 8-byte FP64 or 4-byte FP32 slots), with the partner's
    // type index in the low bits (mixtures), shifted into the register of the last 8; a
    // full register leaves as one uint4 list group.
    auto append = [&](uint32_t sl) {
        w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
        w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
        w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
        w3 = __builtin_amdgcn_alignbit((sl << lsh) | (MULTI ? (uint32_t)Pt[sl] : 0u), w3, 16);
        ++cnt;
        if ((cnt & 7) == 0 && (cnt >> 3) <= maxg) out[(size_t)((cnt >> 3) - 1) * npad] = u32x4{w0, w1, w2, w3};
    };
    for (int r = 0; r < nrows; ++r) {
        const int rb = (int)(key[0] & 0xffffu), re = rb + (int)(key[0] >> 16);
#pragma unroll
        for (int q = 0; q < 8; ++q) key[q] = key[q + 1];
        key[8] = 0u;
        // uniform over the wave by construction; readfirstlane makes the loops scalar
        const int rowmax = wave_max(re - rb);
        for (int cb = 0; cb < rowmax; cb += 64) {
            constexpr int U = PIS_BUILD_UNROLL;
            // trips of U candidates up to the wave's longest window (<= 64); a lane's bits
            // past its own window, and its own atom, are cleared once per chunk below.  A
            // lane whose window is used up reads from slot 0; the others read at most 63
            // slots past their chunk start, inside the staging's padding.
            const int kend = (min(64, rowmax - cb) + U - 1) & ~(U - 1);
            const int sb = rb + cb;
            const int sr = sb < re ? sb : 0;
            lds_cf32* px = lds_opaque(Px + sr);
            lds_cf32* py = lds_opaque(Py + sr);
            lds_cf32* pz = lds_opaque(Pz + sr);
            uint32_t mlo = 0, mhi = 0, blo_m = 0, bhi_m = 0;
            // two halves of 32 candidates: bit positions are scalar, no branch per candidate
            for (int half = 0; half < 2; ++half) {
                const int k0 = 32 * half, k1 = min(kend, k0 + 32);
                uint32_t mi = 0, mm = 0;  // certainly within / within the band's upper edge
                for (int k = k0; k < k1; k += U) {
                    float dx[U], dy[U], dz[U];
                    int tt[U];
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        dx[j] = px[k + j] - qx32;
                        dy[j] = py[k + j] - qy32;
                        dz[j] = pz[k + j] - qz32;
                        tt[j] = MULTI ? (int)Pt[sr + k + j] : 0;
                    }
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        if (PAIRMI && TRI) {
                            mi_tri32(H32, HI32, dx[j], dy[j], dz[j]);
                        } else if (PAIRMI) {
                            dx[j] = mi_fast32(dx[j], L32[0], Li32[0]);
                            dy[j] = mi_fast32(dy[j], L32[1], Li32[1]);
                            dz[j] = mi_fast32(dz[j], L32[2], Li32[2]);
                        }
                        const float r2 = fmaf(dx[j], dx[j], fmaf(dy[j], dy[j], dz[j] * dz[j]));
                        float blo = tlo[0], bhi = thi[0];
                        if (NT == 2) {  // one select per edge: the partner is type 0 or 1
                            blo = tt[j] != 0 ? tlo[1] : blo;
                            bhi = tt[j] != 0 ? thi[1] : bhi;
                        } else if (MULTI) {
                            blo = tt[j] == 1 ? tlo[1] : blo;
                            bhi = tt[j] == 1 ? thi[1] : bhi;
                            blo = tt[j] == 2 ? tlo[2] : blo;
                            bhi = tt[j] == 2 ? thi[2] : bhi;
                            blo = tt[j] == 3 ? tlo[3] : blo;
                            bhi = tt[j] == 3 ? thi[3] : bhi;
                        }
                        if (PIS_BUILD_SHIFTIN) {  // candidate k0 + i ends at bit n - 1 - i
                            shift_in_le(mi, r2, blo);
                            shift_in_le(mm, r2, bhi);
                        } else {
                            mi |= (r2 <= blo ? 1u : 0u) << (k - k0 + j);
                            mm |= (r2 <= bhi ? 1u : 0u) << (k - k0 + j);
                        }
                    }
                }
                if (PIS_BUILD_SHIFTIN) {  // back to candidate k0 + i at bit i
                    const int n = k1 - k0;
                    mi = n > 0 ? __builtin_bitreverse32(mi) >> (32 - n) : 0u;
                    mm = n > 0 ? __builtin_bitreverse32(mm) >> (32 - n) : 0u;
                }
                if (half == 0) {
                    mlo = mi;
                    blo_m = mm & ~mi;
                } else {
                    mhi = mi;
                    bhi_m = mm & ~mi;
                }
            }
            {
                // keep [sb, re) without the atom itself
                const int nv = max(0, min(64, re - sb));
                const uint32_t vlo = nv >= 32 ? 0xffffffffu : ((1u << nv) - 1u);
                const uint32_t vhi = nv >= 64 ? 0xffffffffu : (nv <= 32 ? 0u : ((1u << (nv - 32)) - 1u));
                uint32_t slo = 0xffffffffu, shi = 0xffffffffu;
                const int ds = self - sb;
                if (ds >= 0 && ds < 32) slo = ~(1u << ds);
                if (ds >= 32 && ds < 64) shi = ~(1u << (ds - 32));
                mlo &= vlo & slo;
                mhi &= vhi & shi;
                blo_m &= vlo & slo;
                bhi_m &= vhi & shi;
            }
            if (__any((blo_m | bhi_m) != 0u)) {
                // rare: settle the band candidates with the reference predicate
                for (int half = 0; half < 2; ++half) {
                    uint32_t bm = half ? bhi_m : blo_m;
                    while (bm) {
                        const int kk = __builtin_ctz(bm) + 32 * half;
                        bm &= bm - 1u;
                        const int sl = sb + kk;
                        const PairParam& p = ppl[MULTI ? ti * nt + Pt[sl] : 0];
                        if (exact_within<TRI>(s, st, slot_atom(pref, ecell, g.ne, sl), xi, yi, zi, p.r2_list)) {
                            if (kk < 32)
                                mlo |= 1u << kk;
                            else
                                mhi |= 1u << (kk - 32);
                        }
                    }
                }
            }
            // accepted slots, ascending, into the 8-entry shift register (w0 oldest)
            for (int half = 0; half < 2; ++half) {
                uint32_t m = half ? mhi : mlo;
                const uint32_t hbase = (uint32_t)(sb + 32 * half);
                while (m) {
                    const uint32_t b = (uint32_t)__builtin_ctz(m);
                    m &= m - 1u;
                    append(hbase + b);
                }
            }
        }
    }
    if (!active) return 0;
    const int n_real = cnt;
    if (cnt & 7) {
        while (cnt & 7) {
            w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
            w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
            w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
            w3 = __builtin_amdgcn_alignbit((uint32_t)SCAP << lsh, w3, 16);
            ++cnt;
        }
        if ((cnt >> 3) <= maxg) out[(size_t)((cnt >> 3) - 1) * npad] = u32x4{w0, w1, w2, w3};
    }
    gbl(s.nnb)[a] = (uint16_t)min(n_real, s.maxnb);
    return n_real;
}

// The kind of a system's tiles (KIND_*): box shape and staging mode.
__device__ __forceinline__ int system_kind(const SysDev& s)
{
    const bool rel = s.mode == MODE_TILE_REL;
    return !s.tri ? (rel ? KIND_ORTHO : KIND_ORTHO_RAW) : (rel ? KIND_TRI : KIND_TRI_RAW);
}

// One instantiation per kind (KIND_*), so the common tile-relative kinds carry neither the
// per-pair image of raw staging nor its registers.
template <int SCAP, int KIND>
__global__ __launch_bounds__(kTileThreads, PIS_BUILD_MIN_WAVES) void k_build_list(const SysDev* __restrict__ sys,
                                                              const WorkItem* __restrict__ tiles,
                                                              int ntw)
{
    constexpr bool TRI = KIND == KIND_TRI || KIND == KIND_TRI_RAW;
    constexpr bool REL = KIND == KIND_ORTHO || KIND == KIND_TRI;
    // FP32 staging: half the LDS of the force kernel's, so more tiles share a CU.  The 64
    // slots past the padding slot take the candidate reads that run past a window's end.
    __shared__ float Px[SCAP + 1 + 64], Py[SCAP + 1 + 64], Pz[SCAP + 1 + 64];
    __shared__ uint8_t Pt[SCAP + 1 + 64];
    __shared__ uint8_t smap[SCAP + 1];
    __shared__ int pref[kMaxHalo + 1];
    __shared__ int ecell[kMaxHalo];
    __shared__ int stmp[16];
    __shared__ PairParam ppl[kMaxTypes * kMaxTypes];
    __shared__ uint16_t wtab[(kMaxHalo + 1) * kWinSlices];  // row-window table (orthorhombic, tile-relative)
    __shared__ int ownf[kMaxHalo];  // first atom of each own cell
    const int lane = threadIdx.x & 63;
    if (!block_rebuild_pending<true>(sys, tiles, ntw)) return;
    for (int b = blockIdx.x; b < ntw; b += gridDim.x) {
        const WorkItem w = tiles[xcd_remap(b, ntw)];
        const SysDev& s = sys[w.sys];
        if (!s.ctl[CTL_REBUILD]) continue;
        if (system_kind(s) != KIND) continue;  // another instantiation builds this tile
        const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR]) ^ 1];
        const TileGeom g = tile_geom(s, w.idx);
        const int nt = s.ntypes;
        double cx, cy, cz;
        tile_center(s, g, cx, cy, cz);
        for (int q = threadIdx.x; q < nt * nt; q += blockDim.x) ppl[q] = s.pp[q];
        const int total = stage_tile<SCAP, REL ? STAGE_REL : STAGE_RAW>(s, st, g, cx, cy, cz, Px, Py, Pz, Pt, pref,
                                                                          ecell, stmp, smap);
        if (threadIdx.x == 0) atomicMax(&s.ctl[CTL_MAXSTAGE], total);
        if (total > SCAP) {
            if (threadIdx.x == 0) atomicOr(&s.ctl[CTL_OVF_STAGE], 1);
            __syncthreads();
            continue;
        }
        if (s.mode != MODE_WHOLE) {
            for (int k = threadIdx.x; k < g.sx * g.sy * g.sz; k += blockDim.x) {
                const int lx = k % g.sx, ly = (k / g.sx) % g.sy, lz = k / (g.sx * g.sy);
                ownf[k] = ecell[((lz + 1) * (g.sy + 2) + (ly + 1)) * (g.sx + 2) + (lx + 1)];
            }
        }
        if (PIS_BUILD_WINTAB && KIND == KIND_ORTHO) {
            // the row-window table, one entry per thread: a binary search of the cell's
            // x-sorted slots for the first whose slice (same FP32 faces as the lookups) is
            // >= k.  A search writes only its own entry, so the table is the same whatever
            // the order of the threads, even in a cell whose staged x is not monotone (an
            // atom at the image boundary of a small box staged on the far side, never a
            // partner of the tile's own atoms).
            const int hx = g.sx + 2;
            const float cw32 = (float)s.cellw[0], inv_h = (float)kWinSlices / cw32;
            for (int q = threadIdx.x; q < (g.ne + 1) * kWinSlices; q += blockDim.x) {
                const int e = q / kWinSlices, k = q % kWinSlices;
                int lo = e < g.ne ? pref[e] : total, hi = e < g.ne ? pref[e + 1] : total;
                if (k > 0) {
                    const float face = ((float)(e % hx) - 1.0f - 0.5f * (float)g.sx) * cw32;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (win_slice(Px[mid], face, inv_h) < k)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                }
                wtab[q] = (uint16_t)lo;
            }
            __syncthreads();
        } else {
            __syncthreads();
        }
        const int a0 = s.cell_start[g.c0], a1 = s.cell_start[g.c1];
        int mymax = 0;
        // whole waves iterate together (inactive lanes ride along) so the sweeps stay uniform
        const int span = ((a1 - a0) + 63) & ~63;
        for (int base = a0 + (int)(threadIdx.x & ~63); base < a0 + span; base += blockDim.x) {
            const int a = base + lane;
            const bool active = a < a1;
            // REL: tile-relative staging (plain differences); otherwise raw positions and the
            // per-pair image of the box
            const int n =
                nt == 1   ? build_atom<SCAP, !REL, 1, TRI>(s, st, g, Px, Py, Pz, Pt, pref, ecell, ownf, ppl, wtab,
                                                           total, active, a, cx, cy, cz)
                : nt == 2 ? build_atom<SCAP, !REL, 2, TRI>(s, st, g, Px, Py, Pz, Pt, pref, ecell, ownf, ppl, wtab,
                                                           total, active, a, cx, cy, cz)
                          : build_atom<SCAP, !REL, kMaxTypes, TRI>(s, st, g, Px, Py, Pz, Pt, pref, ecell, ownf,
                                                                   ppl, wtab, total, active, a, cx, cy, cz);
            mymax = max(mymax, n);
        }
        mymax = wave_max(mymax);
        if (lane == 0 && mymax > 0) {
            atomicMax(&s.ctl[CTL_MAXNB], mymax);
            if (mymax > s.maxnb) atomicOr(&s.ctl[CTL_OVF_NB], 1);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// pair force

struct Acc {
    double fx, fy, fz, pe, wp;
};

// LennardJones::compute_potential (lennard_jones.rs:32-55) for one candidate, without a
// branch: `in` selects whether the pair contributes.  FP64: a refined hardware reciprocal
// for 1/r2; a candidate outside the cutoff computes with its own finite r2 and is dropped
// by the selects (raw staging masks padding entries by count, tile-relative staging puts
// the padding slot kPadFar away, so its r2 is finite in FP64 and fails the cutoff).  MIXED
// with FP64 staging (raw staging only): r2 and the cutoff decision stay FP64, the LJ
// arithmetic is FP32, and a candidate outside the cutoff enters it as r2 = +inf, which
// v_rcp_f32 turns into +0 (force factor -0; the energy shift is masked to 0).  Results
// accumulate in FP64.
template <bool MIXED>
__device__ __forceinline__ void lj_accum(const PairParam& p, double dx, double dy, double dz, double r2,
                                         bool in, Acc& acc)
{
    if (!MIXED) {
        double inv = __builtin_amdgcn_rcp(r2);
        inv = fma(inv, fma(-r2, inv, 1.0), inv);
        inv = fma(inv, fma(-r2, inv, 1.0), inv);
        const double sr2 = p.sig2 * inv;
        const double sr6 = sr2 * sr2 * sr2;
        const double fp = in ? fma(p.c48, sr6, -p.c24) * sr6 * inv : 0.0;
        const double u = in ? fma(fma(p.c4, sr6, -p.c4), sr6, -p.ucut) : 0.0;
        acc.fx = fma(-fp, dx, acc.fx);
        acc.fy = fma(-fp, dy, acc.fy);
        acc.fz = fma(-fp, dz, acc.fz);
        acc.pe += u;
        acc.wp = fma(fp, r2, acc.wp);
    } else {
        // one 32-bit select each for r2 and the shift, ahead of the FP64 conversions,
        // instead of two per 64-bit result
        const float r2f = in ? (float)r2 : __builtin_inff();
        const float inv = __builtin_amdgcn_rcpf(r2f);
        const float sr2 = p.sig2f * inv;
        const float sr6 = sr2 * sr2 * sr2;
        const float fpf = fmaf(p.c48f, sr6, -p.c24f) * sr6 * inv;
        const float ucf = in ? p.ucutf : 0.0f;
        const float uf = fmaf(fmaf(p.c4f, sr6, -p.c4f), sr6, -ucf);
        const double fp = (double)fpf;
        acc.fx = fma(-fp, dx, acc.fx);
        acc.fy = fma(-fp, dy, acc.fy);
        acc.fz = fma(-fp, dz, acc.fz);
        acc.pe += (double)uf;
        acc.wp = fma(fp, r2, acc.wp);
    }
}

struct ExactCtx {
    const SysDev* s;
    const State* st;
    const int* pref;
    const int* ecell;
    int ne;
    double xi, yi, zi;
};

typedef __attribute__((address_space(1))) const u32x4 gu32x4;

template <int SCAP, bool MIXED, bool MULTI, bool PAIRMI, bool TRI>
__device__ __forceinline__ void atom_forces(const SysDev& s, const double* Px, const double* Py,
                                            const double* Pz,
                                            const PairParam* ppl, const PairParam& p0, int a, int ti,
                                            double qx, double qy, double qz, const ExactCtx& ex,
                                            Acc& acc)
{
    const int nn = gbl(s.nnb)[a];
    const int ng = (nn + 7) >> 3;
    const gu32x4* L = (const gu32x4*)(s.nbr + a);
    const char* bx = reinterpret_cast<const char*>(Px);
    const char* by = reinterpret_cast<const char*>(Py);
    const char* bz = reinterpret_cast<const char*>(Pz);
    u32x4 un = {0u, 0u, 0u, 0u};
    if (ng > 0) un = L[0];
    for (int gi = 0; gi < ng; ++gi) {
        const u32x4 u = un;
        if (gi + 1 < ng) un = L[(size_t)(gi + 1) * s.npad];
        const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
        const int rem = nn - gi * 8;
#pragma unroll
        for (int h = 0; h < 8; h += PIS_FORCE_BATCH) {
            constexpr int B = PIS_FORCE_BATCH;
            uint32_t off[B];
            double X[B], Y[B], Z[B];
#pragma unroll
            for (int e = 0; e < B; ++e) {
                const int k = h + e;
                off[e] = (wv[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
                const uint32_t ad = MULTI ? off[e] & ~7u : off[e];  // low bits: partner type
                X[e] = *reinterpret_cast<const double*>(bx + ad);
                Y[e] = *reinterpret_cast<const double*>(by + ad);
                Z[e] = *reinterpret_cast<const double*>(bz + ad);
            }
            double R2[B];
            bool IN[B];
            bool band = false;
#pragma unroll
            for (int e = 0; e < B; ++e) {
                double dx = X[e] - qx, dy = Y[e] - qy, dz = Z[e] - qz;
                if (PAIRMI && TRI) {
                    mi_tri_fast(s.box, dx, dy, dz);
                } else if (PAIRMI) {
                    dx = mi_fast(dx, s.L[0], s.Linv[0]);
                    dy = mi_fast(dy, s.L[1], s.Linv[1]);
                    dz = mi_fast(dz, s.L[2], s.Linv[2]);
                }
                X[e] = dx;
                Y[e] = dy;
                Z[e] = dz;
                R2[e] = fma(dx, dx, fma(dy, dy, dz * dz));
                const PairParam& p = MULTI ? ppl[ti * s.ntypes + (off[e] & 7u)] : p0;
                // tile-relative staging: padding entries fail the cutoff by distance
                const bool valid = PAIRMI ? h + e < rem : true;
                if (MIXED) {
                    IN[e] = valid && R2[e] <= p.r2_force;
                } else {
                    IN[e] = valid && R2[e] <= p.f_lo;
                    band |= valid && !IN[e] && R2[e] <= p.f_hi;
                }
            }
            if (!MIXED && band) {
                // the reference's `norm() > rcut` settled exactly inside the rounding band
#pragma unroll
                for (int e = 0; e < B; ++e) {
                    const PairParam& p = MULTI ? ppl[ti * s.ntypes + (off[e] & 7u)] : p0;
                    if ((!PAIRMI || h + e < rem) && !IN[e] && R2[e] <= p.f_hi)
                        IN[e] = exact_within<TRI>(*ex.s, *ex.st,
                                             slot_atom(ex.pref, ex.ecell, ex.ne, (int)(off[e] >> 3)), ex.xi,
                                             ex.yi, ex.zi, p.r2_force);
                }
            }
#pragma unroll
            for (int e = 0; e < B; ++e) {
                const PairParam& p = MULTI ? ppl[ti * s.ntypes + (off[e] & 7u)] : p0;
                lj_accum<MIXED>(p, X[e], Y[e], Z[e], R2[e], IN[e], acc);
            }
        }
    }
}

// FP32 parameters of one type pair for the mixed pair loop with FP32 staging (32 bytes:
// two LDS reads per list entry in mixtures).
// The cutoff of the single-species FP32 loop in three VALU instructions: the compare lands
// in an SGPR pair that selects r^2 or +inf for the reciprocal and is also the carry-in that
// counts the pair into n (n += r2 <= b); written out, the count would take a select and an
// add.  v_cmp_ge_f32(b, r2) is false for NaN.
__device__ __forceinline__ float cut_and_count(float r2, float b, float inf, uint32_t& n)
{
    uint64_t m, c;
    float sel;
    asm("v_cmp_ge_f32_e64 %[m], %[b], %[r2]\n\t"
        "v_cndmask_b32_e64 %[sel], %[inf], %[r2], %[m]\n\t"
        "v_addc_co_u32_e64 %[n], %[c], 0, %[n], %[m]"
        : [m] "=&s"(m), [sel] "=&v"(sel), [n] "+v"(n), [c] "=&s"(c)
        : [b] "v"(b), [r2] "v"(r2), [inf] "v"(inf));
    return sel;
}

struct PairF {
    float r2, sig2, c4, c24, c48, ucut, r2s, pad1;
};

__device__ __forceinline__ PairF pair_f32(const PairParam& p)
{
    return PairF{p.r2f, p.sig2f, p.c4f, p.c24f, p.c48f, p.ucutf, p.r2s32, 0.0f};
}

// The mixed pair loop over one atom's list with FP32 tile-relative staging (SoA Px, Py, Pz
// of SCAP + 1 floats, contiguous): displacement, r^2, cutoff and LJ arithmetic in FP32
// (LennardJones::compute_potential, lennard_jones.rs:32-55, in FP32), the 8 entries of a
// list group summed in FP32 and each group's sums added into the FP64 accumulators.
// Entries are byte offsets of a float slot with the partner type in the two low bits; the
// padding slot lies kPadFar away, so padding entries have r^2 = +inf and fail the cutoff.
// A candidate outside the cutoff enters the reciprocal as +inf (v_rcp_f32 gives +0), so
// its force factor and virial term are -0; the energy shift is masked.
template <int SCAP, bool MULTI>
__device__ __forceinline__ void atom_forces_f32(const SysDev& s, const float* Px, const PairF* prow,
                                                const PairF& p0, int a, float qx, float qy, float qz,
                                                Acc& acc)
{
    constexpr uint32_t YO = (SCAP + 1) * 4, ZO = 2 * (SCAP + 1) * 4;
    const int nn = gbl(s.nnb)[a];
    const int ng = (nn + 7) >> 3;
    const gu32x4* L = (const gu32x4*)(s.nbr + a);
    const char* bx = reinterpret_cast<const char*>(Px);
    u32x4 un = {0u, 0u, 0u, 0u};
    if (ng > 0) un = L[0];
    uint32_t nin = 0;  // single species: pairs within the cutoff
    for (int gi = 0; gi < ng; ++gi) {
        const u32x4 u = un;
        if (gi + 1 < ng) un = L[(size_t)(gi + 1) * s.npad];
        const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
        uint32_t off[8];
        float X[8], Y[8], Z[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            off[e] = (wv[e >> 1] >> ((e & 1) * 16)) & 0xffffu;
            const uint32_t ad = MULTI ? off[e] & ~3u : off[e];
            X[e] = *reinterpret_cast<const float*>(bx + ad);
            Y[e] = *reinterpret_cast<const float*>(bx + ad + YO);
            Z[e] = *reinterpret_cast<const float*>(bx + ad + ZO);
        }
        if constexpr (!MULTI) {
            // one species, coordinates in sigma units: 1/r'^2 is (sigma/r)^2 itself; the
            // energy and the pair virial follow from the sums of (sigma/r)^12 and (sigma/r)^6
            // and the number of pairs in range (combined per atom after the loop), and the
            // force sum is scaled back by 1/sigma there
            float fx = 0.0f, fy = 0.0f, fz = 0.0f, s12 = 0.0f, s6 = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float dx = X[e] - qx, dy = Y[e] - qy, dz = Z[e] - qz;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float sr2 = __builtin_amdgcn_rcpf(cut_and_count(r2, p0.r2s, __builtin_inff(), nin));
                const float sr6 = sr2 * sr2 * sr2;
                const float fp = fmaf(p0.c48, sr6, -p0.c24) * sr6 * sr2;
                fx = fmaf(-fp, dx, fx);
                fy = fmaf(-fp, dy, fy);
                fz = fmaf(-fp, dz, fz);
                s12 = fmaf(sr6, sr6, s12);
                s6 += sr6;
            }
            acc.fx += (double)fx;
            acc.fy += (double)fy;
            acc.fz += (double)fz;
            acc.pe += (double)s12;
            acc.wp += (double)s6;
            continue;
        }
        float fx = 0.0f, fy = 0.0f, fz = 0.0f, pe = 0.0f, wp = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const PairF p = MULTI ? prow[off[e] & 3u] : p0;
            const float dx = X[e] - qx, dy = Y[e] - qy, dz = Z[e] - qz;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const bool in = r2 <= p.r2;
            const float inv = __builtin_amdgcn_rcpf(in ? r2 : __builtin_inff());
            const float sr2 = p.sig2 * inv;
            const float sr6 = sr2 * sr2 * sr2;
            const float w = fmaf(p.c48, sr6, -p.c24) * sr6;  // r . F_ij: the pair virial term
            const float fp = w * inv;
            const float uc = in ? p.ucut : 0.0f;
            const float uu = fmaf(fmaf(p.c4, sr6, -p.c4), sr6, -uc);
            fx = fmaf(-fp, dx, fx);
            fy = fmaf(-fp, dy, fy);
            fz = fmaf(-fp, dz, fz);
            pe += uu;
            wp += w;
        }
        acc.fx += (double)fx;
        acc.fy += (double)fy;
        acc.fz += (double)fz;
        acc.pe += (double)pe;
        acc.wp += (double)wp;
    }
    if constexpr (!MULTI) {
        // U = 4 eps ((s/r)^12 - (s/r)^6) - U(rc) and r . F = 48 eps (s/r)^12 - 24 eps (s/r)^6
        const PairParam& pp = s.pp[0];
        const double s12 = acc.pe, s6 = acc.wp;
        acc.pe = pp.c4 * (s12 - s6) - pp.ucut * (double)nin;
        acc.wp = pp.c48 * s12 - pp.c24 * s6;
        acc.fx *= pp.isig;
        acc.fy *= pp.isig;
        acc.fz *= pp.isig;
    }
}

// LDS of a force workgroup; T is the staged coordinate type.  The FP32 staging reads only
// the FP32 parameter table; from 3072 slots on it also stages without the slot-to-cell map
// (a search of the halo's cell offsets per slot instead), which keeps a workgroup under
// 40 KB so that four share a CU.
template <typename T, int SCAP>
struct ForceLds {
    static constexpr bool F32 = sizeof(T) == 4;
    static constexpr bool MAP = !(F32 && SCAP >= 3072);
    T Px[SCAP + 1], Py[SCAP + 1], Pz[SCAP + 1];
    int pref[kMaxHalo + 1];
    int ecell[kMaxHalo];
    int stmp[16];
    double rtmp[4 * (kTileThreads / 64)];
    PairParam ppl[F32 ? 1 : kMaxTypes * kMaxTypes];
    PairF pf[kMaxTypes * kMaxTypes];
    uint8_t smap[MAP ? SCAP + 1 : 1];
};

template <bool F32>
struct StageType {
    typedef double type;
};
template <>
struct StageType<true> {
    typedef float type;
};


// Registers per lane a force instantiation must fit in (the launch-bounds occupancy target):
// the FP32-staged mixed kernel aims at PIS_FORCE_MIN_WAVES_F32 waves per SIMD.
#ifndef PIS_FORCE_MIN_WAVES_F32
#define PIS_FORCE_MIN_WAVES_F32 4
#endif
template <bool MIXED, int KIND>
struct ForceWaves {
    static constexpr bool F32 = MIXED && (KIND == KIND_ORTHO || KIND == KIND_TRI);
    static constexpr int value = KIND == KIND_TRI_RAW ? 1
                                 : F32                ? PIS_FORCE_MIN_WAVES_F32
                                 : MIXED              ? PIS_FORCE_MIN_WAVES_MIXED
                                                      : PIS_FORCE_MIN_WAVES_FP64;
};

// KIND: one of KIND_ORTHO, KIND_TRI (tile-relative staging), KIND_ORTHO_RAW, KIND_TRI_RAW
// (raw staging with a per-pair image).  Mixed precision with tile-relative staging stages
// FP32 coordinates (atom_forces_f32); every other combination stages FP64.
template <int SCAP, bool MIXED, int KIND>
__global__ __launch_bounds__(kTileThreads, (ForceWaves<MIXED, KIND>::value)) void k_force(
    const SysDev* __restrict__ sys, const WorkItem* __restrict__ tiles, int ntw, int integrate)
{
    constexpr bool REL = KIND == KIND_ORTHO || KIND == KIND_TRI;
    constexpr bool TRI = KIND == KIND_TRI || KIND == KIND_TRI_RAW;
    constexpr bool F32 = MIXED && REL;
    typedef typename StageType<F32>::type T;
    __shared__ ForceLds<T, SCAP> lds;

    const WorkItem w = tiles[xcd_remap(blockIdx.x, ntw)];
    const SysDev& s = sys[w.sys];
    if (system_kind(s) != KIND) return;             // another instantiation takes this tile
    if (integrate && s.ctl[CTL_LEFT] == 0) return;  // frozen system: the whole block leaves
    // read set: the live set, or the freshly sorted one on a rebuild step; a step writes
    // its results into the other set (wt), which becomes live at the end of the step
    // (control words are read with vector loads; readfirstlane makes the block-uniform
    // choices scalar, so the state pointers stay in scalar registers)
    const int cur = __builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR]);
    const int rbf = __builtin_amdgcn_readfirstlane(s.ctl[CTL_REBUILD]);
    const int rset = cur ^ (rbf ? 1 : 0);
    const State st = s.st[rset];
    const State wt = s.st[rset ^ 1];
    const bool rebuilt = rbf != 0;
    // while more steps of this call follow (and no thermostat must act in between), the
    // next step's position update happens here, from the v and F just computed
    const bool ahead = integrate && __builtin_amdgcn_readfirstlane(s.ctl[CTL_REMAIN]) > 1 &&
                       __builtin_amdgcn_readfirstlane(s.ctl[CTL_LEFT]) != 1 && !s.nhc_on;
    const TileGeom g = tile_geom(s, w.idx);
    double cx, cy, cz;
    tile_center(s, g, cx, cy, cz);
    const int nt = s.ntypes;
    for (int q = threadIdx.x; q < nt * nt; q += blockDim.x) {
        if constexpr (!F32) lds.ppl[q] = s.pp[q];
        lds.pf[q] = pair_f32(s.pp[q]);
    }
    // the single-species FP32 loop works in sigma units (an absent pair keeps scale 1)
    const double scale = (F32 && nt == 1 && s.pp[0].r2_force >= 0.0) ? s.pp[0].isig : 1.0;
    const int total = stage_tile<SCAP, REL ? STAGE_REL : STAGE_RAW, T>(
        s, st, g, cx, cy, cz, lds.Px, lds.Py, lds.Pz, (uint8_t*)nullptr, lds.pref, lds.ecell, lds.stmp,
        ForceLds<T, SCAP>::MAP ? lds.smap : (uint8_t*)nullptr, scale);
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    if (total <= SCAP) {
        const PairParam p0 = s.pp[0];  // uniform: scalar loads, no vector registers
        const PairF p0f = pair_f32(p0);
        const int a0 = gbl(s.cell_start)[g.c0], a1 = gbl(s.cell_start)[g.c1];
        for (int a = a0 + (int)threadIdx.x; a < a1; a += blockDim.x) {
            const double xi = gbl(st.x)[a], yi = gbl(st.y)[a], zi = gbl(st.z)[a];
            const int ti = gbl(st.type)[a];
            double qx = xi, qy = yi, qz = zi;
            if (REL) rel_to_center(s, cx, cy, cz, qx, qy, qz);
            Acc acc = {0.0, 0.0, 0.0, 0.0, 0.0};
            if constexpr (F32) {
                const float* Px = reinterpret_cast<const float*>(lds.Px);
                if (nt == 1)
                    atom_forces_f32<SCAP, false>(s, Px, lds.pf, p0f, a, (float)(qx * scale), (float)(qy * scale),
                                                 (float)(qz * scale), acc);
                else
                    atom_forces_f32<SCAP, true>(s, Px, lds.pf + ti * nt, p0f, a, (float)qx, (float)qy, (float)qz,
                                                acc);
            } else {
                const double* Px = reinterpret_cast<const double*>(lds.Px);
                const double* Py = reinterpret_cast<const double*>(lds.Py);
                const double* Pz = reinterpret_cast<const double*>(lds.Pz);
                const ExactCtx ex = {&s, &st, lds.pref, lds.ecell, g.ne, xi, yi, zi};
                if (nt == 1)
                    atom_forces<SCAP, MIXED, false, !REL, TRI>(s, Px, Py, Pz, lds.ppl, p0, a, ti, qx, qy, qz, ex,
                                                               acc);
                else
                    atom_forces<SCAP, MIXED, true, !REL, TRI>(s, Px, Py, Pz, lds.ppl, p0, a, ti, qx, qy, qz, ex,
                                                              acc);
            }
            const double m = s.mass[ti];
            // positions again from memory (L2): not held in registers across the pair loop
            const double xa = gbl(st.x)[a], ya = gbl(st.y)[a], za = gbl(st.z)[a];
            double vx = gbl(st.vx)[a], vy = gbl(st.vy)[a], vz = gbl(st.vz)[a];
            if (integrate) {
                // v += ((a_t + a_{t+dt}) 0.5) dt  (potential.rs:28-30)
                const double dt = s.dt;
                vx = vx + ((gbl(st.fx)[a] / m + acc.fx / m) * 0.5) * dt;
                vy = vy + ((gbl(st.fy)[a] / m + acc.fy / m) * 0.5) * dt;
                vz = vz + ((gbl(st.fz)[a] / m + acc.fz / m) * 0.5) * dt;
                if (s.nhc_on) {  // NVT: the second scaling (potential.rs:49)
                    const double sc = s.nhc[NHC_SCALE];
                    vx = vx * sc;
                    vy = vy * sc;
                    vz = vz * sc;
                }
                gbl(wt.vx)[a] = vx;
                gbl(wt.vy)[a] = vy;
                gbl(wt.vz)[a] = vz;
                gbl(wt.fx)[a] = acc.fx;
                gbl(wt.fy)[a] = acc.fy;
                gbl(wt.fz)[a] = acc.fz;
                if (rebuilt) {  // the other set still has the order before the sort
                    gbl(wt.type)[a] = (uint8_t)ti;
                    gbl(wt.orig)[a] = gbl(st.orig)[a];
                }
                if (ahead) {
                    // the next step's x += v dt + (F/m)/2 dt^2, wrap and skin check, as
                    // k_integrate_pos does them (potential.rs:15-22, simulation_box.rs:29-42)
                    const double dt2 = s.dt2;
                    double x = xa + (vx * dt + (acc.fx / m) * 0.5 * dt2);
                    double y = ya + (vy * dt + (acc.fy / m) * 0.5 * dt2);
                    double z = za + (vz * dt + (acc.fz / m) * 0.5 * dt2);
                    double ddx, ddy, ddz;
                    if (TRI) {
                        wrap_tri(s.box, x, y, z);
                        ddx = x - gbl(s.rx)[a];
                        ddy = y - gbl(s.ry)[a];
                        ddz = z - gbl(s.rz)[a];
                        mi_tri_fast(s.box, ddx, ddy, ddz);
                    } else {
                        double sx = x * s.Linv[0], sy = y * s.Linv[1], sz = z * s.Linv[2];
                        sx = sx - floor(sx);
                        sy = sy - floor(sy);
                        sz = sz - floor(sz);
                        x = s.L[0] * sx;
                        y = s.L[1] * sy;
                        z = s.L[2] * sz;
                        ddx = mi_fast(x - gbl(s.rx)[a], s.L[0], s.Linv[0]);
                        ddy = mi_fast(y - gbl(s.ry)[a], s.L[1], s.Linv[1]);
                        ddz = mi_fast(z - gbl(s.rz)[a], s.L[2], s.Linv[2]);
                    }
                    gbl(wt.x)[a] = x;
                    gbl(wt.y)[a] = y;
                    gbl(wt.z)[a] = z;
                    const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (!(d2 <= s.half_skin2)) {  // also catches NaN
                        atomicOr(&s.ctl[CTL_REBUILD_NEXT], 1);
                        if (d2 != d2) atomicOr(&s.ctl[CTL_NAN], 1);
                    }
                } else {
                    gbl(wt.x)[a] = xa;
                    gbl(wt.y)[a] = ya;
                    gbl(wt.z)[a] = za;
                }
            } else {
                gbl(st.fx)[a] = acc.fx;
                gbl(st.fy)[a] = acc.fy;
                gbl(st.fz)[a] = acc.fz;
            }
            sum[0] += acc.pe;
            sum[1] += 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
            sum[2] += (xa * acc.fx + ya * acc.fy) + za * acc.fz;
            sum[3] += acc.wp;
        }
    } else if (threadIdx.x == 0) {
        atomicOr(&s.ctl[CTL_OVF_STAGE], 1);
    }
    block_sum4(sum, lds.rtmp);
    if (threadIdx.x == 0) {
        double* P = s.partial + 4 * (size_t)w.idx;
        P[0] = sum[0];
        P[1] = sum[1];
        P[2] = sum[2];
        P[3] = sum[3];
    }
}

// Per-system fixed-order sum of the tile partials -> thermo record; optionally flips the
// live buffer set after a rebuild.
// Nose-Hoover chain arithmetic (ensemble/nvt.rs:59-90), one thread, op for op.
__device__ void nhc_forces(double* c, double ke, int64_t n)
{
    const int nc = (int)c[NHC_CHAIN];
    const double kb = c[NHC_KB], T = c[NHC_T];
    c[NHC_G] = 2.0 * ke - ((double)(n * 3)) * kb * T;
    for (int j = 1; j < nc; ++j) c[NHC_G + j] = c[NHC_Q + j - 1] * (c[NHC_XI + j - 1] * c[NHC_XI + j - 1]) - kb * T;
}

__device__ void nhc_half_step(double* c, double dt)
{
    const int j = (int)c[NHC_CHAIN] - 1;
    c[NHC_XI + j] = 0.5 * dt * c[NHC_G + j] / c[NHC_Q + j];
    for (int l = j - 1; l >= 0; --l)
        c[NHC_XI + l] = (0.5 * dt * c[NHC_G + l] / c[NHC_Q + l]) * exp(-0.25 * dt * c[NHC_XI + l + 1]);
}

__device__ double nhc_kinetic(const double* c)
{
    double k = 0.0;
    for (int i = 0; i < (int)c[NHC_CHAIN]; ++i) k += 0.5 * c[NHC_Q + i] * (c[NHC_XI + i] * c[NHC_XI + i]);
    return k;
}

// Sum of a system's per-tile partials (fixed order; identical to k_reduce's sums).
__device__ __forceinline__ void sum_partials(const SysDev& s, double* v, double* rtmp)
{
    v[0] = v[1] = v[2] = v[3] = 0.0;
    const __attribute__((address_space(1))) double* P = gbl(static_cast<const double*>(s.partial));
    // four tiles' partials in flight per thread (the loop is latency-bound otherwise)
    const int bd = blockDim.x;
    int t = threadIdx.x;
    for (; t + 3 * bd < s.ntile; t += 4 * bd) {
        double q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) q[u][k] = P[4 * (t + u * bd) + k];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] += q[u][k];
    }
    for (; t < s.ntile; t += bd)
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] += P[4 * t + k];
    block_sum4(v, rtmp);
}

__device__ void sum_tens(const SysDev& s, double* v, double* tmp);
__device__ double sum_tens_ke(const SysDev& s, double* tmp);

// Start of an NVT step (potential.rs:41-45): KE of the current velocities (the last
// force launch's partials), chain forces and half step, and the velocity scale.  Launched
// on a call's first NVT step and on every NPT step; k_reduce does it for the other NVT
// steps at the end of the step before.
__global__ __launch_bounds__(1024) void k_nhc_pre(const SysDev* __restrict__ sys)
{
    __shared__ double rtmp[64];
    const SysDev& s = sys[blockIdx.x];
    __shared__ double ttmp[1024 + TENS_COUNT];
    if (!s.nhc_on || s.ctl[CTL_LEFT] == 0) return;
    double ke;
    if (s.npt_on) {
        // NPT: the barostat rescaled the velocities since the last force launch; their KE
        // was summed per block right after the scaling
        ke = sum_tens_ke(s, ttmp);
    } else {
        double v[4];
        sum_partials(s, v, rtmp);
        ke = v[1];
    }
    if (threadIdx.x != 0) return;
    double* c = s.nhc;
    nhc_forces(c, ke, s.n);
    nhc_half_step(c, s.dt);
    c[NHC_SCALE] = exp(-0.5 * s.dt * c[NHC_XI]);
}

// Step counter and length of an NVT run (the target ramp's i and total).
__global__ void k_nhc_run(const SysDev* __restrict__ sys, int nsys, double step0, double total)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsys || !sys[k].nhc_on) return;
    sys[k].nhc[NHC_STEP] = step0;
    sys[k].nhc[NHC_TOTAL] = total;
}

// ---------------------------------------------------------------------------------
// MTK barostat (ensemble/npt.rs) around the NVT step (potential.rs:112-135).  3x3
// matrices are column-major; every operation follows the order of the oracle's
// restatement (no library calls), so both produce the same bits from the same inputs.

__device__ void mat3_mul(const double* a, const double* b, double* out)
{
    double t[9];
    for (int c = 0; c < 3; ++c) gemv3(a, b[3 * c], b[3 * c + 1], b[3 * c + 2], t[3 * c], t[3 * c + 1], t[3 * c + 2]);
    for (int k = 0; k < 9; ++k) out[k] = t[k];
}

// exp(A): A halved until its 1-norm is at most 1/2, Taylor series to degree 18, squared
// back (Moler & Van Loan 2003, method 3).
__device__ void mat3_exp(const double* a, double* out)
{
    double norm = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double col = (fabs(a[3 * c]) + fabs(a[3 * c + 1])) + fabs(a[3 * c + 2]);
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
    e[0] = e[4] = e[8] = t[0] = t[4] = t[8] = 1.0;
    for (int j = 1; j <= 18; ++j) {
        mat3_mul(t, x, t);
        for (int k = 0; k < 9; ++k) {
            t[k] = t[k] / (double)j;
            e[k] = e[k] + t[k];
        }
    }
    for (int q = 0; q < sq; ++q) mat3_mul(e, e, e);
    for (int k = 0; k < 9; ++k) out[k] = e[k];
}

// nalgebra 3x3 try_inverse (cofactors, determinant by first-row expansion) and |det|.
__device__ double mat3_det(const double* h)
{
    const double m11 = h[4] * h[8] - h[5] * h[7];
    const double m12 = h[1] * h[8] - h[2] * h[7];
    const double m13 = h[1] * h[5] - h[2] * h[4];
    return h[0] * m11 - h[3] * m12 + h[6] * m13;
}

__device__ void mat3_inverse(const double* h, double* hi)
{
    const double m11 = h[4] * h[8] - h[5] * h[7];
    const double m12 = h[1] * h[8] - h[2] * h[7];
    const double m13 = h[1] * h[5] - h[2] * h[4];
    const double det = h[0] * m11 - h[3] * m12 + h[6] * m13;
    hi[0] = m11 / det;
    hi[3] = -(h[3] * h[8] - h[5] * h[6]) / det;
    hi[6] = (h[3] * h[7] - h[4] * h[6]) / det;
    hi[1] = -m12 / det;
    hi[4] = (h[0] * h[8] - h[2] * h[6]) / det;
    hi[7] = -(h[0] * h[7] - h[1] * h[6]) / det;
    hi[2] = m13 / det;
    hi[5] = -(h[0] * h[5] - h[2] * h[3]) / det;
    hi[8] = (h[0] * h[4] - h[1] * h[3]) / det;
}

// math.rs:39-41: (a + a^T) * 0.5
__device__ void symmetrize3(const double* a, double* out)
{
    double t[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) t[3 * c + r] = (a[3 * c + r] + a[3 * r + c]) * 0.5;
    for (int k = 0; k < 9; ++k) out[k] = t[k];
}

// delta_momentum (npt.rs:44-49): symmetrize((P - P0) * (V * 0.5 * dt)), P = (K + W) / V
__device__ void mtk_add_delta(const double* sums, double* npt, double vol, double dt)
{
    double d[9];
    const double f = vol * 0.5 * dt;
    for (int k = 0; k < 9; ++k) d[k] = ((sums[TENS_K + k] + sums[TENS_W + k]) / vol - npt[NPT_TARGET + k]) * f;
    symmetrize3(d, d);
    for (int k = 0; k < 9; ++k) npt[NPT_MOM + k] += d[k];
}

// MTKBarostat::scale (npt.rs:51-58): exp(symmetrize(momentum / w) * factor * dt)
__device__ void mtk_scale(const double* npt, double dt, double factor, double* out)
{
    double eta[9], a[9];
    for (int k = 0; k < 9; ++k) eta[k] = npt[NPT_MOM + k] / npt[NPT_W];
    symmetrize3(eta, eta);
    for (int k = 0; k < 9; ++k) a[k] = eta[k] * factor * dt;
    mat3_exp(a, out);
}

// Fixed-order sum of the per-block partials of one system (valid in thread 0).
// Sum of the KE column of the partials alone; result valid in thread 0.
__device__ double sum_tens_ke(const SysDev& s, double* tmp /* [blockDim.x] */)
{
    const int t = threadIdx.x;
    const __attribute__((address_space(1))) double* T = gbl(static_cast<const double*>(s.tens));
    double a0 = 0.0, a1 = 0.0;
    int b = t;
    for (; b + (int)blockDim.x < s.nblk; b += 2 * (int)blockDim.x) {
        a0 += T[(size_t)b * TENS_COUNT + TENS_KE];
        a1 += T[(size_t)(b + blockDim.x) * TENS_COUNT + TENS_KE];
    }
    if (b < s.nblk) a0 += T[(size_t)b * TENS_COUNT + TENS_KE];
    tmp[t] = a0 + a1;
    __syncthreads();
    for (int w = (int)blockDim.x >> 1; w > 0; w >>= 1) {
        if (t < w) tmp[t] += tmp[t + w];
        __syncthreads();
    }
    return tmp[0];
}

__device__ void sum_tens(const SysDev& s, double* v, double* tmp /* [blockDim.x + TENS_COUNT] */)
{
    // the partials read as one flat array, coalesced: with a stride that is a multiple of
    // TENS_COUNT every thread keeps one component; a fixed order of additions
    const int stride = ((int)blockDim.x / TENS_COUNT) * TENS_COUNT;
    const int t = threadIdx.x;
    const __attribute__((address_space(1))) double* T = gbl(static_cast<const double*>(s.tens));
    const int total = s.nblk * TENS_COUNT;
    // eight independent loads in flight per thread (the loop is latency-bound otherwise)
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < stride) {
        int i = t;
        for (; i + 7 * stride < total; i += 8 * stride)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] += T[i + u * stride];
        for (int u = 0; i < total; i += stride, ++u) acc[u] += T[i];
    }
    tmp[t] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (t < TENS_COUNT) {
        double x = 0.0;
        for (int r = t; r < stride; r += TENS_COUNT) x += tmp[r];
        tmp[blockDim.x + t] = x;
    }
    __syncthreads();
    if (t == 0)
        for (int k = 0; k < TENS_COUNT; ++k) v[k] = tmp[blockDim.x + k];
}

// Per 256-atom block of an NPT system: phase NPT_SUMS sums v v^T (kinetic_tensor, no
// masses, properties.rs:41-43), r F^T (virial_tensor, properties.rs:45-47) and (1/2) m v^2
// of the live state; NPT_BEGIN first applies the step's velocity scaling and the box
// scaling of the positions (transformations.rs:6-15: s = h_old^-1 r, r = h s) and sums the
// KE only; NPT_END applies the velocity scaling, then sums all three.
enum { NPT_SUMS = 0, NPT_BEGIN = 1, NPT_END = 2 };

__global__ __launch_bounds__(kAtomThreads) void k_npt_atoms(const SysDev* __restrict__ sys,
                                                             const WorkItem* __restrict__ work, int phase)
{
    static_assert(kAtomThreads == 256, "the block sums below split 256 values into 8 x 32");
    __shared__ double red[10 * (kAtomThreads + 1)];
    __shared__ double red2[8 * 10];
    const WorkItem w = work[blockIdx.x];
    const SysDev& s = sys[w.sys];
    if (!s.npt_on || s.ctl[CTL_LEFT] == 0) return;
    const int a = w.idx + threadIdx.x;
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    double v[TENS_COUNT];
    for (int k = 0; k < TENS_COUNT; ++k) v[k] = 0.0;
    if (a < s.n) {
        double vx = gbl(st.vx)[a], vy = gbl(st.vy)[a], vz = gbl(st.vz)[a];
        if (phase != NPT_SUMS) {  // atoms.velocities = &scale * &atoms.velocities
            gemv3(s.npt + NPT_SCALE, vx, vy, vz, vx, vy, vz);
            gbl(st.vx)[a] = vx;
            gbl(st.vy)[a] = vy;
            gbl(st.vz)[a] = vz;
        }
        double x = gbl(st.x)[a], y = gbl(st.y)[a], z = gbl(st.z)[a];
        if (phase == NPT_BEGIN) {
            double f0, f1, f2;
            gemv3(s.npt + NPT_HINV_OLD, x, y, z, f0, f1, f2);
            gemv3(s.box + BOX_H, f0, f1, f2, x, y, z);
            gbl(st.x)[a] = x;
            gbl(st.y)[a] = y;
            gbl(st.z)[a] = z;
        }
        if (phase != NPT_BEGIN) {
            const double f[3] = {gbl(st.fx)[a], gbl(st.fy)[a], gbl(st.fz)[a]};
            const double r[3] = {x, y, z}, u[3] = {vx, vy, vz};
            for (int c = 0; c < 3; ++c)
                for (int q = 0; q < 3; ++q) {
                    v[TENS_K + 3 * c + q] = u[q] * u[c];
                    v[TENS_W + 3 * c + q] = r[q] * f[c];
                }
        }
        const double m = s.mass[gbl(st.type)[a]];
        v[TENS_KE] = 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
    }
    // Block sums in a fixed order.  NPT_BEGIN: the KE alone (the other columns are not
    // read before NPT_END rewrites them).  Otherwise every component: its 256 values go
    // through LDS, ten components at a time, and are added as 8 runs of 32, then the 8 runs.
    const int t = threadIdx.x;
    __attribute__((address_space(1))) double* out = gbl(s.tens) + (size_t)(w.idx / kAtomThreads) * TENS_COUNT;
    if (phase == NPT_BEGIN) {
        double x = v[TENS_KE];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((t & 63) == 0) red2[t >> 6] = x;
        __syncthreads();
        if (t == 0) out[TENS_KE] = (red2[0] + red2[1]) + (red2[2] + red2[3]);
        return;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int k = 0; k < 10; ++k) red[k * (kAtomThreads + 1) + t] = v[10 * r + k];
        __syncthreads();
        if (t < 80) {
            const int k = t % 10, part = t / 10;
            const double* q = red + k * (kAtomThreads + 1) + part * 32;
            double x = 0.0;
            for (int i = 0; i < 32; ++i) x += q[i];
            red2[part * 10 + k] = x;
        }
        __syncthreads();
        if (t < 10) {
            double x = 0.0;
            for (int part = 0; part < 8; ++part) x += red2[part * 10 + t];
            out[10 * r + t] = x;
        }
        __syncthreads();  // red and red2 are rewritten by the next round
    }
}

// Start of an NPT step (potential.rs:118-125), one workgroup per system: the momentum
// update from the pressure tensor of the live state, the step's velocity scaling
// exp(-dt/2 eta_dot), the box h <- exp(dt eta_dot) h with its inverse and volume, and the
// displacement bound under which the list built in an earlier box still holds every pair
// within the cutoff (the reference finds pairs afresh in every force call).  A new box
// thinner than the planned cell grid raises CTL_BOX and freezes the system before anything
// of the step is applied (the host re-plans the grid and resumes from that step).
__global__ __launch_bounds__(1024) void k_npt_begin(const SysDev* __restrict__ sys)
{
    __shared__ double tmp[1024 + TENS_COUNT];
    const SysDev& s = sys[blockIdx.x];
    if (!s.npt_on || s.ctl[CTL_LEFT] == 0) return;
    double v[TENS_COUNT];
    sum_tens(s, v, tmp);
    if (threadIdx.x != 0) return;
    double* npt = s.npt;
    double* box = s.box;
    // the step's barostat update on a copy, committed only when the new box still fits
    // the planned cell grid
    double nl[NPT_COUNT];
    for (int k = 0; k < NPT_COUNT; ++k) nl[k] = npt[k];
    mtk_add_delta(v, nl, box[BOX_VOL], s.dt);
    mtk_scale(nl, s.dt, -0.5, nl + NPT_SCALE);
    double sh[9], h[9], hi[9];
    mtk_scale(nl, s.dt, 1.0, sh);
    for (int k = 0; k < 9; ++k) nl[NPT_HINV_OLD + k] = box[BOX_HINV + k];
    mat3_mul(sh, box + BOX_H, h);
    mat3_inverse(h, hi);
    const double vol = fabs(mat3_det(h));
    if (s.mode != MODE_WHOLE) {
        for (int k = 0; k < 3; ++k) {
            const double* p = h + 3 * ((k + 1) % 3);
            const double* q = h + 3 * ((k + 2) % 3);
            const double cx = p[1] * q[2] - p[2] * q[1], cy = p[2] * q[0] - p[0] * q[2], cz = p[0] * q[1] - p[1] * q[0];
            if (!(vol / sqrt(cx * cx + cy * cy + cz * cz) >= nl[NPT_RCL] * (double)s.nc[k])) {
                // too thin for the cell grid: nothing of this step is applied and the system
                // freezes (every later kernel skips it) until the host re-plans the grid
                s.ctl[CTL_BOX] = 1;
                s.ctl[CTL_LEFT] = 0;
                return;
            }
        }
    }
    for (int k = 0; k < NPT_COUNT; ++k) npt[k] = nl[k];
    for (int k = 0; k < 9; ++k) {
        box[BOX_H + k] = h[k];
        box[BOX_HINV + k] = hi[k];
    }
    box[BOX_VOL] = vol;
    // Keep the list while it stays complete.  With D = h h_ref^-1 and d_i = r_i - D r_i,ref
    // (minimum image), a pair vector is r_ij = D r_ij,ref + d_j - d_i, so |r_ij| <= rc implies
    // |r_ij,ref| <= (rc + 2 max|d|) / sigma_min(D).  sigma_min(D) >= 1 - |D - I|_F, and the
    // list holds every pair within rc + skin at the build: the list is complete while
    // max|d| <= ((1 - |D - I|_F)(rc + skin) - rc) / 2.
    double def[9], dev = 0.0;
    mat3_mul(h, npt + NPT_HREF_INV, def);
    for (int k = 0; k < 9; ++k) {
        const double e = def[k] - ((k & 3) == 0 ? 1.0 : 0.0);
        dev += e * e;
    }
    const double rc = npt[NPT_RCMAX], skin = 2.0 * sqrt(s.half_skin2);
    const double lim = 0.5 * ((1.0 - sqrt(dev)) * (rc + skin) - rc) * (1.0 - 1e-9);
    for (int k = 0; k < 9; ++k) npt[NPT_DEF + k] = def[k];
    if (npt[NPT_LIST_OK] != 1.0 || !(lim > 0.0)) {
        s.ctl[CTL_REBUILD] = 1;
        npt[NPT_SKIN2] = 0.0;
    } else {
        npt[NPT_SKIN2] = fmin(lim * lim, s.half_skin2);
    }
}

// End of an NPT step (potential.rs:129-133) after the velocities were scaled again: the
// second momentum update; the step's thermo record gets the KE, temperature and pressure
// of the final velocities and the NPT Hamiltonian (simulation.rs:106-114): PE + KE + chain
// KE + chain PE (0: eta never moves) + barostat KE (npt.rs:60-62) + barostat PE (npt.rs:64-66).
__global__ __launch_bounds__(1024) void k_npt_end(const SysDev* __restrict__ sys, int record)
{
    __shared__ double tmp[1024 + TENS_COUNT];
    const SysDev& s = sys[blockIdx.x];
    if (!s.npt_on || s.ctl[CTL_LEFT] == 0) return;
    double v[TENS_COUNT];
    sum_tens(s, v, tmp);
    if (threadIdx.x != 0) return;
    double* npt = s.npt;
    const double* box = s.box;
    const double vol = box[BOX_VOL];
    mtk_add_delta(v, npt, vol, s.dt);
    const int k = s.ctl[CTL_STEP] - 1;
    if (!record || k < 0 || k >= s.thermo_cap) return;
    const double* m = npt + NPT_MOM;
    double mt[9], mm[9], pt[9], ph[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            mt[3 * c + r] = m[3 * r + c];
            pt[3 * c + r] = npt[NPT_TARGET + 3 * r + c];
        }
    mat3_mul(m, mt, mm);
    const double bke = ((mm[0] + mm[4]) + mm[8]) / (2.0 * npt[NPT_W]);
    mat3_mul(pt, box + BOX_H, ph);
    const double bpe = (ph[0] + ph[4]) + ph[8];
    const double ke = v[TENS_KE];
    const double trw = (v[TENS_W + 0] + v[TENS_W + 4]) + v[TENS_W + 8];
    __attribute__((address_space(1))) double* T = gbl(reinterpret_cast<double*>(s.thermo + k));
    const double pe = T[0];
    T[1] = ke;
    T[2] = pe + ke + nhc_kinetic(s.nhc) + 0.0 + bke + bpe;
    T[3] = (2.0 * ke) / ((double)(3 * (int64_t)s.n) * s.kb);
    T[4] = (2.0 * ke + trw) / (3.0 * vol);
    T[5] = trw;
}

// A list was built in the current box: the barostat's list check measures deformation
// from here.
__device__ void note_list_box(const SysDev& s)
{
    __attribute__((address_space(1))) double* npt = gbl(s.npt);
    const __attribute__((address_space(1))) double* box = gbl(s.box);
    for (int k = 0; k < 9; ++k) npt[NPT_HREF_INV + k] = box[BOX_HINV + k];
    npt[NPT_LIST_OK] = 1.0;
}

__global__ __launch_bounds__(1024) void k_reduce(const SysDev* __restrict__ sys, int flip, int record,
                                                 int step)
{
    __shared__ double rtmp[64];
    const SysDev& s = sys[blockIdx.x];
    // control words: all read first, then written (one round trip, not one per word), and
    // read before the partial sums so that the two round trips overlap
    __attribute__((address_space(1))) int32_t* ctl = gbl(s.ctl);
    const int c_cur = ctl[CTL_CUR], c_rb = ctl[CTL_REBUILD], c_nrb = ctl[CTL_NREBUILD];
    const int c_step = ctl[CTL_STEP], c_next = ctl[CTL_REBUILD_NEXT];
    const int c_remain = ctl[CTL_REMAIN], c_left = ctl[CTL_LEFT];
    double v[4];
    sum_partials(s, v, rtmp);
    if (step && c_left == 0) return;  // frozen: no step was taken (uniform over the block)
    if (threadIdx.x != 0) return;
    const double pe = 0.5 * v[0], ke = v[1], wabs = v[2], wpair = 0.5 * v[3];
    double ham = pe + ke;
    if (step && s.nhc_on) {
        // end of an NVT step (potential.rs:51-54), H with the chain's energy
        // (simulation.rs:100-105; the chain coordinates eta never move, so its potential
        // energy is 0), then the target ramp (simulation.rs:55-56, nvt.rs:109-118)
        double* c = s.nhc;
        nhc_forces(c, ke, s.n);
        nhc_half_step(c, s.dt);
        ham = ham + nhc_kinetic(c) + 0.0;
        c[NHC_T] = c[NHC_T0] + ((c[NHC_T1] - c[NHC_T0]) / c[NHC_TOTAL]) * c[NHC_STEP];
        c[NHC_STEP] = c[NHC_STEP] + 1.0;
        if (!s.npt_on && c_remain > 1 && c_left != 1) {
            // the next step of this call starts from these same velocities, so its chain
            // update (potential.rs:41-45, what k_nhc_pre does) is this KE's: done here,
            // and that step launches no k_nhc_pre
            nhc_forces(c, ke, s.n);
            nhc_half_step(c, s.dt);
            c[NHC_SCALE] = exp(-0.5 * s.dt * c[NHC_XI]);
        }
    }
    if (record) {
        if (c_step < s.thermo_cap) {
            __attribute__((address_space(1))) double* T = gbl(reinterpret_cast<double*>(s.thermo + c_step));
            T[0] = pe;
            T[1] = ke;
            T[2] = ham;
            T[3] = (2.0 * ke) / ((double)(3 * (int64_t)s.n) * s.kb);  // temperature
            T[4] = (2.0 * ke + wabs) / (3.0 * s.volume);             // pressure
            T[5] = wabs;
            T[6] = wpair;
        }
        ctl[CTL_STEP] = c_step + 1;
    }
    if (!isfinite(pe) || !isfinite(ke)) ctl[CTL_NAN] = 1;
    if (c_rb && (step || flip)) note_list_box(s);
    if (step) {
        // the step wrote into the set it did not read: that set becomes live; the skin
        // check of positions advanced ahead decides the next step's rebuild
        const int rb = c_rb ? 1 : 0;
        ctl[CTL_CUR] = c_cur ^ rb ^ 1;
        ctl[CTL_NREBUILD] = c_nrb + rb;
        ctl[CTL_REBUILD] = c_next;
        ctl[CTL_REBUILD_NEXT] = 0;
        if (c_remain > 0) ctl[CTL_REMAIN] = c_remain - 1;
        if (c_left > 0) ctl[CTL_LEFT] = c_left - 1;
    } else if (flip && c_rb) {
        ctl[CTL_CUR] = c_cur ^ 1;
        ctl[CTL_REBUILD] = 0;
        ctl[CTL_NREBUILD] = c_nrb + 1;
    }
}

// Every system's control words into one array [nsys][CTL_COUNT] (one copy to the host).
__global__ void k_gather_ctl(const SysDev* __restrict__ sys, int nsys, int32_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsys * CTL_COUNT) return;
    out[i] = sys[i / CTL_COUNT].ctl[i % CTL_COUNT];
}

// After a host-requested sort the set the next step writes into still holds the previous
// order's types and original indices: copy them from the live set.
__global__ __launch_bounds__(kAtomThreads) void k_sync_labels(const SysDev* __restrict__ sys,
                                                               const WorkItem* __restrict__ work)
{
    const WorkItem w = work[blockIdx.x];
    const SysDev& s = sys[w.sys];
    const int a = w.idx + threadIdx.x;
    if (a >= s.n) return;
    const int cur = __builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR]);
    gbl(s.st[cur ^ 1].type)[a] = gbl(s.st[cur].type)[a];
    gbl(s.st[cur ^ 1].orig)[a] = gbl(s.st[cur].orig)[a];
}

__global__ void k_flip(const SysDev* __restrict__ sys, int nsys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsys) return;
    const SysDev& s = sys[k];
    if (s.ctl[CTL_REBUILD]) {
        note_list_box(s);
        s.ctl[CTL_CUR] ^= 1;
        s.ctl[CTL_REBUILD] = 0;
        s.ctl[CTL_NREBUILD] += 1;
    }
}

// Test support: the list of every atom as original indices, out[orig * maxnb + k].
template <int SCAP>
__global__ __launch_bounds__(kTileThreads) void k_list_to_orig(const SysDev* __restrict__ sys,
                                                                const WorkItem* __restrict__ tiles,
                                                                int ntw, uint32_t* __restrict__ out,
                                                                uint32_t* __restrict__ counts)
{
    __shared__ double Px[SCAP + 1];
    __shared__ int pref[kMaxHalo + 1];
    __shared__ int ecell[kMaxHalo];
    __shared__ int stmp[16];
    const WorkItem w = tiles[blockIdx.x];
    const SysDev& s = sys[w.sys];
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    const TileGeom g = tile_geom(s, w.idx);
    const int total = stage_tile<SCAP, STAGE_INDEX, double>(s, st, g, 0.0, 0.0, 0.0, Px, nullptr, nullptr,
                                                    nullptr, pref, ecell, stmp);
    if (total > SCAP) return;
    const int* Li = reinterpret_cast<const int*>(Px);
    const int a0 = s.cell_start[g.c0], a1 = s.cell_start[g.c1];
    for (int a = a0 + (int)threadIdx.x; a < a1; a += blockDim.x) {
        const uint32_t oa = st.orig[a];
        const int nn = s.nnb[a];
        for (int k = 0; k < nn; ++k) {
            const uint4 u = s.nbr[(size_t)(k >> 3) * s.npad + a];
            const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
            const uint32_t off = (wv[(k & 7) >> 1] >> (((k & 7) & 1) * 16)) & 0xffffu;
            out[(size_t)oa * s.maxnb + k] = st.orig[Li[off >> s.list_shift]];
        }
        counts[oa] = (uint32_t)nn;
    }
}

// splitmix64's output function (Steele, Lea & Flood 2014; Vigna's constants): the entry
// hash of the list digest.
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Test support: per atom (original index) the list length and the order-independent row
// hash sum(mix64(original index of each entry)) mod 2^64.
template <int SCAP>
__global__ __launch_bounds__(kTileThreads) void k_list_digest(const SysDev* __restrict__ sys,
                                                               const WorkItem* __restrict__ tiles, int ntw,
                                                               uint32_t* __restrict__ counts,
                                                               uint64_t* __restrict__ hashes)
{
    __shared__ int Li[SCAP + 1];
    __shared__ int pref[kMaxHalo + 1];
    __shared__ int ecell[kMaxHalo];
    __shared__ int stmp[16];
    const WorkItem w = tiles[blockIdx.x];
    const SysDev& s = sys[w.sys];
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    const TileGeom g = tile_geom(s, w.idx);
    const int total = stage_tile<SCAP, STAGE_INDEX, int>(s, st, g, 0.0, 0.0, 0.0, Li, nullptr, nullptr, nullptr,
                                                         pref, ecell, stmp);
    if (total > SCAP) return;
    const int a0 = s.cell_start[g.c0], a1 = s.cell_start[g.c1];
    for (int a = a0 + (int)threadIdx.x; a < a1; a += blockDim.x) {
        const int nn = s.nnb[a];
        uint64_t h = 0;
        for (int k = 0; k < nn; ++k) {
            const uint4 u = s.nbr[(size_t)(k >> 3) * s.npad + a];
            const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
            const uint32_t off = (wv[(k & 7) >> 1] >> (((k & 7) & 1) * 16)) & 0xffffu;
            h += mix64((uint64_t)st.orig[Li[off >> s.list_shift]]);
        }
        counts[st.orig[a]] = (uint32_t)nn;
        hashes[st.orig[a]] = h;
    }
}

// Plain streaming copy used to calibrate hardware byte counters: reads and writes exactly
// n doubles.
__global__ __launch_bounds__(256) void k_copy_calibration(const double4* __restrict__ src,
                                                          double4* __restrict__ dst, size_t n4)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------------
// launch wrappers

void launch_copy_calibration(hipStream_t st, const void* src, void* dst, size_t bytes)
{
    const size_t n4 = bytes / 32;
    hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, st, (const double4*)src, (double4*)dst, n4);
}

// Staging capacities with compiled kernel instances.
#define PIS_SCAP_DISPATCH(scap, M) \
    do {                            \
        if ((scap) <= 1024)         \
            M(1024);                \
        else if ((scap) <= 1280)    \
            M(1280);                \
        else if ((scap) <= 1536)    \
            M(1536);                \
        else if ((scap) <= 1792)    \
            M(1792);                \
        else if ((scap) <= 2048)    \
            M(2048);                \
        else if ((scap) <= 3072)    \
            M(3072);                \
        else                        \
            M(4096);                \
    } while (0)

static inline int grid_cap(int n, int cap) { return n < cap ? (n > 0 ? n : 1) : cap; }

void launch_load_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, const double* pos,
                     const double* vel, const double* force, const uint8_t* type)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_load_aos, dim3((n + 255) / 256), dim3(256), 0, st, d_sys, sys, pos, vel,
                       force, type);
}

void launch_store_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, double* pos, double* vel,
                      double* force)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_store_aos, dim3((n + 255) / 256), dim3(256), 0, st, d_sys, sys, pos, vel,
                       force);
}

void launch_set_ctl(hipStream_t st, const SysDev* d_sys, int nsys, int slot, int value)
{
    hipLaunchKernelGGL(k_set_ctl, dim3((nsys + 255) / 256), dim3(256), 0, st, d_sys, nsys, slot, value);
}

void launch_integrate_pos(hipStream_t st, const SysDev* d_sys, const WorkItem* atom_work,
                          int n_atom_work)
{
    hipLaunchKernelGGL(k_integrate_pos, dim3(n_atom_work), dim3(kAtomThreads), 0, st, d_sys, atom_work);
}

void launch_rebuild_sort(hipStream_t st, const SysDev* d_sys, int nsys, const WorkItem* atom_work,
                         int n_atom_work)
{
    // Rebuild kernels loop over their work items, so a modest grid covers any size and
    // costs little on the steps where they return at once.
    const int ga = grid_cap(n_atom_work, PIS_SORT_GRID_CAP);
    hipLaunchKernelGGL(k_bin, dim3(ga), dim3(kAtomThreads), 0, st, d_sys, atom_work, n_atom_work);
    hipLaunchKernelGGL(k_scan, dim3(nsys), dim3(1024), 0, st, d_sys);
    hipLaunchKernelGGL(k_scatter, dim3(ga), dim3(kAtomThreads), 0, st, d_sys, atom_work, n_atom_work);
    hipLaunchKernelGGL(k_rank_permute, dim3(ga), dim3(kAtomThreads), 0, st, d_sys, atom_work,
                       n_atom_work);
}

void launch_build_list(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work, int n_tile_work,
                       int scap, int kinds)
{
    if (n_tile_work <= 0) return;
    const int gt = grid_cap(n_tile_work, PIS_BUILD_GRID_CAP);
#define PIS_BUILD_K(S, K) \
    hipLaunchKernelGGL((k_build_list<S, K>), dim3(gt), dim3(kTileThreads), 0, st, d_sys, tile_work, n_tile_work)
#define PIS_BUILD_O(S) PIS_BUILD_K(S, KIND_ORTHO)
#define PIS_BUILD_OR(S) PIS_BUILD_K(S, KIND_ORTHO_RAW)
#define PIS_BUILD_T(S) PIS_BUILD_K(S, KIND_TRI)
#define PIS_BUILD_TR(S) PIS_BUILD_K(S, KIND_TRI_RAW)
    if (kinds & KIND_ORTHO) PIS_SCAP_DISPATCH(scap, PIS_BUILD_O);
    if (kinds & KIND_ORTHO_RAW) PIS_SCAP_DISPATCH(scap, PIS_BUILD_OR);
    if (kinds & KIND_TRI) PIS_SCAP_DISPATCH(scap, PIS_BUILD_T);
    if (kinds & KIND_TRI_RAW) PIS_SCAP_DISPATCH(scap, PIS_BUILD_TR);
#undef PIS_BUILD_K
#undef PIS_BUILD_O
#undef PIS_BUILD_OR
#undef PIS_BUILD_T
#undef PIS_BUILD_TR
}

void launch_force(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work, int n_tile_work,
                  int scap, bool mixed, int integrate, int kinds)
{
    if (n_tile_work <= 0) return;
    const dim3 g(n_tile_work), b(kTileThreads);
#define PIS_FORCE_LAUNCH(S, M, K) \
    hipLaunchKernelGGL((k_force<S, M, K>), g, b, 0, st, d_sys, tile_work, n_tile_work, integrate)
#define PIS_FORCE_M(S) PIS_FORCE_LAUNCH(S, true, KIND_ORTHO)
#define PIS_FORCE_D(S) PIS_FORCE_LAUNCH(S, false, KIND_ORTHO)
#define PIS_FORCE_MT(S) PIS_FORCE_LAUNCH(S, true, KIND_TRI)
#define PIS_FORCE_DT(S) PIS_FORCE_LAUNCH(S, false, KIND_TRI)
#define PIS_FORCE_MR(S) PIS_FORCE_LAUNCH(S, true, KIND_TRI_RAW)
#define PIS_FORCE_DR(S) PIS_FORCE_LAUNCH(S, false, KIND_TRI_RAW)
#define PIS_FORCE_MO(S) PIS_FORCE_LAUNCH(S, true, KIND_ORTHO_RAW)
#define PIS_FORCE_DO(S) PIS_FORCE_LAUNCH(S, false, KIND_ORTHO_RAW)
    if (kinds & KIND_ORTHO) {
        if (mixed)
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_M);
        else
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_D);
    }
    if (kinds & KIND_ORTHO_RAW) {
        if (mixed)
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_MO);
        else
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_DO);
    }
    if (kinds & KIND_TRI) {
        if (mixed)
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_MT);
        else
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_DT);
    }
    if (kinds & KIND_TRI_RAW) {
        if (mixed)
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_MR);
        else
            PIS_SCAP_DISPATCH(scap, PIS_FORCE_DR);
    }
#undef PIS_FORCE_LAUNCH
#undef PIS_FORCE_M
#undef PIS_FORCE_D
#undef PIS_FORCE_MT
#undef PIS_FORCE_DT
#undef PIS_FORCE_MR
#undef PIS_FORCE_DR
#undef PIS_FORCE_MO
#undef PIS_FORCE_DO
}

void launch_reduce(hipStream_t st, const SysDev* d_sys, int nsys, int flip, int record, int step)
{
    hipLaunchKernelGGL(k_reduce, dim3(nsys), dim3(1024), 0, st, d_sys, flip, record, step);
}

void launch_nhc_pre(hipStream_t st, const SysDev* d_sys, int nsys)
{
    hipLaunchKernelGGL(k_nhc_pre, dim3(nsys), dim3(1024), 0, st, d_sys);
}

void launch_nhc_run(hipStream_t st, const SysDev* d_sys, int nsys, uint64_t step0, uint64_t total)
{
    hipLaunchKernelGGL(k_nhc_run, dim3((nsys + 255) / 256), dim3(256), 0, st, d_sys, nsys, (double)step0,
                       (double)total);
}

void launch_gather_ctl(hipStream_t st, const SysDev* d_sys, int nsys, int32_t* out)
{
    const int n = nsys * CTL_COUNT;
    hipLaunchKernelGGL(k_gather_ctl, dim3((n + 255) / 256), dim3(256), 0, st, d_sys, nsys, out);
}

void launch_npt(hipStream_t st, const SysDev* d_sys, int nsys, const WorkItem* atom_work, int n_atom_work,
                int what, int record)
{
    // what: 0 start-of-step sums, 1 start (barostat, box, scaled state), 2 end of the step
    if (what == 0) {
        hipLaunchKernelGGL(k_npt_atoms, dim3(n_atom_work), dim3(kAtomThreads), 0, st, d_sys, atom_work, (int)NPT_SUMS);
    } else if (what == 1) {
        hipLaunchKernelGGL(k_npt_begin, dim3(nsys), dim3(1024), 0, st, d_sys);
        hipLaunchKernelGGL(k_npt_atoms, dim3(n_atom_work), dim3(kAtomThreads), 0, st, d_sys, atom_work, (int)NPT_BEGIN);
    } else {
        hipLaunchKernelGGL(k_npt_atoms, dim3(n_atom_work), dim3(kAtomThreads), 0, st, d_sys, atom_work, (int)NPT_END);
        hipLaunchKernelGGL(k_npt_end, dim3(nsys), dim3(1024), 0, st, d_sys, record);
    }
}

void launch_sync_labels(hipStream_t st, const SysDev* d_sys, const WorkItem* atom_work, int n_atom_work)
{
    if (n_atom_work <= 0) return;
    hipLaunchKernelGGL(k_sync_labels, dim3(n_atom_work), dim3(kAtomThreads), 0, st, d_sys, atom_work);
}

void launch_flip(hipStream_t st, const SysDev* d_sys, int nsys)
{
    hipLaunchKernelGGL(k_flip, dim3((nsys + 255) / 256), dim3(256), 0, st, d_sys, nsys);
}

void launch_list_digest(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work, int n_tile_work,
                        int scap, uint32_t* counts, uint64_t* hashes)
{
    const dim3 g(n_tile_work), b(kTileThreads);
#define PIS_DIG(S) hipLaunchKernelGGL(k_list_digest<S>, g, b, 0, st, d_sys, tile_work, n_tile_work, counts, hashes)
    PIS_SCAP_DISPATCH(scap, PIS_DIG);
#undef PIS_DIG
}

void launch_list_to_orig(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work,
                         int n_tile_work, int scap, uint32_t* out, uint32_t* counts)
{
    const dim3 g(n_tile_work), b(kTileThreads);
#define PIS_L2O(S) \
    hipLaunchKernelGGL(k_list_to_orig<S>, g, b, 0, st, d_sys, tile_work, n_tile_work, out, counts)
    PIS_SCAP_DISPATCH(scap, PIS_L2O);
#undef PIS_L2O
}


}  // namespace pis
