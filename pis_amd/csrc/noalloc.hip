// noalloc.hip -- pis_compute_potential_noalloc: forces and PE of one system with every
// device buffer in a file-scope array, for callers that cannot allocate device memory.
//
// The array (`scratch`, kScratchBytes per GPU, held until the process ends) is resolved to
// a device address once per device with hipGetSymbolAddress; host code only ever passes
// that address on.  A call zeros the part it uses with a fill kernel, writes the system's
// descriptor and tables with one set-up kernel (kernel arguments, no host-to-device copy),
// loads the caller's device-resident positions and types, and runs the kernels of the
// allocating path (sort, list build, force, reduction) on them, so its forces are the same
// bits as pis_compute_potential's for the same input.  Calls on one device share the array:
// they are serialised by a lock, and each returns after its device work has finished.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.h"

using namespace pis;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // per GPU
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_addr_mu;                    // guards g_scratch
void* g_scratch[kMaxDevices] = {};       // device address of `scratch`, per device
std::mutex g_call_mu[kMaxDevices];       // one call at a time per device

static_assert(sizeof(NoAllocSetup) <= 4096, "set-up kernel arguments must stay within 4 KiB");

// Zeros n16 16-byte words (a kernel, not a memset node, so a captured sequence replays it).
__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, size_t n16)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// The descriptor, work tables, cell numbering (the same functions the host uses for an
// allocated system), pair table, masses and box of one system, from kernel arguments.
__global__ __launch_bounds__(256) void k_noalloc_setup(const NoAllocSetup a)
{
    const SysDev& d = a.dev;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (gid == 0) {
        *a.dev_slot = d;
        d.ctl[CTL_CUR] = 0;
        d.ctl[CTL_REBUILD] = 1;
    }
    for (int i = gid; i < d.ncell; i += stride) {
        const int x = i % d.nc[0], y = (i / d.nc[0]) % d.nc[1], z = i / (d.nc[0] * d.nc[1]);
        d.nat2sorted[i] = sorted_cell_id(d.nc, d.tile, x, y, z);
    }
    for (int t = gid; t <= d.ntile; t += stride)
        d.tile_cell0[t] = t < d.ntile ? tile_first_cell(d.nc, d.tile, t % d.nt[0], (t / d.nt[0]) % d.nt[1],
                                                        t / (d.nt[0] * d.nt[1]))
                                      : d.ncell;
    for (int k = gid; k < a.n_atom_work; k += stride) a.atom_work[k] = WorkItem{0, k * kAtomThreads};
    for (int k = gid; k < a.n_tile_work; k += stride) a.tile_work[k] = WorkItem{0, k};
    for (int k = gid; k < d.ntypes * d.ntypes; k += stride) const_cast<PairParam*>(d.pp)[k] = a.pp[k];
    for (int k = gid; k < d.ntypes; k += stride) const_cast<double*>(d.mass)[k] = a.mass[k];
    for (int k = gid; k < BOX_COUNT; k += stride) d.box[k] = a.box[k];
}

// k_load_aos for device-resident input: positions AoS and 1-based uint32 types, both
// checked here (CTL_BADINPUT bit 0: a type outside 1..ntypes, bit 1: a non-finite
// position); zero velocities and forces.
__global__ void k_load_typed(const SysDev* __restrict__ sys, const double* __restrict__ pos,
                             const uint32_t* __restrict__ types)
{
    const SysDev& s = sys[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    const State st = s.st[0];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const uint32_t t = types[i];
    int bad = 0;
    if (t < 1u || t > (uint32_t)s.ntypes) bad |= 1;
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) bad |= 2;
    if (bad) atomicOr(&s.ctl[CTL_BADINPUT], bad);
    st.x[i] = isfinite(x) ? x : 0.0;
    st.y[i] = isfinite(y) ? y : 0.0;
    st.z[i] = isfinite(z) ? z : 0.0;
    st.vx[i] = st.vy[i] = st.vz[i] = 0.0;
    st.fx[i] = st.fy[i] = st.fz[i] = 0.0;
    st.type[i] = (uint8_t)((bad & 1) ? 0u : t - 1u);
    st.orig[i] = (uint32_t)i;
}

// forces_inout (AoS, original order) += the live set's forces (when given); *pe_out = the
// record's PE (when given).
__global__ void k_store_forces_add(const SysDev* __restrict__ sys, double* __restrict__ force,
                                   double* __restrict__ pe_out)
{
    const SysDev& s = sys[0];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && pe_out) *pe_out = s.thermo[0].pe;
    if (p >= s.n || !force) return;
    const State st = s.st[__builtin_amdgcn_readfirstlane(s.ctl[CTL_CUR])];
    const uint32_t o = st.orig[p];
    force[3 * o] += st.fx[p];
    force[3 * o + 1] += st.fy[p];
    force[3 * o + 2] += st.fz[p];
}

// The control words of the system into `ctl` (host), after everything queued on st.
pis_status read_ctl_sync(hipStream_t st, const int32_t* d_ctl, int32_t* ctl)
{
    PIS_HIP(hipMemcpyAsync(ctl, d_ctl, CTL_COUNT * 4, hipMemcpyDeviceToHost, st));
    PIS_HIP(hipStreamSynchronize(st));
    return PIS_OK;
}

}  // namespace

extern "C" pis_status pis_compute_potential_noalloc(const pis_system_desc* desc, const double* d_pos,
                                                   const uint32_t* d_types, double* d_forces_inout,
                                                   double* d_pe, void* stream)
{
    if (!desc || !d_pos || !d_types) return fail(PIS_ERR_INVALID_ARGUMENT, "desc, positions and types are required");
    int dev = 0;
    PIS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PIS_ERR_UNSUPPORTED, "device index out of range");
    Plan p;
    PIS_TRY(make_plan(desc, &p, false));
    // the allocating path's carving of one system, then the descriptor, the work tables
    // and one thermo record
    const Layout Lo = layout_of(p);
    const int n = (int)p.n;
    const int n_aw = (n + kAtomThreads - 1) / kAtomThreads;
    size_t off = Lo.total;
    const size_t o_sys = off;
    off += align256(sizeof(SysDev));
    const size_t o_aw = off;
    off += align256((size_t)n_aw * sizeof(WorkItem));
    const size_t o_tw = off;
    off += align256((size_t)p.ntile * sizeof(WorkItem));
    const size_t o_th = off;
    off += align256(sizeof(ThermoDev));
    if (off > kScratchBytes)
        return fail(PIS_ERR_CAPACITY, "system of " + std::to_string(n) + " atoms needs " + std::to_string(off) +
                                          " bytes of device scratch; the allocation-free entry point holds " +
                                          std::to_string(kScratchBytes) + " per GPU (use pis_ctx_create)");
    std::lock_guard<std::mutex> call_lock(g_call_mu[dev]);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_addr_mu);
        if (!g_scratch[dev]) {
            void* ptr = nullptr;
            PIS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch[dev] = ptr;
        }
        base = g_scratch[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    char* b = (char*)base;
    NoAllocSetup a{};
    carve_system(p, Lo, b, a.dev);
    a.dev.thermo = (ThermoDev*)(b + o_th);
    a.dev.thermo_cap = 1;
    a.dev_slot = (SysDev*)(b + o_sys);
    a.atom_work = (WorkItem*)(b + o_aw);
    a.tile_work = (WorkItem*)(b + o_tw);
    a.n_atom_work = n_aw;
    a.n_tile_work = p.ntile;
    for (int k = 0; k < p.ntypes * p.ntypes; ++k) a.pp[k] = p.pp[k];
    for (int t = 0; t < p.ntypes; ++t) a.mass[t] = p.mass[t];
    plan_box(p, a.box);
    const SysDev* d_sys = a.dev_slot;
    const size_t n16 = off / 16;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<size_t>((n16 + 255) / 256, 2048)), dim3(256), 0, st,
                       (uint4*)base, n16);
    const int work = std::max(std::max(p.ncell, p.ntile + 1), std::max(n_aw, p.ntile));
    hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned)std::min((work + 255) / 256, 1024)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_load_typed, dim3((unsigned)n_aw), dim3(kAtomThreads), 0, st, d_sys, d_pos, d_types);
    // the list at these positions, as pis_set_state builds it: sort, build, flip, labels;
    // built again with a larger staging class if a tile's halo count comes close to it
    const int kind = plan_kind(p);
    int32_t ctl[CTL_COUNT];
    for (int pass = 0; pass < 2; ++pass) {
        launch_rebuild_sort(st, d_sys, 1, a.atom_work, n_aw);
        launch_build_list(st, d_sys, a.tile_work, p.ntile, p.scap, kind);
        launch_flip(st, d_sys, 1);
        launch_sync_labels(st, d_sys, a.atom_work, n_aw);
        PIS_HIP(hipGetLastError());
        PIS_TRY(read_ctl_sync(st, a.dev.ctl, ctl));
        if (ctl[CTL_BADINPUT] & 1) return fail(PIS_ERR_INVALID_ATOM_TYPE, "an atom type is out of range");
        if (ctl[CTL_BADINPUT] & 2) return fail(PIS_ERR_INVALID_ARGUMENT, "non-finite position");
        const int staged = ctl[CTL_MAXSTAGE];
        if (pass == 0 && p.mode != MODE_WHOLE && staged * 1.02 + 16 > p.scap) {
            const int grown = stage_class_for(staged * 1.1 + 48);
            if (grown < 0)
                return fail(PIS_ERR_CAPACITY, std::to_string(staged) +
                                                  " atoms in one tile's halo exceed the LDS staging limit");
            p.scap = grown;
            launch_set_ctl(st, d_sys, 1, CTL_OVF_STAGE, 0);
            launch_set_ctl(st, d_sys, 1, CTL_OVF_NB, 0);
            launch_set_ctl(st, d_sys, 1, CTL_REBUILD, 1);
            continue;
        }
        break;
    }
    if (ctl[CTL_OVF_STAGE]) return fail(PIS_ERR_CAPACITY, "LDS staging capacity exceeded");
    if (ctl[CTL_OVF_NB])
        return fail(PIS_ERR_CAPACITY, "neighbour capacity " + std::to_string(p.maxnb) + " exceeded (" +
                                          std::to_string(ctl[CTL_MAXNB]) + " needed)");
    launch_force(st, d_sys, a.tile_work, p.ntile, p.scap, p.mixed, 0, kind);
    launch_reduce(st, d_sys, 1, 1, 1, 0);
    hipLaunchKernelGGL(k_store_forces_add, dim3((unsigned)std::max(n_aw, 1)), dim3(kAtomThreads), 0, st, d_sys,
                       d_forces_inout, d_pe);
    PIS_HIP(hipGetLastError());
    PIS_TRY(read_ctl_sync(st, a.dev.ctl, ctl));
    if (ctl[CTL_OVF_STAGE]) return fail(PIS_ERR_CAPACITY, "LDS staging capacity exceeded");
    if (ctl[CTL_NAN]) return fail(PIS_ERR_NAN_ENERGY, "non-finite energy or position");
    return PIS_OK;
}
