// engine.h -- internals of the engine shared by its translation units (engine.hip: plans,
// contexts, the step loop and the C ABI; server.hip: continuous batching of jobs).
#pragma once
#include <string>
#include <vector>

#include "pis_hip.h"
#include "common.h"
#include "kernels.h"

namespace pis {

// Records `msg` as the calling thread's last error and returns `s`.
pis_status fail(pis_status s, const std::string& msg);
const std::string& last_error();

#define PIS_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? PIS_ERR_OUT_OF_MEMORY : PIS_ERR_HIP,     \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

#define PIS_TRY(expr)                    \
    do {                                 \
        pis_status s_ = (expr);          \
        if (s_ != PIS_OK) return s_;     \
    } while (0)

// Everything the host decides about one system before touching the device.
struct Plan {
    int64_t n = 0;
    int ntypes = 0;
    int nc[3] = {1, 1, 1}, tile[3] = {1, 1, 1}, nt[3] = {1, 1, 1};
    int tile_req[3] = {0, 0, 0};  // a caller's tile (pis_system_desc::tile); zeros: the model
    int ncell = 1, ntile = 1, mode = MODE_WHOLE, maxnb = 8, scap = 1024, npad = 64;
    double L[3] = {0, 0, 0}, Linv[3] = {0, 0, 0}, cellw[3] = {0, 0, 0};
    double rcmax = 0, rcl = 0, skin = 0, density = 0, volume = 0, kb = 1, wslack = 0;
    bool mixed = false;
    // general box (off-diagonal h): h, h^-1 as given (column-major); L holds the
    // perpendicular widths, and qmax bounds |x| of a wrapped position per component
    bool tri = false;
    double h[9] = {0}, hinv[9] = {0};
    double qmax = 0;
    uint32_t nbcap_req = 0;  // list capacity the caller asked for (0: from the density)
    double halo_scale = 1.0;  // tile halos are planned for density * halo_scale atoms
    double grid_margin = 0;   // cells are at least rcl * (1 + grid_margin) wide
    std::vector<PairParam> pp;
    std::vector<double> mass;
    std::vector<uint8_t> type0;
};


// need_types = false: desc->type_ids is not read (the types arrive in device memory).
pis_status make_plan(const pis_system_desc* d, Plan* p, bool need_types = true);
// Smallest LDS staging class holding `need` atoms (-1 if none does).
int stage_class_for(double need);

// Offsets of every array of a system's device memory (256-byte aligned) and the total.
struct Layout {
    size_t st[2][11];
    size_t r, ctl, cc, cs, n2s, tc, ac, ss, sk, sx, nbr, nnb, part, pp, m, nhc, box, npt, tens;
    size_t total;
};
size_t align256(size_t x);
Layout layout_of(const Plan& p);
// Points a system descriptor into memory at b carved as Lo, scalar fields from the plan.
void carve_system(const Plan& p, const Layout& Lo, char* b, SysDev& d);
// The box as the device keeps it (BOX_COUNT doubles): h, h^-1 and the volume.
void plan_box(const Plan& p, double* bx);

struct SysHost {
    Plan plan;
    void* mem = nullptr;
    size_t mem_bytes = 0;
    ThermoDev* thermo = nullptr;
    int thermo_cap = 0;
    bool pooled = false;  // mem and thermo belong to a caller's pool: never freed here
    SysDev dev{};
};

}  // namespace pis

struct pis_ctx {
    using SysHost = pis::SysHost;
    using WorkItem = pis::WorkItem;
    using SysDev = pis::SysDev;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<SysHost> sys;
    SysDev* d_sys = nullptr;
    size_t sys_cap = 0;  // SysDev slots allocated at d_sys
    WorkItem* d_atom_work = nullptr;
    int n_atom_work = 0;
    size_t atom_work_cap = 0;  // WorkItem slots allocated at d_atom_work
    WorkItem* d_tile_work = nullptr;
    int n_tile_work = 0;
    size_t tile_work_cap = 0;  // WorkItem slots allocated at d_tile_work
    // tiles grouped by LDS staging capacity: one launch per class
    struct TileClass {
        int scap;
        int kind;  // KIND_* of every tile in the class
        int first;
        int count;
        int lane;  // 0: `stream`, k > 0: side[k-1]
    };
    std::vector<TileClass> classes;
    int scap = 1024;  // largest class (single-system helpers)
    bool mixed = false;
    bool multi = false;
    // captured step sequences, keyed by whether thermo is recorded
    bool use_graph = false;
    // captured steps, [record + 2 nvt + 4 steady]: a call's first step (with the position
    // update kernel) and its later NVE steps (positions already advanced by the force kernel)
    hipGraphExec_t graph[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // a membership change (work tables, classes) leaves the captured steps stale: the next
    // call captures the step again and instantiates it; the host time this takes is counted
    bool graph_stale[8] = {false, false, false, false, false, false, false, false};
    uint64_t graph_instantiations = 0;
    double graph_build_s = 0.0;
    bool nhc_ready = false;  // pis_nvt_init has set every system's chain
    bool npt_ready = false;  // pis_npt_init has set every system's barostat
    // staging classes after the first run their list build and force on streams of their
    // own (default priority), forked from and joined back into `stream` by events
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    int nlanes = 1;
    int max_lanes = pis::kMaxLanes;  // streams the classes may spread over (the first is `stream`)
    void* d_xfer = nullptr;
    size_t xfer_bytes = 0;
    // trajectory snapshots: positions copied on `stream`, sent to the host on snap_stream
    // (default priority) while later steps run; one in flight at a time
    hipStream_t snap_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    double* d_snap = nullptr;
    size_t snap_bytes = 0;
    bool snap_pending = false;
    // statistics / timing
    uint64_t steps = 0, force_launches = 0;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double force_ms = 0.0;
    uint64_t timed_launches = 0;
};

namespace pis {

// Device memory of one system (one allocation, carved) and its tables.
pis_status alloc_system(SysHost& h);
// Every system's descriptor to the device copy the kernels read.
pis_status upload_sys(pis_ctx* c);
void free_ctx(pis_ctx* c);
// Tiles of all systems grouped by staging class, one table upload; drops captured graphs.
pis_status group_tiles(pis_ctx* c);
// Side streams and events for `nlanes` lanes (created once, kept until the context ends).
pis_status reserve_lanes(pis_ctx* c, int nlanes);
pis_status ensure_xfer(pis_ctx* c, size_t bytes);
pis_status read_ctl(pis_ctx* c, uint32_t k, int32_t* ctl);
pis_status write_ctl(pis_ctx* c, uint32_t k, const int32_t* ctl);
pis_status enqueue_classes(pis_ctx* c, bool build, int force);
// KIND_* bits of the context's systems (which build/force instantiations a launch needs)
inline int plan_kind(const Plan& p)
{
    const bool rel = p.mode == MODE_TILE_REL;
    return !p.tri ? (rel ? KIND_ORTHO : KIND_ORTHO_RAW) : (rel ? KIND_TRI : KIND_TRI_RAW);
}
// List entry shift of a system (SysDev::list_shift): FP32 staging in the force kernel for
// mixed precision with tile-relative staging, FP64 otherwise.
inline int plan_list_shift(const Plan& p) { return p.mixed && p.mode == MODE_TILE_REL ? 2 : 3; }
inline int ctx_kinds(const pis_ctx* c)
{
    int k = 0;
    for (const auto& h : c->sys) k |= plan_kind(h.plan);
    return k;
}
pis_status enqueue_rebuild(pis_ctx* c);
pis_status run_rebuild(pis_ctx* c);
// Status of every system (or only system `only`) from its device flags.
pis_status check_flags(pis_ctx* c, int only = -1);
pis_status ensure_thermo(pis_ctx* c, int cap);
pis_status sys_set_state(pis_ctx* c, uint32_t k, const double* pos, const double* vel, const double* force);
pis_status sys_get_state(pis_ctx* c, uint32_t k, double* pos, double* vel, double* force);
// An empty context (systems are added later) with the given precision.
pis_status ctx_create_empty(bool mixed, pis_ctx** out);
// Appends one planned system (device memory allocated, tables rebuilt) as index size()-1.
pis_status ctx_add(pis_ctx* c, Plan&& plan);
// Removes the systems at the given indices (any order); later indices shift down.
pis_status ctx_remove(pis_ctx* c, std::vector<size_t> idx);
// Moves the systems at the given indices out of the context without freeing them (their
// device memory and state stay as they are), in ascending index order.
pis_status ctx_take(pis_ctx* c, std::vector<size_t> idx, std::vector<SysHost>* out);
// Appends systems whose device memory is set up (taken from another context, or placed
// with alloc_system_at) as the last indices.
pis_status ctx_put(pis_ctx* c, std::vector<SysHost>&& add);
// Places a planned system in caller-owned device memory (system_bytes(plan, thermo_records)
// bytes, 256-byte aligned): arrays, zeros and tables enqueued on `st`, the thermo ring
// right after the arrays.  The system is marked pooled.
pis_status alloc_system_at(SysHost& h, void* mem, int thermo_records, hipStream_t st);
// Device bytes a system with this plan occupies (its allocation plus `thermo_records`).
size_t system_bytes(const Plan& p, int thermo_records);
// Enqueues `nsteps` steps (record: thermo per step; nvt: the chain is applied) of every
// system on the context's stream, from captured graphs when use_graph is set; the first
// step runs the position-update kernel, later NVE steps find their positions advanced by
// the previous step's force kernel.
pis_status enqueue_steps(pis_ctx* c, uint64_t nsteps, bool record, bool nvt);

}  // namespace pis
