// server.hip -- continuous batching of NVE simulation jobs on one GPU (pis_server_*).
//
// One worker thread owns three contexts, each with its own stream:
//   * `ctx`, the running batch: every resident job's system, advanced together in rounds
//     of `round_steps` batched steps replayed from captured graphs (a job whose step
//     budget, ctl[CTL_LEFT], runs out inside a round is frozen by the kernels for the rest
//     of it);
//   * `actx`, admission: while a round runs, queued jobs that fit the HBM budget and the
//     slot limit are placed, loaded, sorted, listed and given their initial forces here;
//   * `rctx`, collection: jobs that finished in the last round move here, and their final
//     state and thermo leave the device while the next round runs.
// Between rounds, finished jobs move ctx -> rctx and admitted ones actx -> ctx.  A system's
// device memory stays where it is; only the contexts' work tables change, and the captured
// round's step graphs are captured again (a few hundred microseconds of host time, once per
// round that changed the resident set).  Every system lives in one block of a device memory pool that grows by slabs and is freed only with the server,
// so neither admission nor retirement allocates or frees device memory per job.  A
// system's arithmetic does not depend on which other systems share its launches, so a
// job's result is bit-identical to the same job run alone.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"

using namespace pis;

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    Plan plan;
    std::vector<double> pos, vel;
    double dt = 0.0;
    uint64_t nsteps = 0;
    size_t bytes = 0;     // its block of the pool: the system and round_steps thermo records
    void* mem = nullptr;  // that block while the job is on the device
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    pis_status status = PIS_OK;
    std::string error;
    pis_thermo last{};
    std::vector<double> out_pos, out_vel, out_force;
    bool done = false;
};

// Device memory in slabs of at least `slab_bytes`, handed out first-fit in 256-byte units
// and taken back with adjacent free ranges merged.  Slabs are freed with the pool only
// (hipFree synchronises the device).
class DevicePool {
public:
    explicit DevicePool(size_t slab_bytes) : slab_bytes_(slab_bytes) {}
    ~DevicePool()
    {
        for (auto& s : slabs_) (void)hipFree(s.base);
    }
    DevicePool(const DevicePool&) = delete;
    DevicePool& operator=(const DevicePool&) = delete;

    pis_status alloc(size_t bytes, void** out)
    {
        bytes = align256(bytes);
        for (auto& s : slabs_)
            for (auto it = s.free.begin(); it != s.free.end(); ++it)
                if (it->second >= bytes) {
                    const size_t off = it->first, len = it->second;
                    s.free.erase(it);
                    if (len > bytes) s.free[off + bytes] = len - bytes;
                    *out = s.base + off;
                    return PIS_OK;
                }
        Slab s;
        s.size = std::max(slab_bytes_, bytes);
        PIS_HIP(hipMalloc((void**)&s.base, s.size));
        if (s.size > bytes) s.free[bytes] = s.size - bytes;
        *out = s.base;
        reserved_ += s.size;
        slabs_.push_back(std::move(s));
        return PIS_OK;
    }

    void release(void* p, size_t bytes)
    {
        bytes = align256(bytes);
        char* q = (char*)p;
        for (auto& s : slabs_) {
            if (q < s.base || q >= s.base + s.size) continue;
            size_t off = (size_t)(q - s.base), len = bytes;
            auto next = s.free.lower_bound(off);
            if (next != s.free.end() && off + len == next->first) {  // merge with the range after
                len += next->second;
                next = s.free.erase(next);
            }
            if (next != s.free.begin()) {  // and with the range before
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    prev->second += len;
                    return;
                }
            }
            s.free[off] = len;
            return;
        }
    }

    size_t reserved() const { return reserved_; }

private:
    struct Slab {
        char* base = nullptr;
        size_t size = 0;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Slab> slabs_;
    size_t slab_bytes_;
    size_t reserved_ = 0;
};

}  // namespace

struct pis_server {
    pis_ctx* ctx = nullptr;   // running batch
    pis_ctx* actx = nullptr;  // admission
    pis_ctx* rctx = nullptr;  // collection
    std::vector<std::shared_ptr<Job>> resident;  // parallel to ctx->sys
    std::vector<std::shared_ptr<Job>> admitted;  // parallel to actx->sys
    std::vector<std::shared_ptr<Job>> retiring;  // parallel to rctx->sys
    std::deque<std::shared_ptr<Job>> pending;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;  // submitted and not yet collected
    std::unique_ptr<DevicePool> pool;
    std::mutex mu;  // guards the queue, jobs, budget and statistics
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    uint64_t budget = 0, used = 0;
    uint32_t max_systems = 0, round_steps = 0;
    bool mixed = false;
    uint64_t next_id = 1;
    Clock::time_point t0;
    int32_t* d_ctl = nullptr;
    size_t d_ctl_cap = 0;
    // statistics
    uint64_t submitted = 0, completed = 0, failed = 0, rounds = 0, atom_steps = 0, peak_bytes = 0;
    uint32_t peak_resident = 0;
    size_t n_resident = 0;  // resident.size() as last published under the lock
    uint64_t graph_inst = 0, pool_bytes = 0;
    double graph_s = 0.0;
    pis_status fatal = PIS_OK;
    std::string fatal_msg;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Ends a job (success or failure) and wakes its waiter; called with the lock held.
void finish(pis_server* sv, const std::shared_ptr<Job>& j, pis_status st, const std::string& msg)
{
    j->status = st;
    j->error = msg;
    j->t_done = sv->now();
    j->done = true;
    j->pos.clear();
    j->pos.shrink_to_fit();
    j->vel.clear();
    j->vel.shrink_to_fit();
    if (st == PIS_OK)
        ++sv->completed;
    else
        ++sv->failed;
    sv->cv_done.notify_all();
}

// Gives a job's pool block back and takes it off the budget; called with the lock held
// once no launch can touch the block any more.
void release_block(pis_server* sv, Job& j)
{
    if (j.mem) sv->pool->release(j.mem, j.bytes);
    j.mem = nullptr;
    sv->used -= j.bytes;
}

std::vector<size_t> all_of(const pis_ctx* c)
{
    std::vector<size_t> v(c->sys.size());
    for (size_t k = 0; k < v.size(); ++k) v[k] = k;
    return v;
}

// Places one queued job in the admission context: its block of the pool, the arrays and
// tables, the state (sorted and listed) and its step budget.  Its initial forces come
// with the admission context's next force launch (admit_forces).
pis_status admit_one(pis_server* sv, const std::shared_ptr<Job>& j)
{
    pis_ctx* a = sv->actx;
    PIS_TRY(sv->pool->alloc(j->bytes, &j->mem));
    std::vector<SysHost> add(1);
    add[0].plan = j->plan;
    PIS_TRY(alloc_system_at(add[0], j->mem, (int)sv->round_steps, a->stream));
    add[0].dev.dt = j->dt;
    add[0].dev.dt2 = j->dt * j->dt;  // dt.powi(2)
    const pis_status put = ctx_put(a, std::move(add));
    sv->admitted.push_back(j);  // the system is in the context now, whatever put returned
    PIS_TRY(put);
    const uint32_t k = (uint32_t)a->sys.size() - 1;
    PIS_TRY(sys_set_state(a, k, j->pos.data(), j->vel.data(), nullptr));
    int32_t ctl[CTL_COUNT];
    PIS_TRY(read_ctl(a, k, ctl));
    ctl[CTL_LEFT] = (int32_t)j->nsteps;
    ctl[CTL_STEP] = 0;
    return write_ctl(a, k, ctl);
}

// Initial forces of every admitted system (one force launch per class of the admission
// context, integrate off), waited for.
pis_status admit_forces(pis_server* sv)
{
    pis_ctx* a = sv->actx;
    if (a->sys.empty()) return PIS_OK;
    PIS_TRY(enqueue_classes(a, false, 0));
    PIS_HIP(hipGetLastError());
    PIS_HIP(hipStreamSynchronize(a->stream));
    return PIS_OK;
}

// The collection context's jobs: final state and last thermo record to the host, then
// the jobs end and their blocks go back to the pool.
pis_status collect_all(pis_server* sv)
{
    pis_ctx* r = sv->rctx;
    if (r->sys.empty()) return PIS_OK;
    for (size_t k = 0; k < r->sys.size(); ++k) {
        Job& j = *sv->retiring[k];
        const size_t n = (size_t)r->sys[k].plan.n;
        j.out_pos.resize(3 * n);
        j.out_vel.resize(3 * n);
        j.out_force.resize(3 * n);
        PIS_TRY(sys_get_state(r, (uint32_t)k, j.out_pos.data(), j.out_vel.data(), j.out_force.data()));
        int32_t ctl[CTL_COUNT];
        PIS_TRY(read_ctl(r, (uint32_t)k, ctl));
        const int rec = ctl[CTL_STEP] - 1;
        if (rec >= 0 && rec < r->sys[k].thermo_cap) {
            ThermoDev t;
            PIS_HIP(hipMemcpyAsync(&t, r->sys[k].thermo + rec, sizeof(ThermoDev), hipMemcpyDeviceToHost, r->stream));
            PIS_HIP(hipStreamSynchronize(r->stream));
            j.last = pis_thermo{t.pe, t.ke, t.etot, t.temperature, t.pressure, t.virial_abs, t.virial_pair};
        }
    }
    std::vector<SysHost> gone;
    PIS_TRY(ctx_take(r, all_of(r), &gone));
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& j : sv->retiring) {
        sv->atom_steps += (uint64_t)j->plan.n * j->nsteps;
        release_block(sv, *j);
        finish(sv, j, PIS_OK, "");
    }
    sv->retiring.clear();
    return PIS_OK;
}

// Control words of every resident system, in one gather launch and one copy.
pis_status gather_ctl(pis_server* sv, std::vector<int32_t>& ctl)
{
    pis_ctx* c = sv->ctx;
    const size_t need = c->sys.size() * CTL_COUNT;
    if (sv->d_ctl_cap < need) {
        if (sv->d_ctl) PIS_HIP(hipFree(sv->d_ctl));
        sv->d_ctl = nullptr;
        sv->d_ctl_cap = 0;
        PIS_HIP(hipMalloc(&sv->d_ctl, 2 * need * 4));
        sv->d_ctl_cap = 2 * need;
    }
    launch_gather_ctl(c->stream, c->d_sys, (int)c->sys.size(), sv->d_ctl);
    ctl.resize(need);
    PIS_HIP(hipMemcpyAsync(ctl.data(), sv->d_ctl, need * 4, hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

// After a round (its stream idle): jobs whose budget ran out move to the collection
// context; failed ones (capacity, non-finite values) end with their error at once.
pis_status retire(pis_server* sv)
{
    pis_ctx* c = sv->ctx;
    if (c->sys.empty()) return PIS_OK;
    std::vector<int32_t> ctl;
    PIS_TRY(gather_ctl(sv, ctl));
    struct Out {
        size_t k;
        pis_status st;
        std::string msg;
    };
    std::vector<Out> out;
    for (size_t k = 0; k < c->sys.size(); ++k) {
        const int32_t* q = &ctl[k * CTL_COUNT];
        if (q[CTL_OVF_STAGE])
            out.push_back({k, PIS_ERR_CAPACITY, "LDS staging capacity exceeded"});
        else if (q[CTL_OVF_NB])
            out.push_back({k, PIS_ERR_CAPACITY, "neighbour capacity " + std::to_string(c->sys[k].plan.maxnb) +
                                                    " exceeded (" + std::to_string(q[CTL_MAXNB]) + " needed)"});
        else if (q[CTL_NAN])
            out.push_back({k, PIS_ERR_NAN_ENERGY, "non-finite energy or position"});
        else if (q[CTL_LEFT] == 0)
            out.push_back({k, PIS_OK, ""});
    }
    if (out.empty()) return PIS_OK;
    std::vector<size_t> idx;
    for (const auto& o : out) idx.push_back(o.k);
    std::vector<SysHost> taken;
    PIS_TRY(ctx_take(c, idx, &taken));  // in ascending index order, like `out`
    std::vector<SysHost> to_collect;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (size_t q = 0; q < out.size(); ++q) {
            auto& j = sv->resident[out[q].k];
            if (out[q].st != PIS_OK) {
                release_block(sv, *j);
                finish(sv, j, out[q].st, out[q].msg);
            } else {
                to_collect.push_back(std::move(taken[q]));
                sv->retiring.push_back(j);
            }
        }
    }
    for (size_t q = out.size(); q-- > 0;) sv->resident.erase(sv->resident.begin() + (ptrdiff_t)out[q].k);
    return ctx_put(sv->rctx, std::move(to_collect));
}

// The next queued job if it fits the slots and the budget (arrival order: a head that does
// not fit holds back those behind it), taken off the queue and charged; null otherwise.
std::shared_ptr<Job> next_admissible(pis_server* sv)
{
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop || sv->pending.empty()) return nullptr;
    if (sv->resident.size() + sv->admitted.size() >= sv->max_systems) return nullptr;
    auto j = sv->pending.front();
    if (sv->used + j->bytes > sv->budget) return nullptr;
    sv->pending.pop_front();
    sv->used += j->bytes;
    sv->peak_bytes = std::max(sv->peak_bytes, sv->used);
    return j;
}

// A device-level failure ends every job and stops the server.
void fail_all(pis_server* sv, pis_status st)
{
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->fatal = st;
    sv->fatal_msg = last_error();
    for (auto* v : {&sv->resident, &sv->admitted, &sv->retiring})
        for (auto& j : *v)
            if (!j->done) finish(sv, j, st, sv->fatal_msg);
    for (auto& j : sv->pending) finish(sv, j, st, sv->fatal_msg);
    sv->resident.clear();
    sv->admitted.clear();
    sv->retiring.clear();
    sv->pending.clear();
    sv->n_resident = 0;
    sv->stop = true;
}

void worker_main(pis_server* sv)
{
    (void)hipSetDevice(sv->ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] {
                return sv->stop || !sv->pending.empty() || !sv->resident.empty() || !sv->retiring.empty();
            });
            if (sv->stop) break;
        }
        pis_status st = PIS_OK;
        // a round of the resident batch on its stream
        const bool running = !sv->resident.empty();
        if (running) {
            pis_ctx* c = sv->ctx;
            launch_set_ctl(c->stream, c->d_sys, (int)c->sys.size(), CTL_STEP, 0);
            st = enqueue_steps(c, (uint64_t)sv->round_steps, true, false);
            if (st == PIS_OK) {
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) st = fail(PIS_ERR_HIP, std::string("round launch: ") + hipGetErrorString(e));
            }
            c->steps += sv->round_steps;
        }
        // meanwhile, on the other streams: the last round's finished jobs leave the device
        // and queued jobs come on, until the round is over (at least one admission a round)
        if (st == PIS_OK) st = collect_all(sv);
        while (st == PIS_OK) {
            auto j = next_admissible(sv);
            if (!j) break;
            j->t_admit = sv->now();
            const pis_status as = admit_one(sv, j);
            if (as != PIS_OK) {
                // the job fails alone: its system (if it got that far) leaves the
                // admission context, whose stream is idle after the failed call
                const std::string msg = last_error();
                if (!sv->admitted.empty() && sv->admitted.back() == j) {
                    std::vector<SysHost> gone;
                    st = ctx_take(sv->actx, {sv->actx->sys.size() - 1}, &gone);
                    sv->admitted.pop_back();
                }
                std::lock_guard<std::mutex> lk(sv->mu);
                release_block(sv, *j);
                finish(sv, j, as, msg);
                continue;
            }
            if (running && hipStreamQuery(sv->ctx->stream) != hipErrorNotReady) break;
        }
        if (st == PIS_OK) st = admit_forces(sv);
        // end of the round: finished jobs out, admitted ones in
        if (st == PIS_OK && running) {
            const hipError_t e = hipStreamSynchronize(sv->ctx->stream);
            if (e != hipSuccess) st = fail(PIS_ERR_HIP, std::string("round: ") + hipGetErrorString(e));
            if (st == PIS_OK) st = retire(sv);
        }
        if (st == PIS_OK && !sv->admitted.empty()) {
            std::vector<SysHost> moving;
            st = ctx_take(sv->actx, all_of(sv->actx), &moving);
            if (st == PIS_OK) st = ctx_put(sv->ctx, std::move(moving));
            if (st == PIS_OK) {
                sv->resident.insert(sv->resident.end(), sv->admitted.begin(), sv->admitted.end());
                sv->admitted.clear();
            }
        }
        if (st != PIS_OK) {
            fail_all(sv, st);
            break;
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        if (running) ++sv->rounds;
        sv->n_resident = sv->resident.size();
        sv->peak_resident = std::max(sv->peak_resident, (uint32_t)sv->resident.size());
        sv->graph_inst = sv->ctx->graph_instantiations;
        sv->graph_s = sv->ctx->graph_build_s;
        sv->pool_bytes = sv->pool->reserved();
    }
}

}  // namespace

extern "C" {

pis_status pis_server_create(uint64_t hbm_budget_bytes, uint32_t max_systems, uint32_t round_steps,
                             int32_t precision, pis_server** out)
{
    if (!out) return fail(PIS_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (max_systems == 0 || round_steps == 0 || round_steps > 100000 || hbm_budget_bytes == 0)
        return fail(PIS_ERR_INVALID_ARGUMENT, "need max_systems > 0, 0 < round_steps <= 100000, budget > 0");
    if (precision != PIS_PRECISION_FP64 && precision != PIS_PRECISION_MIXED)
        return fail(PIS_ERR_INVALID_ARGUMENT, "unknown precision");
    const bool mixed = precision == PIS_PRECISION_MIXED;
    pis_ctx *c = nullptr, *a = nullptr, *r = nullptr;
    pis_status st = ctx_create_empty(mixed, &c);
    if (st == PIS_OK) st = ctx_create_empty(mixed, &a);
    if (st == PIS_OK) st = ctx_create_empty(mixed, &r);
    if (st != PIS_OK) {
        free_ctx(c);
        free_ctx(a);
        free_ctx(r);
        return st;
    }
    c->use_graph = true;
    // The device serves a few hardware queues per process and streams share them in turn:
    // four streams in all, created together -- the running batch on two lanes, admission
    // and collection on one each -- so the round's launches do not queue behind admission
    // or collection work.
    c->max_lanes = 2;
    a->max_lanes = 1;
    r->max_lanes = 1;
    st = reserve_lanes(c, 2);
    if (st != PIS_OK) {
        free_ctx(c);
        free_ctx(a);
        free_ctx(r);
        return st;
    }
    pis_server* sv = new pis_server();
    sv->ctx = c;
    sv->actx = a;
    sv->rctx = r;
    // slabs of 1/16 of the budget, at least 256 MiB: a handful of hipMalloc calls a run
    sv->pool.reset(new DevicePool(std::max<size_t>((size_t)256 << 20, hbm_budget_bytes / 16)));
    sv->budget = hbm_budget_bytes;
    sv->max_systems = max_systems;
    sv->round_steps = round_steps;
    sv->mixed = mixed;
    sv->t0 = Clock::now();
    sv->worker = std::thread(worker_main, sv);
    *out = sv;
    return PIS_OK;
}

pis_status pis_server_submit(pis_server* sv, const pis_system_desc* desc, const double* pos, const double* vel,
                             double dt, uint64_t nsteps, uint64_t* job_id)
{
    if (!sv || !desc || !pos || !job_id) return fail(PIS_ERR_INVALID_ARGUMENT, "null argument");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PIS_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (nsteps == 0 || nsteps > (uint64_t)INT32_MAX) return fail(PIS_ERR_INVALID_ARGUMENT, "need 0 < nsteps < 2^31");
    auto j = std::make_shared<Job>();
    PIS_TRY(make_plan(desc, &j->plan));
    if (j->plan.mixed != sv->mixed)
        return fail(PIS_ERR_INVALID_ARGUMENT, "job precision differs from the server's");
    const size_t n = (size_t)j->plan.n;
    for (size_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i])) return fail(PIS_ERR_INVALID_ARGUMENT, "non-finite position");
    j->pos.assign(pos, pos + 3 * n);
    if (vel)
        j->vel.assign(vel, vel + 3 * n);
    else
        j->vel.assign(3 * n, 0.0);
    j->dt = dt;
    j->nsteps = nsteps;
    j->bytes = align256(system_bytes(j->plan, (int)sv->round_steps));
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop)
        return fail(sv->fatal != PIS_OK ? sv->fatal : PIS_ERR_INVALID_ARGUMENT, "server stopped: " + sv->fatal_msg);
    if (j->bytes > sv->budget)
        return fail(PIS_ERR_OUT_OF_MEMORY, "job needs " + std::to_string(j->bytes) + " bytes, above the server budget");
    j->id = sv->next_id++;
    j->t_submit = sv->now();
    sv->jobs[j->id] = j;
    sv->pending.push_back(j);
    ++sv->submitted;
    *job_id = j->id;
    sv->cv_work.notify_one();
    return PIS_OK;
}

pis_status pis_server_wait(pis_server* sv, uint64_t job_id, double timeout_s, pis_job_result* result, double* pos,
                           double* vel, double* force)
{
    if (!sv) return fail(PIS_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(job_id);
    if (it == sv->jobs.end()) return fail(PIS_ERR_NOT_FOUND, "unknown or already collected job");
    std::shared_ptr<Job> j = it->second;
    auto ready = [&] { return j->done; };
    if (timeout_s < 0.0)
        sv->cv_done.wait(lk, ready);
    else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return fail(PIS_ERR_TIMEOUT, "job " + std::to_string(job_id) + " still running");
    sv->jobs.erase(it);
    lk.unlock();
    if (result) {
        result->job_id = j->id;
        result->status = j->status;
        result->reserved = 0;
        result->n_atoms = (uint64_t)j->plan.n;
        result->steps = j->nsteps;
        result->submit_s = j->t_submit;
        result->admit_s = j->t_admit;
        result->done_s = j->t_done;
        result->last = j->last;
    }
    if (j->status == PIS_OK) {
        const size_t m = j->out_pos.size();
        if (pos) std::copy(j->out_pos.begin(), j->out_pos.begin() + (ptrdiff_t)m, pos);
        if (vel) std::copy(j->out_vel.begin(), j->out_vel.begin() + (ptrdiff_t)m, vel);
        if (force) std::copy(j->out_force.begin(), j->out_force.begin() + (ptrdiff_t)m, force);
    } else {
        fail(j->status, j->error);
    }
    return PIS_OK;
}

pis_status pis_server_get_stats(pis_server* sv, pis_server_stats* out)
{
    if (!sv || !out) return fail(PIS_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    out->submitted = sv->submitted;
    out->completed = sv->completed;
    out->failed = sv->failed;
    out->rounds = sv->rounds;
    out->atom_steps = sv->atom_steps;
    out->resident = sv->n_resident;
    out->pending = sv->pending.size();
    out->bytes_in_use = sv->used;
    out->peak_bytes = sv->peak_bytes;
    out->peak_resident = sv->peak_resident;
    out->reserved = 0;
    out->graph_instantiations = sv->graph_inst;
    out->graph_build_s = sv->graph_s;
    out->pool_bytes = sv->pool_bytes;
    return PIS_OK;
}

pis_status pis_server_destroy(pis_server* sv)
{
    if (!sv) return PIS_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        for (auto& j : sv->pending) finish(sv, j, PIS_ERR_INVALID_ARGUMENT, "server destroyed");
        sv->pending.clear();
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (auto* v : {&sv->resident, &sv->admitted, &sv->retiring})
            for (auto& j : *v)
                if (!j->done) finish(sv, j, PIS_ERR_INVALID_ARGUMENT, "server destroyed");
    }
    if (sv->d_ctl) (void)hipFree(sv->d_ctl);
    // the contexts first (they leave pooled systems' memory alone), then the pool's slabs
    free_ctx(sv->ctx);
    free_ctx(sv->actx);
    free_ctx(sv->rctx);
    sv->pool.reset();
    delete sv;
    return PIS_OK;
}

}  // extern "C"
