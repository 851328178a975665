// server.hip -- continuous batching of NVE simulation jobs on one GPU (pis_server_*).
//
// One worker thread owns a context whose systems are the resident jobs.  It loops over
// rounds: retire the jobs whose step budget ran out (final state and thermo copied to
// the host), admit queued jobs in arrival order while the HBM budget and the slot limit
// allow (upload, list build, initial forces of that system alone), then advance every
// resident job by up to `round_steps` steps in one batched step sequence per step.  A
// job's step budget lives on the device (ctl[CTL_LEFT]): a job that finishes inside a
// round is frozen by the kernels for the rest of it.  A system's arithmetic does not
// depend on which other systems share its launches, so a job's result is bit-identical
// to the same job run alone.
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
    size_t bytes = 0;
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    pis_status status = PIS_OK;
    std::string error;
    pis_thermo last{};
    std::vector<double> out_pos, out_vel, out_force;
    bool done = false;
};

}  // namespace

struct pis_server {
    pis_ctx* ctx = nullptr;
    std::vector<std::shared_ptr<Job>> resident;  // parallel to ctx->sys
    std::deque<std::shared_ptr<Job>> pending;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;  // submitted and not yet collected
    std::mutex mu;
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

// Puts an admitted job on the device: its system (appended), state, list and initial
// forces (a force launch over this system's tiles only), and its step budget.
pis_status admit(pis_server* sv, const std::shared_ptr<Job>& j)
{
    pis_ctx* c = sv->ctx;
    PIS_TRY(ctx_add(c, Plan(j->plan)));
    const uint32_t k = (uint32_t)c->sys.size() - 1;
    c->sys[k].dev.dt = j->dt;
    c->sys[k].dev.dt2 = j->dt * j->dt;  // dt.powi(2)
    PIS_TRY(ensure_thermo(c, (int)sv->round_steps));
    PIS_TRY(upload_sys(c));
    PIS_TRY(sys_set_state(c, k, j->pos.data(), j->vel.data(), nullptr));
    // staging growth in sys_set_state may have regrouped the tiles: find this system's
    // tiles (its class lists them last, systems being grouped in index order)
    const Plan& p = c->sys[k].plan;
    for (const auto& q : c->classes) {
        if (q.scap != p.scap || q.kind != plan_kind(p)) continue;
        launch_force(c->stream, c->d_sys, c->d_tile_work + q.first + (q.count - p.ntile), p.ntile, q.scap,
                     c->mixed, 0, plan_kind(p));
    }
    int32_t ctl[CTL_COUNT];
    PIS_TRY(read_ctl(c, k, ctl));
    ctl[CTL_LEFT] = (int32_t)j->nsteps;
    ctl[CTL_STEP] = 0;
    PIS_TRY(write_ctl(c, k, ctl));
    PIS_HIP(hipGetLastError());
    return PIS_OK;
}

// Copies a finished job's final state and last thermo record to the host.
pis_status collect(pis_server* sv, uint32_t k, const int32_t* ctl, Job& j)
{
    pis_ctx* c = sv->ctx;
    const size_t n = (size_t)c->sys[k].plan.n;
    j.out_pos.resize(3 * n);
    j.out_vel.resize(3 * n);
    j.out_force.resize(3 * n);
    PIS_TRY(sys_get_state(c, k, j.out_pos.data(), j.out_vel.data(), j.out_force.data()));
    const int rec = ctl[CTL_STEP] - 1;
    if (rec >= 0 && rec < c->sys[k].thermo_cap) {
        ThermoDev t;
        PIS_HIP(hipMemcpy(&t, c->sys[k].thermo + rec, sizeof(ThermoDev), hipMemcpyDeviceToHost));
        j.last = pis_thermo{t.pe, t.ke, t.etot, t.temperature, t.pressure, t.virial_abs, t.virial_pair};
    }
    return PIS_OK;
}

// One round: `steps` batched NVE steps of every resident system (frozen ones skipped by
// the kernels), replayed from a captured graph.
pis_status run_round(pis_server* sv, int steps)
{
    pis_ctx* c = sv->ctx;
    const int ns = (int)c->sys.size();
    launch_set_ctl(c->stream, c->d_sys, ns, CTL_STEP, 0);
    PIS_TRY(enqueue_steps(c, (uint64_t)steps, true, false));
    PIS_HIP(hipGetLastError());
    c->steps += (uint64_t)steps;
    return PIS_OK;
}

pis_status gather_ctl(pis_server* sv, std::vector<int32_t>& ctl)
{
    pis_ctx* c = sv->ctx;
    const size_t need = c->sys.size() * CTL_COUNT;
    if (sv->d_ctl_cap < need) {
        if (sv->d_ctl) PIS_HIP(hipFree(sv->d_ctl));
        sv->d_ctl = nullptr;
        PIS_HIP(hipMalloc(&sv->d_ctl, need * 4));
        sv->d_ctl_cap = need;
    }
    launch_gather_ctl(c->stream, c->d_sys, (int)c->sys.size(), sv->d_ctl);
    ctl.resize(need);
    PIS_HIP(hipMemcpyAsync(ctl.data(), sv->d_ctl, need * 4, hipMemcpyDeviceToHost, c->stream));
    PIS_HIP(hipStreamSynchronize(c->stream));
    return PIS_OK;
}

// Retires finished or failed resident jobs; returns their count.
pis_status retire(pis_server* sv, int* nretired)
{
    pis_ctx* c = sv->ctx;
    *nretired = 0;
    if (c->sys.empty()) return PIS_OK;
    std::vector<int32_t> ctl;
    PIS_TRY(gather_ctl(sv, ctl));
    std::vector<size_t> gone;
    for (size_t k = 0; k < c->sys.size(); ++k) {
        const int32_t* q = &ctl[k * CTL_COUNT];
        auto& j = sv->resident[k];
        pis_status st = PIS_OK;
        std::string msg;
        if (q[CTL_OVF_STAGE]) {
            st = PIS_ERR_CAPACITY;
            msg = "LDS staging capacity exceeded";
        } else if (q[CTL_OVF_NB]) {
            st = PIS_ERR_CAPACITY;
            msg = "neighbour capacity " + std::to_string(c->sys[k].plan.maxnb) + " exceeded";
        } else if (q[CTL_NAN]) {
            st = PIS_ERR_NAN_ENERGY;
            msg = "non-finite energy or position";
        } else if (q[CTL_LEFT] != 0) {
            continue;  // still running
        }
        if (st == PIS_OK) {
            pis_status cs = collect(sv, (uint32_t)k, q, *j);
            if (cs != PIS_OK) {
                st = cs;
                msg = last_error();
            }
        }
        gone.push_back(k);
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->used -= j->bytes;
        if (st == PIS_OK) sv->atom_steps += (uint64_t)c->sys[k].plan.n * j->nsteps;
        finish(sv, j, st, msg);
    }
    if (!gone.empty()) {
        PIS_TRY(ctx_remove(c, gone));
        for (size_t g = gone.size(); g-- > 0;) sv->resident.erase(sv->resident.begin() + (ptrdiff_t)gone[g]);
    }
    *nretired = (int)gone.size();
    return PIS_OK;
}

void worker_main(pis_server* sv)
{
    (void)hipSetDevice(sv->ctx->device);
    for (;;) {
        std::vector<std::shared_ptr<Job>> admit_now;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !sv->resident.empty(); });
            if (sv->stop) break;
            // arrival order: the head waits until it fits, later jobs wait behind it
            while (!sv->pending.empty() && sv->resident.size() + admit_now.size() < sv->max_systems &&
                   sv->used + sv->pending.front()->bytes <= sv->budget) {
                admit_now.push_back(sv->pending.front());
                sv->used += sv->pending.front()->bytes;
                sv->pending.pop_front();
            }
            sv->peak_bytes = std::max(sv->peak_bytes, sv->used);
        }
        for (auto& j : admit_now) {
            j->t_admit = sv->now();
            sv->resident.push_back(j);
            pis_status st = admit(sv, j);
            if (st != PIS_OK) {
                const std::string msg = last_error();
                // drop the half-admitted system (if it was appended)
                if (sv->ctx->sys.size() == sv->resident.size()) (void)ctx_remove(sv->ctx, {sv->resident.size() - 1});
                sv->resident.pop_back();
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->used -= j->bytes;
                finish(sv, j, st, msg);
            }
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->n_resident = sv->resident.size();
            sv->peak_resident = std::max(sv->peak_resident, (uint32_t)sv->resident.size());
        }
        if (sv->resident.empty()) continue;
        pis_status st = run_round(sv, (int)sv->round_steps);
        int nret = 0;
        if (st == PIS_OK) st = retire(sv, &nret);
        if (st != PIS_OK) {
            // a device-level failure ends every resident and queued job
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->fatal = st;
            sv->fatal_msg = last_error();
            for (auto& j : sv->resident) finish(sv, j, st, sv->fatal_msg);
            for (auto& j : sv->pending) finish(sv, j, st, sv->fatal_msg);
            sv->resident.clear();
            sv->pending.clear();
            sv->stop = true;
            break;
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        ++sv->rounds;
        sv->n_resident = sv->resident.size();
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
    pis_ctx* c = nullptr;
    PIS_TRY(ctx_create_empty(precision == PIS_PRECISION_MIXED, &c));
    c->use_graph = true;
    pis_server* sv = new pis_server();
    sv->ctx = c;
    sv->budget = hbm_budget_bytes;
    sv->max_systems = max_systems;
    sv->round_steps = round_steps;
    sv->mixed = precision == PIS_PRECISION_MIXED;
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
    j->bytes = system_bytes(j->plan, (int)sv->round_steps);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return fail(sv->fatal != PIS_OK ? sv->fatal : PIS_ERR_INVALID_ARGUMENT, "server stopped: " + sv->fatal_msg);
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
        for (auto& j : sv->resident)
            if (!j->done) finish(sv, j, PIS_ERR_INVALID_ARGUMENT, "server destroyed");
    }
    if (sv->d_ctl) (void)hipFree(sv->d_ctl);
    free_ctx(sv->ctx);
    delete sv;
    return PIS_OK;
}

}  // extern "C"
