// sph_server.cpp -- the serving loop of libsph_hip.so (sph_server_* in
// include/sph_hip.h): continuous batching of "advance scene S by k steps" jobs
// over one batch context.
//
// One scheduler thread owns the context and makes every HIP call. Each round it
// (1) admits queued jobs in arrival order while the pool has a free range for
// them; a job that does not fit yet may be passed by later jobs that fit
// (backfill) until it has been passed max_bypass times, after which admission
// waits for it, so no job starves; (2) advances all running jobs together by
// min(chunk, fewest steps left) steps and (3) retires the jobs whose steps are
// done: reads their state back by id, frees their slots. A job is stepped
// exactly `steps` times in the batch, and a scene of a batch computes what it
// computes alone, so its result does not depend on the other jobs.
//
// Jobs are shared between the job table, the queue, the running list and any
// thread waiting for them, so a job collected by one thread stays valid for
// another that is still waiting on it.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sph_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    sph_scene_desc desc{};
    std::vector<float> pos, vel;   // initial state until admitted, the result once done
    uint32_t n = 0;
    uint32_t steps = 0;
    bool has_emitter = false;
    sph_emitter em{};
    int state = SPH_JOB_QUEUED;
    uint32_t scene = 0;
    uint32_t remaining = 0;
    uint32_t bypassed = 0;          // later jobs admitted while this one did not fit
    double submit_s = 0, admit_s = 0, done_s = 0;
    std::string error;
};

using JobPtr = std::shared_ptr<Job>;

}  // namespace

namespace sph {
sph_status report(sph_status code, const std::string &msg);   // sets sph_last_error() (sph_capi.hip)
sph_status host_exception() noexcept;                          // the status of a caught C++ exception
}
using sph::report;

struct sph_server {
    sph_server_params prm{};
    sph_ctx *ctx = nullptr;
    Clock::time_point t0 = Clock::now();
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work;   // wakes the scheduler
    std::condition_variable cv_done;   // wakes waiters
    bool stop = false;
    std::string broken;   // why the scheduler stopped on a host error (empty: running)
    uint64_t next_id = 1;
    std::deque<JobPtr> queue;
    std::vector<JobPtr> running;
    std::unordered_map<uint64_t, JobPtr> jobs;
    sph_server_stats st{};
    uint32_t pool = 0;
    uint32_t max_bypass = 32;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void run();
    void run_loop();
    void fail_all(const std::string &why);   // with mu held
};

namespace {

// Admits queued jobs in arrival order while they fit. A job that does not fit
// yet (no free range of slots or no free scene id) is passed by the later jobs
// that do, each of which counts as one bypass of every job still waiting before
// it; once a waiting job has been bypassed max_bypass times nothing behind it is
// admitted until it fits.
void admit(sph_server *s, std::unique_lock<std::mutex> &lk)
{
    std::vector<JobPtr> waiting;   // jobs before the current one that did not fit
    size_t i = 0;
    while (i < s->queue.size()) {
        for (const JobPtr &w : waiting)
            if (w->bypassed >= s->max_bypass) return;   // it holds the line
        JobPtr j = s->queue[i];
        lk.unlock();
        uint32_t scene = 0;
        sph_status st = sph_scene_add(s->ctx, &j->desc, &scene);
        if (st == SPH_ERR_CAPACITY) {
            lk.lock();
            if (j->bypassed >= s->max_bypass) return;   // it holds the line
            waiting.push_back(j);
            ++i;
            continue;
        }
        const bool added = st == SPH_OK;
        std::string err;
        if (!st) st = sph_scene_upload(s->ctx, scene, j->pos.data(), j->vel.empty() ? nullptr : j->vel.data(),
                                       nullptr, j->n);
        if (!st && j->has_emitter) st = sph_scene_set_emitter(s->ctx, scene, &j->em);
        if (st) {
            err = sph_last_error();
            if (added) (void) sph_scene_remove(s->ctx, scene);
        }
        lk.lock();
        s->queue.erase(s->queue.begin() + (std::ptrdiff_t) i);
        if (st) {
            j->state = SPH_JOB_FAILED;
            j->error = err;
            j->done_s = s->now();
            s->st.failed++;
            s->cv_done.notify_all();
            continue;
        }
        for (const JobPtr &w : waiting) w->bypassed++;
        j->scene = scene;
        j->remaining = j->steps;
        j->state = SPH_JOB_RUNNING;
        j->admit_s = s->now();
        std::vector<float>().swap(j->pos);
        std::vector<float>().swap(j->vel);
        s->running.push_back(j);
    }
}

// Reads a finished job's particles back by id (the device writes them into the
// library's pinned buffer) and frees its slots. False if the slots could not be
// freed (the context is unusable).
bool retire(sph_server *s, const JobPtr &j)
{
    const float *pp = nullptr, *vp = nullptr;
    uint32_t n = 0;
    sph_status st = sph_scene_readback(s->ctx, j->scene, 1, &pp, &vp, &n);
    std::vector<float> pos, vel;
    if (!st) {
        pos.assign(pp, pp + 3 * (size_t) n);
        vel.assign(vp, vp + 3 * (size_t) n);
    }
    const std::string err = st ? sph_last_error() : "";
    const bool freed = sph_scene_remove(s->ctx, j->scene) == SPH_OK;
    std::lock_guard<std::mutex> g(s->mu);
    j->done_s = s->now();
    if (st) {
        j->state = SPH_JOB_FAILED;
        j->error = err;
        s->st.failed++;
    } else {
        j->pos.swap(pos);
        j->vel.swap(vel);
        j->n = n;
        j->state = SPH_JOB_DONE;
        s->st.completed++;
    }
    s->cv_done.notify_all();
    return freed;
}

}  // namespace

// Fails every queued and running job and refuses new ones (mu held).
void sph_server::fail_all(const std::string &why)
{
    broken = why;
    for (const JobPtr &j : queue) {
        j->state = SPH_JOB_FAILED;
        j->error = broken;
        j->done_s = now();
        st.failed++;
    }
    for (const JobPtr &j : running) {
        if (j->state != SPH_JOB_RUNNING) continue;
        j->state = SPH_JOB_FAILED;
        j->error = broken;
        j->done_s = now();
        st.failed++;
    }
    queue.clear();
    running.clear();
    cv_done.notify_all();
}

// The scheduler thread. A C++ exception in it (a host allocation failure) fails
// every queued and running job and stops the server instead of ending the process.
void sph_server::run()
{
    try {
        run_loop();
    } catch (...) {
        (void) sph::host_exception();
        std::lock_guard<std::mutex> g(mu);
        fail_all(std::string("the scheduler stopped: ") + sph_last_error());
    }
}

void sph_server::run_loop()
{
    std::unique_lock<std::mutex> lk(mu);
    const uint32_t chunk = prm.chunk_steps ? prm.chunk_steps : 8u;
    while (true) {
        cv_work.wait(lk, [&] { return stop || !queue.empty() || !running.empty(); });
        if (stop) break;
        const double t_adm = now();
        admit(this, lk);
        st.admin_s += now() - t_adm;
        if (running.empty()) {
            // the first queued job does not fit an empty pool: it never will
            if (!queue.empty()) {
                JobPtr j = queue.front();
                queue.pop_front();
                j->state = SPH_JOB_FAILED;
                j->error = "job does not fit the pool";
                j->done_s = now();
                st.failed++;
                cv_done.notify_all();
            }
            continue;
        }
        uint32_t k = chunk;
        for (const JobPtr &j : running) k = std::min(k, j->remaining);
        std::vector<JobPtr> batch = running;
        lk.unlock();
        uint64_t live = 0;
        for (const JobPtr &j : batch) {
            uint32_t n = 0;
            if (!sph_scene_count(ctx, j->scene, &n)) live += n;
        }
        const double t_a = now();
        sph_status s1 = sph_step(ctx, k, nullptr);
        if (!s1) s1 = sph_sync(ctx);
        const double t_b = now();
        const std::string err = s1 ? sph_last_error() : "";
        bool usable = true;
        if (s1) {
            // the chunk failed: its jobs fail and give their slots back
            for (const JobPtr &j : batch) usable = sph_scene_remove(ctx, j->scene) == SPH_OK && usable;
            std::lock_guard<std::mutex> g(mu);
            for (const JobPtr &j : batch) {
                j->state = SPH_JOB_FAILED;
                j->error = err;
                j->done_s = now();
                st.failed++;
            }
        } else {
            for (const JobPtr &j : batch) {
                j->remaining -= k;
                if (j->remaining == 0) usable = retire(this, j) && usable;
            }
        }
        const double t_c = now();
        uint32_t free_slots = 0;
        (void) sph_batch_info(ctx, nullptr, nullptr, &free_slots, nullptr);
        lk.lock();
        st.free_slots = free_slots;
        st.busy_s += t_b - t_a;
        st.admin_s += t_c - t_b;
        if (!s1) {
            st.batch_steps += k;
            st.particle_steps += live * k;
        }
        running.erase(std::remove_if(running.begin(), running.end(),
                                     [](const JobPtr &j) { return j->state != SPH_JOB_RUNNING; }),
                      running.end());
        cv_done.notify_all();
        if (!usable) {
            fail_all("the scheduler stopped: a failed job's slots could not be freed: " + err);
            break;
        }
    }
}

extern "C" {

sph_status sph_server_create(const sph_server_params *p, sph_server **out)
{
    try {
        if (!p || !out) return report(SPH_ERR_INVALID, "sph_server_create: null argument");
        *out = nullptr;
        if (p->max_jobs == 0) return report(SPH_ERR_INVALID, "sph_server_create: max_jobs must be > 0");
        const uint64_t per = sph_slot_bytes(p->max_neighbours);
        const uint64_t slots = std::min<uint64_t>(p->hbm_budget / per, 1ull << 31);
        if (slots < 256) return report(SPH_ERR_INVALID, "sph_server_create: budget below 256 slots");
        auto s = std::make_unique<sph_server>();
        s->prm = *p;
        sph_batch_params bp{};
        bp.pool = (uint32_t) slots;
        bp.max_scenes = p->max_jobs;
        bp.flags = p->flags;
        bp.max_neighbours = p->max_neighbours;
        sph_status st = sph_batch_create(&bp, nullptr, 0, &s->ctx);
        if (st) return st;
        if (p->variant >= 0 && (st = sph_set_variant(s->ctx, p->variant, 0))) {
            sph_destroy(s->ctx);
            return st;
        }
        if (p->graphs) (void) sph_set_graphs(s->ctx, 1);
        s->pool = (uint32_t) slots;
        s->max_bypass = p->max_bypass ? p->max_bypass : 32u;
        s->st.pool = s->pool;
        s->st.free_slots = s->pool;
        sph_server *raw = s.release();
        raw->worker = std::thread([raw] { raw->run(); });
        *out = raw;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

void sph_server_destroy(sph_server *s)
{
    try {
        if (!s) return;
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        if (s->worker.joinable()) s->worker.join();
        sph_destroy(s->ctx);
        delete s;
    } catch (...) {
        (void) sph::host_exception();
    }
}

sph_status sph_server_submit(sph_server *s, const sph_job_desc *d, uint64_t *job_id)
{
    try {
        if (!s || !d || !job_id) return report(SPH_ERR_INVALID, "sph_server_submit: null argument");
        if (d->n > 0 && !d->pos3) return report(SPH_ERR_INVALID, "sph_server_submit: null positions");
        if (d->n > d->scene.capacity) return report(SPH_ERR_INVALID, "sph_server_submit: n above the capacity");
        if (d->steps == 0) return report(SPH_ERR_INVALID, "sph_server_submit: steps must be > 0");
        if ((uint64_t) d->scene.capacity > s->pool)
            return report(SPH_ERR_CAPACITY, "sph_server_submit: capacity " + std::to_string(d->scene.capacity) +
                                                " above the pool of " + std::to_string(s->pool) + " slots");
        auto j = std::make_shared<Job>();
        j->desc = d->scene;
        j->n = d->n;
        j->steps = d->steps;
        j->pos.assign(d->pos3, d->pos3 + 3 * (size_t) d->n);
        if (d->vel3) j->vel.assign(d->vel3, d->vel3 + 3 * (size_t) d->n);
        if (d->emitter) {
            j->has_emitter = true;
            j->em = *d->emitter;
        }
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->broken.empty()) return report(SPH_ERR_STATE, "sph_server_submit: " + s->broken);
        j->id = s->next_id++;
        j->submit_s = s->now();
        *job_id = j->id;
        s->queue.push_back(j);
        s->jobs[j->id] = j;
        s->st.submitted++;
        s->cv_work.notify_all();
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

static void fill_info(const Job &j, sph_job_info *info)
{
    info->state = j.state;
    info->count = j.n;
    info->submit_s = j.submit_s;
    info->admit_s = j.admit_s;
    info->done_s = j.done_s;
}

sph_status sph_server_job(sph_server *s, uint64_t id, sph_job_info *info)
{
    try {
        if (!s || !info) return report(SPH_ERR_INVALID, "sph_server_job: null argument");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return report(SPH_ERR_INVALID, "sph_server_job: unknown job " + std::to_string(id));
        fill_info(*it->second, info);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_server_wait(sph_server *s, uint64_t id, double timeout_s, sph_job_info *info)
{
    try {
        if (!s) return report(SPH_ERR_INVALID, "sph_server_wait: null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return report(SPH_ERR_INVALID, "sph_server_wait: unknown job " + std::to_string(id));
        const JobPtr j = it->second;   // stays valid if another thread collects the job meanwhile
        auto done = [&] { return j->state == SPH_JOB_DONE || j->state == SPH_JOB_FAILED; };
        if (timeout_s < 0)
            s->cv_done.wait(lk, done);
        else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
            return report(SPH_ERR_STATE, "sph_server_wait: timed out");
        if (info) fill_info(*j, info);
        if (j->state == SPH_JOB_FAILED) return report(SPH_ERR_FAULT, "job " + std::to_string(id) + " failed: " + j->error);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_server_result(sph_server *s, uint64_t id, float *pos3, float *vel3, uint32_t n)
{
    try {
        if (!s) return report(SPH_ERR_INVALID, "sph_server_result: null server");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return report(SPH_ERR_INVALID, "sph_server_result: unknown job " + std::to_string(id));
        Job &j = *it->second;
        if (j.state == SPH_JOB_FAILED) {
            const std::string why = "job " + std::to_string(id) + " failed: " + j.error;
            s->jobs.erase(it);
            return report(SPH_ERR_FAULT, why);
        }
        if (j.state != SPH_JOB_DONE) return report(SPH_ERR_STATE, "sph_server_result: job not done");
        if (n != j.n) return report(SPH_ERR_INVALID, "sph_server_result: n = " + std::to_string(n) + ", job holds " +
                                                        std::to_string(j.n));
        if (pos3) std::memcpy(pos3, j.pos.data(), j.pos.size() * sizeof(float));
        if (vel3) std::memcpy(vel3, j.vel.data(), j.vel.size() * sizeof(float));
        s->jobs.erase(it);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_server_stats_get(sph_server *s, sph_server_stats *out)
{
    try {
        if (!s || !out) return report(SPH_ERR_INVALID, "sph_server_stats_get: null argument");
        std::lock_guard<std::mutex> g(s->mu);
        *out = s->st;
        out->queued = s->queue.size();
        out->running = s->running.size();
        out->elapsed_s = s->now();
        out->pool = s->pool;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

}  // extern "C"
