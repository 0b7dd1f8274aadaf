// Continuous-batching server over a pool (include/hipgran.h, "Server").
//
// Producers submit requests; the host work of each (bond search, pair-state
// rows, state records) runs in hg_server_submit on the producer's thread.
// One worker thread owns the pool and advances it one rebuild interval per
// round (at most `quantum` steps, fewer when a request finishes sooner):
//   1. the interval's launches are queued for the resident requests
//      (pool_begin) and the worker does not wait for them;
//   2. while they run, on the pool's copy stream: the requests that finished
//      in the previous round are read back and their ranges freed, and the
//      queued requests are uploaded in arrival order as long as the head fits
//      the arena and a slot is free (they join the launch order next round);
//   3. the worker waits for the interval (pool_end). A request whose system
//      faulted (NaN/Inf, or a table overflow) fails alone with HG_ERR_NUMERIC
//      or HG_ERR_CAPACITY; a request that has run all its steps leaves the
//      launch order, to be read back during the next round.
// Spring state stays on the device for the whole life of a request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hg_internal.h"
#include "pool_internal.h"

namespace {

using clock_type = std::chrono::steady_clock;

struct request {
    uint64_t id = 0;
    size_t n = 0;
    hg::pool_prep prep;               // host work done at submission
    std::vector<double> x, v, th, w;  // the final state
    std::vector<double> verts, vfacet;
    size_t nf = 0;
    uint64_t nsteps = 0;
    int state = HG_REQ_QUEUED;
    hg_status status = HG_OK;  // why it failed
    int32_t slot = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::string error;
};

}  // namespace

struct hg_server {
    hg_params params{};
    hg_server_config cfg{};
    hg_pool *pool = nullptr;
    size_t max_particles = 0;
    int device = 0;
    clock_type::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work;  // worker: new submissions or stop
    std::condition_variable cv_done;  // waiters: a request finished
    std::deque<uint64_t> queue;       // queued ids, arrival order
    std::map<uint64_t, request> reqs;  // nodes stay put: the worker keeps pointers to live requests
    uint64_t next_id = 1;
    bool stop = false;
    std::string fatal;  // set when the worker stopped on an error
    hg_server_stats stats{};
    size_t tuned_particles = 0;  // resident particles when the pool was last tuned
    std::thread worker;

    double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
    void run();
    void stop_all(const std::string &why, const std::vector<request *> &resident);
};

// Fails every request not yet finished, after an error the pool cannot go on
// from (a HIP error); the caller holds no lock.
void hg_server::stop_all(const std::string &why, const std::vector<request *> &resident) {
    std::lock_guard<std::mutex> lk(mu);
    fatal = why;
    for (request *r : resident) {
        r->state = HG_REQ_FAILED;
        r->status = HG_ERR_HIP;
        r->error = why;
    }
    for (uint64_t id : queue) {
        request &r = reqs[id];
        r.state = HG_REQ_FAILED;
        r.status = HG_ERR_HIP;
        r.error = why;
    }
    queue.clear();
    cv_done.notify_all();
}

void hg_server::run() {
    if (hipSetDevice(device) != hipSuccess) {
        stop_all("hipSetDevice failed in the server worker", {});
        return;
    }
    const hipStream_t cs = hg::pool_copy_stream(pool);
    std::vector<request *> running, retiring, admitted;
    auto all_resident = [&] {
        std::vector<request *> v(running);
        v.insert(v.end(), retiring.begin(), retiring.end());
        v.insert(v.end(), admitted.begin(), admitted.end());
        return v;
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !running.empty() || !retiring.empty(); });
            if (stop) return;
        }
        // 1. the interval of the resident requests, queued
        uint32_t len = 0;
        if (!running.empty()) {
            uint64_t max_len = cfg.quantum;
            for (request *r : running) {
                size_t n = 0;
                double t = 0;
                uint64_t k = 0;
                (void)hg_pool_info(pool, r->slot, &n, &t, &k);
                max_len = std::min<uint64_t>(max_len, r->nsteps - k);
            }
            hg::pool_reset_profile(pool);
            if (hg::pool_begin(pool, cfg.dt, max_len, cfg.rebuild_every, &len) != HG_OK) {
                stop_all(hg_last_error(), all_resident());
                return;
            }
        }
        // 2a. while it runs: the requests that finished last round
        const bool any_retired = !retiring.empty();
        for (request *r : retiring) {
            hg_status st = hg::pool_read_state(pool, r->slot, r->x.data(), r->v.data(), r->th.data(), r->w.data(), cs);
            const std::string err = st == HG_OK ? std::string() : hg_last_error();
            if (hg::pool_release(pool, r->slot, cs) != HG_OK && st == HG_OK) {
                stop_all(hg_last_error(), all_resident());
                return;
            }
            std::lock_guard<std::mutex> lk(mu);
            r->slot = -1;
            r->t_done = now();
            if (st == HG_OK) {
                r->state = HG_REQ_DONE;
                stats.done++;
            } else {
                r->state = HG_REQ_FAILED;
                r->status = st;
                r->error = err;
                stats.failed++;
            }
        }
        retiring.clear();
        if (any_retired) cv_done.notify_all();
        // 2b. admission in arrival order while the head fits
        for (;;) {
            request *r = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (queue.empty()) break;
                r = &reqs[queue.front()];
            }
            int32_t slot = -1;
            hg_status st = hg::pool_commit(pool, r->prep, &slot, cs);
            std::string err = st == HG_OK ? std::string() : hg_last_error();
            // no room now: wait for a retirement, unless nothing is resident
            // (then the request can never be admitted and fails)
            const bool resident = !running.empty() || !admitted.empty() || len;
            if (st == HG_ERR_CAPACITY && resident) break;
            if (st == HG_OK && r->nf) {
                st = hg::pool_set_facets(pool, slot, r->nf, r->verts.data(), r->vfacet.data(), cs);
                if (st != HG_OK) {
                    err = hg_last_error();
                    (void)hg::pool_release(pool, slot, cs);
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            queue.pop_front();
            if (st != HG_OK) {
                r->state = HG_REQ_FAILED;
                r->status = st;
                r->error = err;
                r->t_done = now();
                stats.failed++;
                cv_done.notify_all();
                continue;
            }
            r->prep = hg::pool_prep();  // the host copies are on the device now
            r->slot = slot;
            r->state = HG_REQ_RUNNING;
            r->t_admit = now();
            admitted.push_back(r);
        }
        // 3. the interval's end: faults, finished requests
        if (len) {
            std::vector<int32_t> faulted;
            double busy = 0.0;
            if (hg::pool_end(pool, faulted) != HG_OK || hg_pool_last_busy(pool, &busy) != HG_OK) {
                stop_all(hg_last_error(), all_resident());
                return;
            }
            std::vector<request *> still;
            std::vector<int32_t> release;
            {
                std::lock_guard<std::mutex> lk(mu);
                stats.rounds++;
                stats.device_ms += busy;
                for (request *r : running) {
                    stats.particle_steps += uint64_t(r->n) * len;
                    int fault = 0;
                    (void)hg_pool_fault(pool, r->slot, &fault);
                    if (fault) {
                        r->state = HG_REQ_FAILED;
                        r->status = (fault & HG_FAULT_NUMERIC) ? HG_ERR_NUMERIC : HG_ERR_CAPACITY;
                        r->error = (fault & HG_FAULT_NUMERIC)
                                       ? "the request's state became NaN or infinite"
                                       : "the request overflowed its neighbour list or pair-state table";
                        r->t_done = now();
                        stats.failed++;
                        release.push_back(r->slot);
                        r->slot = -1;
                        continue;
                    }
                    size_t n = 0;
                    double t = 0;
                    uint64_t k = 0;
                    (void)hg_pool_info(pool, r->slot, &n, &t, &k);
                    if (k >= r->nsteps) {
                        hg::pool_park(pool, r->slot);
                        retiring.push_back(r);
                    } else {
                        still.push_back(r);
                    }
                }
            }
            running.swap(still);
            for (int32_t s : release)
                if (hg::pool_release(pool, s, cs) != HG_OK) {
                    stop_all(hg_last_error(), all_resident());
                    return;
                }
            if (!release.empty()) cv_done.notify_all();
        }
        running.insert(running.end(), admitted.begin(), admitted.end());
        admitted.clear();
        {
            std::lock_guard<std::mutex> lk(mu);
            stats.queued = queue.size();
            stats.running = running.size() + retiring.size();
            // re-tune the pool's launch split when the resident particle
            // count has moved by half since the last tuning (the split never
            // changes results, only speed)
            size_t resident = 0;
            for (request *r : running) resident += r->n;
            if (resident && (resident > tuned_particles + tuned_particles / 2 || 2 * resident < tuned_particles)) {
                (void)hg_pool_autotune(pool, 1);
                tuned_particles = resident;
            }
        }
    }
}

#define HG_TRY_S(expr)                \
    do {                              \
        hg_status st_ = (expr);       \
        if (st_ != HG_OK) return st_; \
    } while (0)

extern "C" {

hg_status hg_server_create(const hg_params *params, const hg_server_config *cfg, hg_server **out) {
    using namespace hg;
    if (!params || !cfg || !out) return fail(HG_ERR_INVALID, "hg_server_create: params, cfg and out are required");
    *out = nullptr;
    if (cfg->quantum == 0 || !(cfg->dt > 0.0) || cfg->max_requests == 0)
        return fail(HG_ERR_INVALID, "hg_server_create: quantum, dt and max_requests must be positive");
    const size_t bpp = hg_pool_bytes_per_particle(params, cfg->max_facets);
    if (bpp == 0) return fail(HG_ERR_INVALID, "hg_server_create: invalid params");
    const size_t max_particles = size_t(cfg->budget_bytes / bpp);
    if (max_particles == 0) return fail(HG_ERR_INVALID, "hg_server_create: budget_bytes below one particle");
    auto *S = new hg_server();
    S->params = *params;
    S->cfg = *cfg;
    S->max_particles = max_particles;
    S->t0 = clock_type::now();
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return fail(HG_ERR_NO_DEVICE, "hg_server_create: no current HIP device");
    }
    hg_status st = hg_pool_create(params, max_particles, cfg->max_requests, cfg->max_facets, &S->pool);
    if (st != HG_OK) {
        delete S;
        return st;
    }
    S->worker = std::thread([S] { S->run(); });
    *out = S;
    return HG_OK;
}

hg_status hg_server_destroy(hg_server *S) {
    if (!S) return HG_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (S->pool) hg_pool_destroy(S->pool);
    delete S;
    return HG_OK;
}

hg_status hg_server_submit(hg_server *S, size_t n, const double *x0, const double *v0, const double *theta0,
                           const double *omega0, size_t nf, const double *verts, const double *v_facet,
                           uint64_t nsteps, uint64_t *id) {
    using namespace hg;
    if (!S || !x0 || !id) return fail(HG_ERR_INVALID, "hg_server_submit: server, x0 and id are required");
    if (n == 0 || nsteps == 0) return fail(HG_ERR_INVALID, "hg_server_submit: empty request");
    if (nf > S->cfg.max_facets || (nf && (!verts || !v_facet)))
        return fail(HG_ERR_INVALID, "hg_server_submit: bad facets");
    if (n > S->max_particles)
        return fail(HG_ERR_CAPACITY, "hg_server_submit: " + std::to_string(n) + " particles exceed the budget of " +
                                         std::to_string(S->max_particles));
    request r;
    r.n = n;
    r.nsteps = nsteps;
    // the host work of the admission, here on the producer's thread
    HG_TRY_S(pool_prepare(S->params, n, x0, v0, theta0, omega0, nullptr, 0, r.prep));
    for (auto *f : {&r.x, &r.v, &r.th, &r.w}) f->assign(3 * n, 0.0);
    r.nf = nf;
    if (nf) {
        r.verts.assign(verts, verts + 9 * nf);
        r.vfacet.assign(v_facet, v_facet + 3 * nf);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->fatal.empty()) return fail(HG_ERR_HIP, "hg_server_submit: server stopped: " + S->fatal);
        r.id = S->next_id++;
        r.t_submit = S->now();
        *id = r.id;
        S->queue.push_back(r.id);
        S->reqs.emplace(r.id, std::move(r));
        S->stats.submitted++;
        S->stats.queued = S->queue.size();
    }
    S->cv_work.notify_one();
    return HG_OK;
}

hg_status hg_server_status(hg_server *S, uint64_t id, int *state) {
    using namespace hg;
    if (!S || !state) return fail(HG_ERR_INVALID, "hg_server_status: server and state are required");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HG_ERR_INVALID, "hg_server_status: unknown or released request");
    *state = it->second.state;
    return HG_OK;
}

hg_status hg_server_wait(hg_server *S, uint64_t id, double timeout_s, double *x, double *v, double *theta,
                         double *omega, double *times) {
    using namespace hg;
    if (!S) return fail(HG_ERR_INVALID, "hg_server_wait: server is null");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HG_ERR_INVALID, "hg_server_wait: unknown or released request");
    auto finished = [&] {
        int s = it->second.state;
        return s == HG_REQ_DONE || s == HG_REQ_FAILED || !S->fatal.empty();
    };
    if (timeout_s < 0) {
        S->cv_done.wait(lk, finished);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) {
        return fail(HG_ERR_CAPACITY, "hg_server_wait: timed out");
    }
    request &r = it->second;
    if (r.state != HG_REQ_DONE) {
        const std::string e = r.error.empty() ? S->fatal : r.error;
        const hg_status st = r.status != HG_OK ? r.status : HG_ERR_HIP;
        S->reqs.erase(it);
        return fail(st, "hg_server_wait: request failed: " + e);
    }
    const size_t bytes = 3 * r.n * sizeof(double);
    if (x) std::memcpy(x, r.x.data(), bytes);
    if (v) std::memcpy(v, r.v.data(), bytes);
    if (theta) std::memcpy(theta, r.th.data(), bytes);
    if (omega) std::memcpy(omega, r.w.data(), bytes);
    if (times) {
        times[0] = r.t_submit;
        times[1] = r.t_admit;
        times[2] = r.t_done;
    }
    S->reqs.erase(it);
    return HG_OK;
}

hg_status hg_server_get_stats(hg_server *S, hg_server_stats *out) {
    using namespace hg;
    if (!S || !out) return fail(HG_ERR_INVALID, "hg_server_get_stats: server and out are required");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->elapsed_s = S->now();
    return HG_OK;
}

}  // extern "C"
