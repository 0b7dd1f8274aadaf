// Continuous-batching server over a pool (include/hipgran.h, "Server").
//
// One worker thread owns the pool. Each round it admits queued requests in
// arrival order while the head fits the arena and a system slot is free,
// advances every resident request by one quantum (one hg_pool_step: a
// persistent kernel launch per rebuild interval), and retires the requests
// that have run their steps: final state copied to host memory, arena range
// and slot freed for the next request. Submissions from the producer thread
// overlap with the device work of earlier requests.
#include <hip/hip_runtime.h>

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

namespace {

using clock_type = std::chrono::steady_clock;

struct request {
    uint64_t id = 0;
    size_t n = 0;
    std::vector<double> x, v, th, w;  // inputs, then the final state
    std::vector<double> verts, vfacet;
    size_t nf = 0;
    uint64_t nsteps = 0;
    int state = HG_REQ_QUEUED;
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
    std::map<uint64_t, request> reqs;
    std::vector<uint64_t> running;
    uint64_t next_id = 1;
    bool stop = false;
    std::string fatal;  // set when the worker stopped on an error
    hg_server_stats stats{};
    size_t tuned_particles = 0;  // resident particles when the pool was last tuned
    std::thread worker;

    double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
    void run();
};

void hg_server::run() {
    if (hipSetDevice(device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = "hipSetDevice failed in the server worker";
        cv_done.notify_all();
        return;
    }
    for (;;) {
        // admission: strict arrival order while the head fits
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !running.empty(); });
            if (stop) return;
            while (!queue.empty()) {
                request &r = reqs[queue.front()];
                int32_t slot = -1;
                hg_status st = hg_pool_add(pool, r.n, r.x.data(), r.v.data(), r.th.data(), r.w.data(), &slot);
                // no room now: wait for a retirement, unless nothing is resident
                // (then the request can never be admitted and fails)
                if (st == HG_ERR_CAPACITY && !running.empty()) break;
                if (st == HG_OK && r.nf) st = hg_pool_set_facets(pool, slot, r.nf, r.verts.data(), r.vfacet.data());
                queue.pop_front();
                if (st != HG_OK) {
                    r.state = HG_REQ_FAILED;
                    r.error = hg_last_error();
                    if (slot >= 0) (void)hg_pool_remove(pool, slot);
                    stats.failed++;
                    cv_done.notify_all();
                    continue;
                }
                r.slot = slot;
                r.state = HG_REQ_RUNNING;
                r.t_admit = now();
                running.push_back(r.id);
            }
            stats.queued = queue.size();
            stats.running = running.size();
            if (running.empty()) continue;
            // re-tune the pool's launch split when the resident particle count
            // has moved by half since the last tuning (the split never changes
            // results, only speed)
            size_t resident = 0;
            for (uint64_t id : running) resident += reqs[id].n;
            if (resident > tuned_particles + tuned_particles / 2 || 2 * resident < tuned_particles) {
                (void)hg_pool_autotune(pool, 1);
                tuned_particles = resident;
            }
        }
        // one quantum for every resident request (the lock is not held, so
        // submissions proceed meanwhile)
        hg_status st = hg_pool_step(pool, cfg.dt, cfg.quantum, cfg.rebuild_every);
        double chunk_ms = 0, rebuild_ms = 0;
        uint64_t launches = 0;
        if (st == HG_OK) st = hg_pool_last_step_profile(pool, &chunk_ms, &rebuild_ms, &launches, nullptr);
        std::lock_guard<std::mutex> lk(mu);
        if (st != HG_OK) {
            fatal = hg_last_error();
            for (uint64_t id : running) {
                reqs[id].state = HG_REQ_FAILED;
                reqs[id].error = fatal;
            }
            running.clear();
            cv_done.notify_all();
            return;
        }
        stats.rounds++;
        stats.device_ms += chunk_ms + rebuild_ms;
        std::vector<uint64_t> still;
        for (uint64_t id : running) {
            request &r = reqs[id];
            stats.particle_steps += uint64_t(r.n) * cfg.quantum;
            size_t n = 0;
            double t = 0;
            uint64_t k = 0;
            hg_status s2 = hg_pool_info(pool, r.slot, &n, &t, &k);
            if (s2 == HG_OK && k < r.nsteps) {
                still.push_back(id);
                continue;
            }
            if (s2 == HG_OK) s2 = hg_pool_get_state(pool, r.slot, r.x.data(), r.v.data(), r.th.data(), r.w.data());
            (void)hg_pool_remove(pool, r.slot);
            r.slot = -1;
            r.t_done = now();
            if (s2 == HG_OK) {
                r.state = HG_REQ_DONE;
                stats.done++;
            } else {
                r.state = HG_REQ_FAILED;
                r.error = hg_last_error();
                stats.failed++;
            }
        }
        running.swap(still);
        stats.running = running.size();
        cv_done.notify_all();
    }
}

extern "C" {

hg_status hg_server_create(const hg_params *params, const hg_server_config *cfg, hg_server **out) {
    using namespace hg;
    if (!params || !cfg || !out) return fail(HG_ERR_INVALID, "hg_server_create: params, cfg and out are required");
    *out = nullptr;
    if (cfg->quantum == 0 || !(cfg->dt > 0.0) || cfg->max_requests == 0)
        return fail(HG_ERR_INVALID, "hg_server_create: quantum, dt and max_requests must be positive");
    if (params->pair_mode == HG_PAIRS_NEIGHBOR_LIST && cfg->rebuild_every && cfg->quantum % cfg->rebuild_every)
        return fail(HG_ERR_INVALID, "hg_server_create: quantum must be a multiple of rebuild_every");
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
    if (nsteps % S->cfg.quantum)
        return fail(HG_ERR_INVALID, "hg_server_submit: nsteps must be a multiple of the server quantum");
    if (nf > S->cfg.max_facets || (nf && (!verts || !v_facet)))
        return fail(HG_ERR_INVALID, "hg_server_submit: bad facets");
    if (n > S->max_particles)
        return fail(HG_ERR_CAPACITY, "hg_server_submit: " + std::to_string(n) + " particles exceed the budget of " +
                                         std::to_string(S->max_particles));
    request r;
    r.n = n;
    r.nsteps = nsteps;
    auto take = [n](std::vector<double> &dst, const double *src) {
        dst.assign(3 * n, 0.0);
        if (src) std::memcpy(dst.data(), src, 3 * n * sizeof(double));
    };
    take(r.x, x0);
    take(r.v, v0);
    take(r.th, theta0);
    take(r.w, omega0);
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
        std::string e = r.error.empty() ? S->fatal : r.error;
        S->reqs.erase(it);
        return fail(HG_ERR_HIP, "hg_server_wait: request failed: " + e);
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
