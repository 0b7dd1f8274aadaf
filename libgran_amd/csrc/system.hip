// C ABI of the single-system DEM path (include/hipgran.h): lifecycle, state
// transfer, stepping and inspection. Device work is queued on the system's own
// stream; entry points that return data synchronise that stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "hg_internal.h"
#include "snapshot.h"

namespace hg {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }
hg_status fail(hg_status st, const std::string &msg) {
    g_err = msg;
    return st;
}

namespace {

template <typename T>
hg_status dalloc(hg_system *S, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HG_ERR_OOM : HG_ERR_HIP,
                    std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes: " +
                        hipGetErrorString(e));
    S->allocations.push_back(*p);
    return HG_OK;
}

#define HG_TRY(expr)                   \
    do {                               \
        hg_status st_ = (expr);        \
        if (st_ != HG_OK) return st_;  \
    } while (0)

hg_status check_device() {
    int dev = 0;
    HG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HG_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(HG_ERR_NO_DEVICE, std::string("hipgran is built for gfx950; current device is ") +
                                          prop.gcnArchName);
    return HG_OK;
}

// AoS triples -> 4-double records (and back).
std::vector<double> pack4(const double *src, size_t n) {
    std::vector<double> out(4 * n, 0.0);
    if (src)
        for (size_t i = 0; i < n; i++) {
            out[4 * i] = src[3 * i];
            out[4 * i + 1] = src[3 * i + 1];
            out[4 * i + 2] = src[3 * i + 2];
        }
    return out;
}

hg_status download3(hg_system *S, const double *dev, double *host3) {
    if (!host3) return HG_OK;
    std::vector<double> tmp(4 * S->n);
    HG_HIP(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    for (size_t i = 0; i < S->n; i++) {
        host3[3 * i] = tmp[4 * i];
        host3[3 * i + 1] = tmp[4 * i + 1];
        host3[3 * i + 2] = tmp[4 * i + 2];
    }
    return HG_OK;
}

hg_status upload3(hg_system *S, double *dev, const double *host3) {
    std::vector<double> tmp = pack4(host3, S->n);
    HG_HIP(hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    return HG_OK;
}

uint32_t next_pow2(uint64_t v) {
    uint32_t p = 1;
    while (p < v && p < (1u << 30)) p <<= 1;
    return p;
}

hipError_t get_event(hg_system *S, hipEvent_t *e) {
    if (!S->event_pool.empty()) {
        *e = S->event_pool.back();
        S->event_pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(e);
}

// Queues `launch` on the system stream, bracketed by events when profiling.
template <typename F>
hipError_t timed(hg_system *S, int phase, F &&launch) {
    if (!S->profiling) return launch();
    hipEvent_t a, b;
    hipError_t e;
    if ((e = get_event(S, &a)) != hipSuccess) return e;
    if ((e = get_event(S, &b)) != hipSuccess) return e;
    if ((e = hipEventRecord(a, S->stream)) != hipSuccess) return e;
    if ((e = launch()) != hipSuccess) return e;
    if ((e = hipEventRecord(b, S->stream)) != hipSuccess) return e;
    S->pending.push_back({phase, a, b});
    return hipSuccess;
}

// After a stream synchronisation: fold finished ranges into the totals.
hg_status collect_profile(hg_system *S) {
    for (auto &r : S->pending) {
        float ms = 0.f;
        HG_HIP(hipEventElapsedTime(&ms, r.begin, r.end));
        S->prof_ms[r.phase] += double(ms);
        S->prof_launches[r.phase]++;
        S->event_pool.push_back(r.begin);
        S->event_pool.push_back(r.end);
    }
    S->pending.clear();
    return HG_OK;
}

// Device flags -> status. Clears them for the next call.
hg_status check_flags(hg_system *S, const char *where) {
    S->flags_unchecked = false;
    HG_HIP(hipMemcpyAsync(S->flags_host, S->flags, FLAG_COUNT * sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    if (!S->pending.empty()) {
        hg_status st = collect_profile(S);
        if (st != HG_OK) return st;
    }
    const int nl = S->flags_host[FLAG_NL_OVERFLOW], pr = S->flags_host[FLAG_PAIR_OVERFLOW];
    const int nlmax = S->flags_host[FLAG_NL_MAXCOUNT], prmax = S->flags_host[FLAG_PAIR_MAXCOUNT];
    const int numeric = S->flags_host[FLAG_NUMERIC];
    if (nl || pr || numeric) {
        HG_HIP(launch_zero_u32(reinterpret_cast<uint32_t *>(S->flags), FLAG_COUNT, S->stream));
        HG_HIP(hipStreamSynchronize(S->stream));
        // a truncated table is reported first: it can be what made the state non-finite
        if (nl)
            return fail(HG_ERR_CAPACITY, std::string(where) + ": neighbour list needs " + std::to_string(nlmax) +
                                             " slots per particle, capacity is " + std::to_string(S->cap_n));
        if (pr)
            return fail(HG_ERR_CAPACITY, std::string(where) + ": pair-state table needs " + std::to_string(prmax) +
                                             " slots per particle, capacity is " + std::to_string(S->cap_s));
        return fail(HG_ERR_NUMERIC, std::string(where) + ": a velocity or angular velocity became NaN or infinite");
    }
    return HG_OK;
}

// Reports the flags of steps hg_step_async left unchecked, before a call
// that reads results.
hg_status settle(hg_system *S, const char *where) {
    return S->flags_unchecked ? check_flags(S, where) : HG_OK;
}

// The steps of hg_step; `check` false leaves the flag read to the next call
// that reads results (hg_step_async).
hg_status step_impl(hg_system *S, double dt, uint64_t nsteps, uint32_t rebuild_every, bool check);

sys_view make_view(hg_system *S) {
    sys_view v;
    v.n = uint32_t(S->n);
    v.x = S->x;
    v.v = S->v;
    v.w = S->w;
    v.a = S->a;
    v.al = S->al;
    v.nb = S->nb;
    v.ncnt = S->ncnt;
    v.cap_n = S->cap_n;
    v.pt.pj_old = S->pj[S->cur];
    v.pt.ps_old = S->ps[S->cur];
    v.pt.pcnt_old = S->pcnt[S->cur];
    v.pt.pj_new = S->pj[S->cur ^ 1];
    v.pt.ps_new = S->ps[S->cur ^ 1];
    v.pt.pcnt_new = S->pcnt[S->cur ^ 1];
    v.pt.cap = S->cap_s;
    v.pt.stride = S->n;
    v.pt.bj = S->bj;
    v.pt.bstr = S->bstr;
    v.pt.bcnt = S->bcnt;
    v.pt.lconn = S->lconn;
    v.pt.lslot = S->lslot;
    v.facets = S->facets;
    v.ss = S->ss;
    v.diss = S->track_diss ? S->diss : nullptr;
    v.nf = S->nf;
    v.flags = S->flags;
    return v;
}

// One acceleration pass (queued; no synchronisation).
hg_status force_pass(hg_system *S) {
    if (S->p.binary_model == HG_BINARY_SINTER_LEGACY && !S->legacy_ready)
        return fail(HG_ERR_INVALID, "the legacy sinter_functor's connectors are not set (hg_set_legacy_connectors)");
    sys_view v = make_view(S);
    HG_HIP(timed(S, PHASE_FORCE, [&] { return launch_force(S->p, v, S->stream); }));
    if (S->p.binary_model != HG_BINARY_NONE) S->cur ^= 1;
    return HG_OK;
}

hg_status validate(const hg_params *p, size_t n) {
    if (!p) return fail(HG_ERR_INVALID, "params is null");
    if (n == 0 || n >= (size_t(1) << 31)) return fail(HG_ERR_INVALID, "n must be in [1, 2^31)");
    if (p->binary_model < HG_BINARY_NONE || p->binary_model > HG_BINARY_SINTER_LEGACY)
        return fail(HG_ERR_INVALID, "unknown binary_model");
    if (p->binary_model == HG_BINARY_SINTER_LEGACY && p->precision != HG_PRECISION_FP64)
        return fail(HG_ERR_INVALID, "the legacy sinter_functor runs in HG_PRECISION_FP64 only");
    if (p->hamaker_model < HG_HAMAKER_NONE || p->hamaker_model > HG_HAMAKER_SIMPLIFIED)
        return fail(HG_ERR_INVALID, "unknown hamaker_model");
    if (p->pair_mode != HG_PAIRS_ALL && p->pair_mode != HG_PAIRS_NEIGHBOR_LIST)
        return fail(HG_ERR_INVALID, "unknown pair_mode");
    if (p->pair_mode == HG_PAIRS_NEIGHBOR_LIST && !(p->r_verlet > 0.0))
        return fail(HG_ERR_INVALID, "neighbour-list mode needs r_verlet > 0");
    if (p->precision != HG_PRECISION_FP64 && p->precision != HG_PRECISION_MIXED)
        return fail(HG_ERR_INVALID, "unknown precision");
    if (p->precision == HG_PRECISION_MIXED && p->pair_mode != HG_PAIRS_NEIGHBOR_LIST)
        return fail(HG_ERR_INVALID, "a single system runs HG_PRECISION_MIXED in neighbour-list mode only; "
                                    "all-pairs systems run it through a pool");
    if (p->sinter_breaking != 0 && (p->sinter_breaking != 1 || p->binary_model != HG_BINARY_SINTER))
        return fail(HG_ERR_INVALID, "sinter_breaking is 0, or 1 with HG_BINARY_SINTER");
    return HG_OK;
}

}  // namespace
}  // namespace hg

using namespace hg;

extern "C" {

const char *hg_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace hg {
namespace {
// hg_create; `given_bonds` (with `given_skipped`) replaces the bond search on
// x0 (a system reloaded from a snapshot keeps the bonds it was built with).
hg_status create_impl(const hg_params *params, size_t n, const double *x0, const double *v0, const double *theta0,
                      const double *omega0, const std::vector<std::pair<uint32_t, uint32_t>> *given_bonds,
                      size_t given_skipped, hg_system **out) {
    if (!out) return fail(HG_ERR_INVALID, "out is null");
    *out = nullptr;
    HG_TRY(validate(params, n));
    if (!x0) return fail(HG_ERR_INVALID, "x0 is null");
    HG_TRY(check_device());
    hg_system *S = new hg_system();
    S->p = *params;
    S->n = n;
    HG_HIP(hipGetDevice(&S->device));
    auto cleanup_on = [&](hg_status st) {
        if (st != HG_OK) {
            std::string msg = g_err;
            hg_destroy(S);
            g_err = msg;
        }
        return st;
    };
#define HG_TRYC(expr)                              \
    do {                                           \
        hg_status st_ = (expr);                    \
        if (st_ != HG_OK) return cleanup_on(st_);  \
    } while (0)
#define HG_HIPC(call)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return cleanup_on(fail(HG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)));      \
    } while (0)

    HG_HIPC(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    double **fields[6] = {&S->x, &S->v, &S->th, &S->w, &S->a, &S->al};
    for (auto f : fields) HG_TRYC(dalloc(S, f, 4 * n));
    HG_TRYC(upload3(S, S->x, x0));
    HG_TRYC(upload3(S, S->v, v0));
    HG_TRYC(upload3(S, S->th, theta0));
    HG_TRYC(upload3(S, S->w, omega0));
    HG_HIPC(launch_zero_f64(S->a, 4 * n, S->stream));
    HG_HIPC(launch_zero_f64(S->al, 4 * n, S->stream));
    HG_TRYC(dalloc(S, &S->flags, FLAG_COUNT));
    HG_HIPC(hipHostMalloc(reinterpret_cast<void **>(&S->flags_host), FLAG_COUNT * sizeof(int)));
    HG_HIPC(hipMemsetAsync(S->flags, 0, FLAG_COUNT * sizeof(int), S->stream));
    HG_TRYC(dalloc(S, &S->red, 16));

    // Bond topology (alt_sinter ctor) and the initial pair-state table.
    std::vector<std::vector<uint32_t>> rows(n);
    if (S->p.binary_model == HG_BINARY_SINTER) {
        if (given_bonds) {
            S->bonds = *given_bonds;
            S->bonds_skipped = given_skipped;
        } else {
            sinter_topology_host(n, x0, S->p.sinter.r_part, S->p.sinter.d_crit, S->bonds, S->bonds_skipped,
                                 !S->p.sinter_breaking);
        }
        for (auto &b : S->bonds) {
            rows[b.first].push_back(b.second | kBondBit);
            rows[b.second].push_back(b.first | kBondBit);
        }
        for (auto &r : rows)
            std::sort(r.begin(), r.end(), [](uint32_t a, uint32_t b) { return (a & kIdxMask) < (b & kIdxMask); });
    }
    size_t max_bonds = 0;
    for (auto &r : rows) max_bonds = std::max(max_bonds, r.size());
    if (S->p.binary_model != HG_BINARY_NONE) {
        S->cap_s = S->p.pair_capacity ? S->p.pair_capacity : uint32_t((max_bonds + 24 + 3) / 4 * 4);
        S->cap_s = std::max<uint32_t>(S->cap_s, kTagPrefetch);
        if (S->cap_s < max_bonds)
            return cleanup_on(fail(HG_ERR_CAPACITY, "pair_capacity is below the largest bond count " +
                                                        std::to_string(max_bonds)));
        for (int b = 0; b < 2; b++) {
            HG_TRYC(dalloc(S, &S->pj[b], size_t(S->cap_s) * n));
            HG_TRYC(dalloc(S, &S->ps[b], size_t(S->cap_s) * kSpringRec * n));
            HG_TRYC(dalloc(S, &S->pcnt[b], n));
        }
        std::vector<uint32_t> pj(size_t(S->cap_s) * n, kNoEntry), cnt(n);
        for (size_t i = 0; i < n; i++) {
            cnt[i] = uint32_t(rows[i].size());
            for (size_t s = 0; s < rows[i].size(); s++) pj[s * n + i] = rows[i][s];
        }
        HG_HIPC(hipMemcpyAsync(S->pj[0], pj.data(), pj.size() * 4, hipMemcpyHostToDevice, S->stream));
        HG_HIPC(hipMemcpyAsync(S->pcnt[0], cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, S->stream));
        HG_HIPC(launch_zero_f64(S->ps[0], size_t(S->cap_s) * kSpringRec * n, S->stream));
        HG_HIPC(hipStreamSynchronize(S->stream));
    }
    if (S->p.binary_model == HG_BINARY_SINTER_LEGACY) {
        // sinter_functor ctor bond search (sinter_bridge.h:146-161): every
        // pair with |gap| <= critical_separation, no cycle check; the
        // connectors come with hg_set_legacy_connectors
        if (given_bonds) {
            S->bonds = *given_bonds;
            S->bonds_skipped = given_skipped;
        } else {
            sinter_topology_host(n, x0, S->p.contact.r_part, S->p.sinter.d_crit, S->bonds, S->bonds_skipped, false,
                                 true);
        }
        std::vector<std::vector<uint32_t>> part(n);
        for (auto &b : S->bonds) {
            part[b.first].push_back(b.second);
            part[b.second].push_back(b.first);
        }
        size_t mb = 0;
        for (auto &r : part) mb = std::max(mb, r.size());
        if (mb > kMaxBondsPerParticle)
            return cleanup_on(fail(HG_ERR_CAPACITY, "a particle has " + std::to_string(mb) +
                                                        " sinter bonds; at most " +
                                                        std::to_string(kMaxBondsPerParticle)));
        S->bcap = kMaxBondsPerParticle;
        std::vector<uint32_t> bj(size_t(S->bcap) * n, kNoEntry), cnt(n, 0), slot(size_t(S->bcap) * n, 0);
        for (size_t i = 0; i < n; i++) {  // rows come out ascending: bonds are in row order
            cnt[i] = uint32_t(part[i].size());
            for (size_t k = 0; k < part[i].size(); k++) bj[k * n + i] = part[i][k];
        }
        for (size_t i = 0; i < n; i++)
            for (size_t k = 0; k < part[i].size(); k++) {
                const uint32_t j = part[i][k];
                const auto it = std::lower_bound(part[j].begin(), part[j].end(), uint32_t(i));
                slot[k * n + i] = uint32_t(it - part[j].begin());
            }
        HG_TRYC(dalloc(S, &S->bj, size_t(S->bcap) * n));
        HG_TRYC(dalloc(S, &S->bcnt, n));
        HG_TRYC(dalloc(S, &S->lslot, size_t(S->bcap) * n));
        HG_TRYC(dalloc(S, &S->lconn, size_t(S->bcap) * kConnRec * n));
        HG_HIPC(hipMemcpyAsync(S->bj, bj.data(), bj.size() * 4, hipMemcpyHostToDevice, S->stream));
        HG_HIPC(hipMemcpyAsync(S->bcnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, S->stream));
        HG_HIPC(hipMemcpyAsync(S->lslot, slot.data(), slot.size() * 4, hipMemcpyHostToDevice, S->stream));
        HG_HIPC(launch_zero_f64(S->lconn, size_t(S->bcap) * kConnRec * n, S->stream));
        HG_HIPC(hipStreamSynchronize(S->stream));
        S->legacy_ready = S->bonds.empty();
    }
    if (S->p.binary_model == HG_BINARY_SINTER && S->p.sinter_breaking) {
        if (max_bonds > kMaxBondsPerParticle)
            return cleanup_on(fail(HG_ERR_CAPACITY, "a particle has " + std::to_string(max_bonds) +
                                                        " bonds; neck strengths hold at most " +
                                                        std::to_string(kMaxBondsPerParticle)));
        S->bcap = kMaxBondsPerParticle;
        HG_TRYC(dalloc(S, &S->bj, size_t(S->bcap) * n));
        HG_TRYC(dalloc(S, &S->bstr, size_t(S->bcap) * n));
        HG_TRYC(dalloc(S, &S->bcnt, n));
        std::vector<double> inf(S->bonds.size(), std::numeric_limits<double>::infinity());
        HG_TRYC(hg_set_neck_strengths(S, inf.data()));
    }

    if (S->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST) {
        // Default capacity: close-packing bound of spheres of radius r whose
        // centres lie within r_verlet, with 30% headroom, rounded up to 8.
        const double r = S->p.contact.r_part > 0 ? S->p.contact.r_part
                                                 : (S->p.sinter.r_part > 0 ? S->p.sinter.r_part : S->p.hamaker.r_part);
        uint64_t cap = S->p.neighbor_capacity;
        if (cap == 0) {
            double ratio = r > 0 ? S->p.r_verlet / r : 64.0;
            double bound = 0.7405 * (std::pow(ratio + 1.0, 3.0) - 1.0) * 1.3;
            cap = uint64_t(std::ceil(std::max(bound, 8.0) / 8.0) * 8.0);
        }
        cap = std::min<uint64_t>(cap, n > 1 ? n - 1 : 1);
        S->cap_n = uint32_t(std::max<uint64_t>(cap, 1));
        HG_TRYC(dalloc(S, &S->nb, size_t(S->cap_n) * n));
        HG_TRYC(dalloc(S, &S->ncnt, n));
        S->hash_size = next_pow2(std::max<uint64_t>(2 * n, 1024));
        HG_TRYC(dalloc(S, &S->cell_of, n));
        HG_TRYC(dalloc(S, &S->cell_items, n));
        HG_TRYC(dalloc(S, &S->bucket_count, S->hash_size));
        HG_TRYC(dalloc(S, &S->bucket_start, S->hash_size));
        HG_TRYC(dalloc(S, &S->bucket_fill, S->hash_size));
        HG_TRYC(dalloc(S, &S->scan_tmp, scan_tmp_words(S->hash_size)));
        HG_TRYC(dalloc(S, &S->bbox, 8));
        HG_HIPC(launch_neighbor_build(*S, S->stream));
        HG_TRYC(check_flags(S, "hg_create"));
    }
    HG_HIPC(hipStreamSynchronize(S->stream));
    *out = S;
    return HG_OK;
#undef HG_TRYC
#undef HG_HIPC
}
}  // namespace
}  // namespace hg

extern "C" {

hg_status hg_create(const hg_params *params, size_t n, const double *x0, const double *v0, const double *theta0,
                    const double *omega0, hg_system **out) {
    return create_impl(params, n, x0, v0, theta0, omega0, nullptr, 0, out);
}

hg_status hg_destroy(hg_system *S) {
    if (!S) return HG_OK;
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    for (void *p : S->allocations) (void)hipFree(p);
    for (auto &r : S->pending) {
        (void)hipEventDestroy(r.begin);
        (void)hipEventDestroy(r.end);
    }
    for (auto e : S->event_pool) (void)hipEventDestroy(e);
    if (S->flags_host) (void)hipHostFree(S->flags_host);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return HG_OK;
}

hg_status hg_set_facets(hg_system *S, size_t nf, const double *verts, const double *v_facet) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (nf > size_t(kMaxFacets)) return fail(HG_ERR_INVALID, "at most 16 facets are supported");
    if (nf > 0 && (!verts || !v_facet)) return fail(HG_ERR_INVALID, "verts and v_facet are required");
    std::vector<facet_dev> F(nf);
    for (size_t f = 0; f < nf; f++) {
        const double *q = verts + 9 * f;
        for (int k = 0; k < 3; k++) F[f].vert[k] = d3{q[3 * k], q[3 * k + 1], q[3 * k + 2]};
        F[f].v_facet = d3{v_facet[3 * f], v_facet[3 * f + 1], v_facet[3 * f + 2]};
        // triangular_facet.h:23-29
        d3 u = F[f].vert[1] - F[f].vert[0], w = F[f].vert[2] - F[f].vert[0];
        F[f].area = 1.0 / 2.0 * norm(cross(u, w));
        F[f].normal = normalized(cross(u, w));
        F[f].pad[0] = F[f].pad[1] = 0.0;
    }
    if (!S->facets) HG_TRY(dalloc(S, &S->facets, kMaxFacets));
    if (!S->ss && nf > 0) HG_TRY(dalloc(S, &S->ss, size_t(kMaxFacets) * 9 * S->n));
    if (nf > 0) {
        HG_HIP(hipMemcpyAsync(S->facets, F.data(), nf * sizeof(facet_dev), hipMemcpyHostToDevice, S->stream));
        HG_HIP(launch_zero_f64(S->ss, nf * 9 * S->n, S->stream));
    }
    HG_HIP(hipStreamSynchronize(S->stream));
    S->nf = int(nf);
    S->acc_valid = false;
    return HG_OK;
}

hg_status hg_get_facets(hg_system *S, double *verts, double *v_facet) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (S->nf == 0) return HG_OK;
    std::vector<facet_dev> F(S->nf);
    HG_HIP(hipMemcpyAsync(F.data(), S->facets, F.size() * sizeof(facet_dev), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    for (int f = 0; f < S->nf; f++) {
        for (int k = 0; k < 3; k++) {
            if (verts) {
                verts[9 * f + 3 * k] = F[f].vert[k].x;
                verts[9 * f + 3 * k + 1] = F[f].vert[k].y;
                verts[9 * f + 3 * k + 2] = F[f].vert[k].z;
            }
        }
        if (v_facet) {
            v_facet[3 * f] = F[f].v_facet.x;
            v_facet[3 * f + 1] = F[f].v_facet.y;
            v_facet[3 * f + 2] = F[f].v_facet.z;
        }
    }
    return HG_OK;
}

hg_status hg_advance_facet(hg_system *S, size_t f, const double *v_facet, double dt) {
    if (!S || !v_facet) return fail(HG_ERR_INVALID, "system and v_facet are required");
    if (f >= size_t(S->nf)) return fail(HG_ERR_INVALID, "no such facet");
    HG_HIP(launch_advance_facet(S->facets + f, v_facet, dt, S->stream));
    return HG_OK;
}

hg_status hg_rebuild_neighbors(hg_system *S) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (S->p.pair_mode != HG_PAIRS_NEIGHBOR_LIST) return fail(HG_ERR_INVALID, "system has no neighbour list");
    HG_HIP(launch_neighbor_build(*S, S->stream));
    return check_flags(S, "hg_rebuild_neighbors");
}

hg_status hg_compute_accelerations(hg_system *S) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(force_pass(S));
    S->acc_valid = true;
    return check_flags(S, "hg_compute_accelerations");
}

hg_status hg_step(hg_system *S, double dt, uint64_t nsteps, uint32_t rebuild_every) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_step"));
    return step_impl(S, dt, nsteps, rebuild_every, true);
}

hg_status hg_step_async(hg_system *S, double dt, uint64_t nsteps, uint32_t rebuild_every) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    return step_impl(S, dt, nsteps, rebuild_every, false);
}

}  // extern "C"

namespace hg {
namespace {
hg_status step_impl(hg_system *S, double dt, uint64_t nsteps, uint32_t rebuild_every, bool check) {
    if (nsteps == 0) return HG_OK;
    if (!S->acc_valid) {
        HG_TRY(force_pass(S));
        S->acc_valid = true;
    }
    const double half = dt / 2.0;
    const uint32_t n = uint32_t(S->n);
    const bool nl = S->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST;
    const int nf_advance = S->p.facet_advance ? 0 : S->nf;  // facets the step moves itself
    if (n <= kSmallSystemMax && S->p.precision == HG_PRECISION_FP64 && !S->profiling && !S->per_step &&
        !S->track_diss) {
        // small systems: the steps between list rebuilds in one launch of one
        // workgroup instead of three dependent launches per step
        if (S->p.binary_model == HG_BINARY_SINTER_LEGACY && !S->legacy_ready)
            return fail(HG_ERR_INVALID, "the legacy sinter_functor's connectors are not set (hg_set_legacy_connectors)");
        if (n <= kTinySystemMax && !S->tiny) HG_TRY(dalloc(S, &S->tiny, size_t(n) * std::max(n, S->cap_n) * kSpringRec));
        uint64_t s = 0;
        while (s < nsteps) {
            if (nl && rebuild_every && (S->step_index % rebuild_every) == 0)
                HG_HIP(launch_neighbor_build(*S, S->stream));
            uint64_t len = nsteps - s;
            if (nl && rebuild_every) len = std::min<uint64_t>(len, rebuild_every - S->step_index % rebuild_every);
            len = std::min<uint64_t>(len, uint64_t(1) << 30);
            const sys_view even = make_view(S);
            S->cur ^= 1;
            const sys_view odd = make_view(S);
            S->cur ^= 1;
            HG_HIP(launch_small_steps(S->p, even, odd, S->x, S->v, S->th, S->w, S->facets, uint32_t(len), s > 0,
                                      s + len == nsteps, half, dt, nf_advance, S->lconn, S->bcnt, S->tiny,
                                      S->stream));
            if (S->p.binary_model != HG_BINARY_NONE) S->cur ^= int(len & 1u);
            for (uint64_t k = 0; k < len; k++) S->t = S->t + dt;
            S->step_index += len;
            s += len;
        }
        if (!check) {
            S->flags_unchecked = true;
            return HG_OK;
        }
        return check_flags(S, "hg_step");
    }
    for (uint64_t s = 0; s < nsteps; s++) {
        if (nl && rebuild_every && (S->step_index % rebuild_every) == 0)
            HG_HIP(timed(S, PHASE_REBUILD, [&] { return launch_neighbor_build(*S, S->stream); }));
        HG_HIP(timed(S, PHASE_KICK, [&] {
            hipError_t e = launch_kick_drift(S->x, S->v, S->th, S->w, S->a, S->al, n, s > 0, half, half, dt,
                                             S->facets, nf_advance, dt, S->stream);
            if (e == hipSuccess && S->lconn) e = launch_legacy_move(S->x, S->v, S->w, n, dt, S->lconn, S->bcnt, S->stream);
            return e;
        }));
        HG_TRY(force_pass(S));
        S->t = S->t + dt;
        S->step_index++;
    }
    HG_HIP(timed(S, PHASE_KICK, [&] {
        return launch_final_kick(S->v, S->w, S->a, S->al, n, half, S->facets, nf_advance, dt, S->flags, S->stream);
    }));
    if (!check) {
        S->flags_unchecked = true;
        return HG_OK;
    }
    return check_flags(S, "hg_step");
}
}  // namespace
}  // namespace hg

extern "C" {

hg_status hg_get_state(hg_system *S, double *x, double *v, double *theta, double *omega) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_get_state"));
    HG_TRY(download3(S, S->x, x));
    HG_TRY(download3(S, S->v, v));
    HG_TRY(download3(S, S->th, theta));
    HG_TRY(download3(S, S->w, omega));
    return HG_OK;
}

hg_status hg_set_state(hg_system *S, const double *x, const double *v, const double *theta, const double *omega) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_set_state"));
    if (x) HG_TRY(upload3(S, S->x, x));
    if (v) HG_TRY(upload3(S, S->v, v));
    if (theta) HG_TRY(upload3(S, S->th, theta));
    if (omega) HG_TRY(upload3(S, S->w, omega));
    S->acc_valid = false;
    return HG_OK;
}

hg_status hg_get_accelerations(hg_system *S, double *a, double *alpha) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_get_accelerations"));
    HG_TRY(download3(S, S->a, a));
    HG_TRY(download3(S, S->al, alpha));
    return HG_OK;
}

hg_status hg_reset_accelerations(hg_system *S) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_HIP(launch_zero_f64(S->a, 4 * S->n, S->stream));
    HG_HIP(launch_zero_f64(S->al, 4 * S->n, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    S->acc_valid = true;
    return HG_OK;
}

hg_status hg_get_time(hg_system *S, double *t, uint64_t *step_index) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (t) *t = S->t;
    if (step_index) *step_index = S->step_index;
    return HG_OK;
}

hg_status hg_get_neighbors(hg_system *S, uint64_t *offsets, uint32_t *idx, size_t cap, size_t *total) {
    if (!S || !offsets) return fail(HG_ERR_INVALID, "system and offsets are required");
    HG_TRY(settle(S, "hg_get_neighbors"));
    if (S->p.pair_mode != HG_PAIRS_NEIGHBOR_LIST) return fail(HG_ERR_INVALID, "system has no neighbour list");
    const size_t n = S->n;
    std::vector<uint32_t> nb(size_t(S->cap_n) * n), cnt(n);
    HG_HIP(hipMemcpyAsync(nb.data(), S->nb, nb.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(cnt.data(), S->ncnt, cnt.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    size_t tot = 0;
    for (size_t i = 0; i < n; i++) {
        offsets[i] = tot;
        for (uint32_t k = 0; k < cnt[i]; k++) {
            if (idx && tot < cap) idx[tot] = nb[size_t(k) * n + i];
            tot++;
        }
    }
    offsets[n] = tot;
    if (total) *total = tot;
    return HG_OK;
}

hg_status hg_get_bonds(hg_system *S, uint32_t *pairs, size_t cap, size_t *nbonds, size_t *nskipped) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    for (size_t b = 0; pairs && b < S->bonds.size() && b < cap; b++) {
        pairs[2 * b] = S->bonds[b].first;
        pairs[2 * b + 1] = S->bonds[b].second;
    }
    if (nbonds) *nbonds = S->bonds.size();
    if (nskipped) *nskipped = S->bonds_skipped;
    return HG_OK;
}

hg_status hg_set_legacy_connectors(hg_system *S, size_t nbonds, const uint32_t *pairs, const double *connectors) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (S->p.binary_model != HG_BINARY_SINTER_LEGACY)
        return fail(HG_ERR_INVALID, "hg_set_legacy_connectors: the system is not HG_BINARY_SINTER_LEGACY");
    if (nbonds != S->bonds.size())
        return fail(HG_ERR_INVALID, "hg_set_legacy_connectors: the system has " + std::to_string(S->bonds.size()) +
                                        " bonds, " + std::to_string(nbonds) + " given");
    if (nbonds && (!pairs || !connectors)) return fail(HG_ERR_INVALID, "hg_set_legacy_connectors: pairs and connectors are required");
    for (size_t b = 0; b < nbonds; b++)
        if (pairs[2 * b] != S->bonds[b].first || pairs[2 * b + 1] != S->bonds[b].second)
            return fail(HG_ERR_INVALID, "hg_set_legacy_connectors: pair " + std::to_string(b) +
                                            " is not the system's bond (hg_get_bonds order)");
    const size_t n = S->n;
    std::vector<uint32_t> bj(size_t(S->bcap) * n), cnt(n);
    HG_HIP(hipMemcpyAsync(bj.data(), S->bj, bj.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(cnt.data(), S->bcnt, n * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    auto slot_of = [&](uint32_t i, uint32_t j) {
        uint32_t k = 0;
        while (k < cnt[i] && bj[size_t(k) * n + i] != j) k++;
        return k;
    };
    std::vector<double> rec(size_t(S->bcap) * kConnRec * n, 0.0);
    for (size_t b = 0; b < nbonds; b++) {
        const uint32_t i = pairs[2 * b], j = pairs[2 * b + 1];
        double *ri = rec.data() + (size_t(slot_of(i, j)) * n + i) * kConnRec;
        double *rj = rec.data() + (size_t(slot_of(j, i)) * n + j) * kConnRec;
        for (int c = 0; c < 15; c++) {
            ri[c] = connectors[30 * b + c];
            rj[c] = connectors[30 * b + 15 + c];
        }
    }
    HG_HIP(hipMemcpyAsync(S->lconn, rec.data(), rec.size() * 8, hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    S->legacy_ready = true;
    return HG_OK;
}

hg_status hg_get_legacy_connectors(hg_system *S, double *connectors, size_t cap, size_t *nbonds) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_get_legacy_connectors"));
    if (S->p.binary_model != HG_BINARY_SINTER_LEGACY)
        return fail(HG_ERR_INVALID, "hg_get_legacy_connectors: the system is not HG_BINARY_SINTER_LEGACY");
    if (nbonds) *nbonds = S->bonds.size();
    if (!connectors || cap < S->bonds.size()) return HG_OK;
    const size_t n = S->n;
    std::vector<uint32_t> bj(size_t(S->bcap) * n), cnt(n);
    std::vector<double> rec(size_t(S->bcap) * kConnRec * n);
    HG_HIP(hipMemcpyAsync(bj.data(), S->bj, bj.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(cnt.data(), S->bcnt, n * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(rec.data(), S->lconn, rec.size() * 8, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    auto slot_of = [&](uint32_t i, uint32_t j) {
        uint32_t k = 0;
        while (k < cnt[i] && bj[size_t(k) * n + i] != j) k++;
        return k;
    };
    for (size_t b = 0; b < S->bonds.size(); b++) {
        const uint32_t i = S->bonds[b].first, j = S->bonds[b].second;
        const double *ri = rec.data() + (size_t(slot_of(i, j)) * n + i) * kConnRec;
        const double *rj = rec.data() + (size_t(slot_of(j, i)) * n + j) * kConnRec;
        for (int c = 0; c < 15; c++) {
            connectors[30 * b + c] = ri[c];
            connectors[30 * b + 15 + c] = rj[c];
        }
    }
    return HG_OK;
}

hg_status hg_get_springs(hg_system *S, int which, uint32_t *ij, double *springs, size_t cap, size_t *total) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_TRY(settle(S, "hg_get_springs"));
    if (total) *total = 0;
    if (S->p.binary_model == HG_BINARY_NONE) return HG_OK;
    const size_t n = S->n, cs = S->cap_s;
    std::vector<uint32_t> pj(cs * n), cnt(n);
    std::vector<double> ps(cs * kSpringRec * n);
    HG_HIP(hipMemcpyAsync(pj.data(), S->pj[S->cur], pj.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(ps.data(), S->ps[S->cur], ps.size() * 8, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipMemcpyAsync(cnt.data(), S->pcnt[S->cur], cnt.size() * 4, hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    size_t tot = 0;
    for (size_t i = 0; i < n; i++)
        for (uint32_t s = 0; s < cnt[i]; s++) {
            const uint32_t tag = pj[size_t(s) * n + i];
            const bool bonded = (tag & kBondBit) != 0;
            if (bonded != (which == 1)) continue;
            if (tot < cap) {
                if (ij) {
                    ij[2 * tot] = uint32_t(i);
                    ij[2 * tot + 1] = tag & kIdxMask;
                }
                if (springs)
                    for (int c = 0; c < 9; c++) springs[9 * tot + c] = ps[(size_t(s) * n + i) * kSpringRec + c];
            }
            tot++;
        }
    if (total) *total = tot;
    return HG_OK;
}

hg_status hg_profile_enable(hg_system *S, int enable) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_HIP(hipStreamSynchronize(S->stream));
    HG_TRY(collect_profile(S));
    for (int k = 0; k < 4; k++) {
        S->prof_launches[k] = 0;
        S->prof_ms[k] = 0.0;
    }
    S->profiling = enable != 0;
    return HG_OK;
}

hg_status hg_set_dissipation_tracking(hg_system *S, int enable) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    if (enable) {
        if (S->p.precision != HG_PRECISION_FP64 ||
            (S->p.binary_model != HG_BINARY_CONTACT && S->p.binary_model != HG_BINARY_SINTER) ||
            S->p.sinter_breaking)
            return fail(HG_ERR_INVALID, "hg_set_dissipation_tracking: FP64 systems with the contact or alt_sinter "
                                        "functor (bonds that do not break) only");
        if (!S->diss) HG_TRY(dalloc(S, &S->diss, 4 * S->n));
        HG_HIP(launch_zero_f64(S->diss, 4 * S->n, S->stream));
    }
    S->track_diss = enable != 0;
    return HG_OK;
}

hg_status hg_get_dissipated_work(hg_system *S, double *out) {
    if (!S || !out) return fail(HG_ERR_INVALID, "system and out are required");
    if (!S->diss) return fail(HG_ERR_INVALID, "hg_get_dissipated_work: tracking was never enabled");
    std::vector<double> h(4 * S->n);
    HG_HIP(hipMemcpyAsync(h.data(), S->diss, h.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    if (S->flags_unchecked) HG_TRY(check_flags(S, "hg_get_dissipated_work"));
    for (int c = 0; c < 4; c++) {
        double sum = 0.0;
        for (size_t i = 0; i < S->n; i++) sum = sum + h[4 * i + c];
        out[c] = sum;
    }
    out[4] = ((out[0] + out[1]) + out[2]) + out[3];
    return HG_OK;
}

hg_status hg_set_per_step_launches(hg_system *S, int enable) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    S->per_step = enable != 0;
    return HG_OK;
}

hg_status hg_profile_read(hg_system *S, uint64_t *launches, double *device_ms) {
    if (!S) return fail(HG_ERR_INVALID, "system is null");
    HG_HIP(hipStreamSynchronize(S->stream));
    HG_TRY(collect_profile(S));
    for (int k = 0; k < 4; k++) {
        if (launches) launches[k] = S->prof_launches[k];
        if (device_ms) device_ms[k] = S->prof_ms[k];
    }
    return HG_OK;
}

hg_status hg_sinter_topology(size_t n, const double *x0, double r_part, double d_crit, uint32_t *pairs, size_t cap,
                             size_t *nbonds, size_t *nskipped) {
    if (n > 0 && !x0) return fail(HG_ERR_INVALID, "x0 is null");
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped = 0;
    sinter_topology_host(n, x0, r_part, d_crit, bonds, skipped);
    for (size_t b = 0; pairs && b < bonds.size() && b < cap; b++) {
        pairs[2 * b] = bonds[b].first;
        pairs[2 * b + 1] = bonds[b].second;
    }
    if (nbonds) *nbonds = bonds.size();
    if (nskipped) *nskipped = skipped;
    return HG_OK;
}

hg_status hg_breaking_sinter_topology(size_t n, const double *x0, double r_part, double d_crit, uint32_t *pairs,
                                      size_t cap, size_t *nbonds) {
    if (n > 0 && !x0) return fail(HG_ERR_INVALID, "x0 is null");
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped = 0;
    sinter_topology_host(n, x0, r_part, d_crit, bonds, skipped, false);
    for (size_t b = 0; pairs && b < bonds.size() && b < cap; b++) {
        pairs[2 * b] = bonds[b].first;
        pairs[2 * b + 1] = bonds[b].second;
    }
    if (nbonds) *nbonds = bonds.size();
    return HG_OK;
}

hg_status hg_set_neck_strengths(hg_system *S, const double *strengths) {
    if (!S || !strengths) return fail(HG_ERR_INVALID, "system and strengths are required");
    if (!S->bcnt) return fail(HG_ERR_INVALID, "hg_set_neck_strengths: the system's bonds do not break "
                                              "(params.sinter_breaking)");
    const size_t n = S->n;
    std::vector<uint32_t> bj(size_t(S->bcap) * n, kNoEntry), cnt(n, 0);
    std::vector<double> bs(size_t(S->bcap) * n, 0.0);
    for (size_t b = 0; b < S->bonds.size(); b++) {
        for (int end = 0; end < 2; end++) {
            const uint32_t i = end ? S->bonds[b].second : S->bonds[b].first;
            const uint32_t j = end ? S->bonds[b].first : S->bonds[b].second;
            bj[size_t(cnt[i]) * n + i] = j;
            bs[size_t(cnt[i]) * n + i] = strengths[b];
            cnt[i]++;
        }
    }
    HG_HIP(hipMemcpyAsync(S->bj, bj.data(), bj.size() * 4, hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipMemcpyAsync(S->bstr, bs.data(), bs.size() * 8, hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipMemcpyAsync(S->bcnt, cnt.data(), n * 4, hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    return HG_OK;
}

hg_status hg_diagnostics(hg_system *S, double mass, double inertia, double *out) {
    if (!S || !out) return fail(HG_ERR_INVALID, "system and out are required");
    HG_TRY(settle(S, "hg_diagnostics"));
    HG_HIP(launch_diagnostics(S->x, S->v, S->w, uint32_t(S->n), mass, inertia, S->red, S->stream));
    HG_HIP(hipMemcpyAsync(out, S->red, 9 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    return HG_OK;
}

hg_status hg_energy(hg_system *S, double mass, double inertia, double *out) {
    if (!S || !out) return fail(HG_ERR_INVALID, "system and out are required");
    HG_TRY(settle(S, "hg_energy"));
    if (!S->energy) HG_TRY(dalloc(S, &S->energy, size_t(8) * S->n));
    const bool nl = S->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST, bin = S->p.binary_model != HG_BINARY_NONE;
    energy_view E{S->x,
                  S->v,
                  S->w,
                  0u,
                  uint32_t(S->n),
                  S->n,
                  nl ? S->nb : nullptr,
                  nl ? S->ncnt : nullptr,
                  bin ? S->pj[S->cur] : nullptr,
                  bin ? S->pcnt[S->cur] : nullptr,
                  bin ? S->ps[S->cur] : nullptr,
                  S->lconn ? S->bj : nullptr,
                  S->lconn ? S->bcnt : nullptr,
                  S->lslot,
                  S->lconn};
    if (S->p.binary_model == HG_BINARY_SINTER_LEGACY && !S->legacy_ready)
        return fail(HG_ERR_INVALID, "hg_energy: the legacy sinter_functor's connectors are not set");
    HG_HIP(launch_energy(S->p, E, mass, inertia, S->energy, S->stream));
    std::vector<double> h(size_t(8) * S->n);
    HG_HIP(hipMemcpyAsync(h.data(), S->energy, h.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    sum_energy(h.data(), S->n, out);
    return HG_OK;
}

}  // extern "C"

namespace {
// Device rows of a system (n particles, stride n) copied to the host.
template <typename T>
hg_status download_all(hg_system *S, const T *dev, size_t count, std::vector<T> &host) {
    host.resize(count);
    if (count == 0) return HG_OK;
    HG_HIP(hipMemcpyAsync(host.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    return HG_OK;
}
template <typename T>
hg_status upload_all(hg_system *S, T *dev, const std::vector<T> &host) {
    if (host.empty()) return HG_OK;
    HG_HIP(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, S->stream));
    HG_HIP(hipStreamSynchronize(S->stream));
    return HG_OK;
}
}  // namespace

extern "C" {

hg_status hg_save(hg_system *S, void *buf, size_t cap, size_t *size) {
    if (!S || !size) return fail(HG_ERR_INVALID, "hg_save: system and size are required");
    HG_TRY(settle(S, "hg_save"));
    const uint32_t n = uint32_t(S->n);
    snapshot z;
    z.p = S->p;
    z.n = n;
    z.t = S->t;
    z.step_index = S->step_index;
    z.acc_valid = S->acc_valid ? 1u : 0u;
    for (auto [dev, host] : {std::pair<const double *, std::vector<double> *>{S->x, &z.x}, {S->v, &z.v},
                             {S->th, &z.th}, {S->w, &z.w}, {S->a, &z.a}, {S->al, &z.al}}) {
        host->resize(3 * size_t(n));
        HG_TRY(download3(S, dev, host->data()));
    }
    if (S->nb) {
        std::vector<uint32_t> nb;
        z.has_nl = true;
        HG_TRY(download_all(S, S->nb, size_t(S->cap_n) * n, nb));
        HG_TRY(download_all(S, S->ncnt, n, z.nl_count));
        rows_to_csr(nb.data(), z.nl_count.data(), n, 1, 1, z.nl_ids);
    }
    if (S->cap_s) {
        // the current half of the pair-state table (with the legacy functor:
        // the contact springs of its unbonded pairs)
        std::vector<uint32_t> pj;
        std::vector<double> ps;
        z.has_table = true;
        HG_TRY(download_all(S, S->pj[S->cur], size_t(S->cap_s) * n, pj));
        HG_TRY(download_all(S, S->ps[S->cur], size_t(S->cap_s) * kSpringRec * n, ps));
        HG_TRY(download_all(S, S->pcnt[S->cur], n, z.pt_count));
        rows_to_csr(pj.data(), z.pt_count.data(), n, 1, 1, z.pt_tags);
        rows_to_csr(ps.data(), z.pt_count.data(), n, kSpringRec, 9, z.pt_springs);
    }
    for (auto &b : S->bonds) {
        z.bonds.push_back(b.first);
        z.bonds.push_back(b.second);
    }
    z.skipped = S->bonds_skipped;
    z.nf = uint32_t(S->nf);
    if (S->nf > 0) {
        HG_TRY(download_all(S, S->facets, size_t(S->nf), z.facets));
        HG_TRY(download_all(S, S->ss, size_t(S->nf) * 9 * n, z.ss));
    }
    if (S->bstr) {
        std::vector<uint32_t> bj;
        std::vector<double> bs;
        z.has_neck = true;
        HG_TRY(download_all(S, S->bj, size_t(S->bcap) * n, bj));
        HG_TRY(download_all(S, S->bstr, size_t(S->bcap) * n, bs));
        HG_TRY(download_all(S, S->bcnt, n, z.neck_count));
        rows_to_csr(bj.data(), z.neck_count.data(), n, 1, 1, z.neck_j);
        rows_to_csr(bs.data(), z.neck_count.data(), n, 1, 1, z.neck_str);
    }
    if (S->lconn) {
        std::vector<uint32_t> bj, sl;
        std::vector<double> lc;
        if (!S->legacy_ready) return fail(HG_ERR_INVALID, "hg_save: the legacy connectors are not set");
        z.has_legacy = true;
        HG_TRY(download_all(S, S->bj, size_t(S->bcap) * n, bj));
        HG_TRY(download_all(S, S->lslot, size_t(S->bcap) * n, sl));
        HG_TRY(download_all(S, S->lconn, size_t(S->bcap) * kConnRec * n, lc));
        HG_TRY(download_all(S, S->bcnt, n, z.leg_count));
        rows_to_csr(bj.data(), z.leg_count.data(), n, 1, 1, z.leg_j);
        rows_to_csr(sl.data(), z.leg_count.data(), n, 1, 1, z.leg_slot);
        rows_to_csr(lc.data(), z.leg_count.data(), n, kConnRec, 15, z.leg_conn);
    }
    const std::vector<uint8_t> bytes = encode_snapshot(z);
    *size = bytes.size();
    if (!buf) return HG_OK;
    if (cap < bytes.size())
        return fail(HG_ERR_CAPACITY, "hg_save: the snapshot needs " + std::to_string(bytes.size()) + " bytes");
    std::memcpy(buf, bytes.data(), bytes.size());
    return HG_OK;
}

hg_status hg_load(const void *buf, size_t size, hg_system **out) {
    if (!out) return fail(HG_ERR_INVALID, "hg_load: out is required");
    *out = nullptr;
    snapshot z;
    HG_TRY(decode_snapshot(buf, size, z));
    const uint32_t n = uint32_t(z.n);
    std::vector<std::pair<uint32_t, uint32_t>> bonds(z.bonds.size() / 2);
    for (size_t b = 0; b < bonds.size(); b++) bonds[b] = {z.bonds[2 * b], z.bonds[2 * b + 1]};
    hg_system *S = nullptr;
    HG_TRY(create_impl(&z.p, n, z.x.data(), z.v.data(), z.th.data(), z.w.data(), &bonds, z.skipped, &S));
    auto fill = [&]() -> hg_status {
        HG_TRY(upload3(S, S->a, z.a.data()));
        HG_TRY(upload3(S, S->al, z.al.data()));
        if (z.has_nl != (S->nb != nullptr) || (z.has_table != (S->cap_s != 0)) || z.has_neck != (S->bstr != nullptr) ||
            z.has_legacy != (S->lconn != nullptr))
            return fail(HG_ERR_INVALID, "hg_load: the snapshot's tables do not match its parameters");
        if (z.has_nl) {
            if (max_count(z.nl_count) > S->cap_n)
                return fail(HG_ERR_CAPACITY, "hg_load: a neighbour row exceeds the list capacity");
            std::vector<uint32_t> nb;
            csr_to_rows(z.nl_count, z.nl_ids, n, S->cap_n, 1, 1, 0u, nb);
            HG_TRY(upload_all(S, S->nb, nb));
            HG_TRY(upload_all(S, S->ncnt, z.nl_count));
        }
        if (z.has_table) {
            if (max_count(z.pt_count) > S->cap_s)
                return fail(HG_ERR_CAPACITY, "hg_load: a pair-state row exceeds the table capacity");
            std::vector<uint32_t> pj;
            std::vector<double> ps;
            csr_to_rows(z.pt_count, z.pt_tags, n, S->cap_s, 1, 1, kNoEntry, pj);
            csr_to_rows(z.pt_count, z.pt_springs, n, S->cap_s, kSpringRec, 9, 0.0, ps);
            S->cur = 0;
            HG_TRY(upload_all(S, S->pj[0], pj));
            HG_TRY(upload_all(S, S->ps[0], ps));
            HG_TRY(upload_all(S, S->pcnt[0], z.pt_count));
        }
        if (z.nf > 0) {
            std::vector<double> verts(9 * size_t(z.nf)), vel(3 * size_t(z.nf));
            HG_TRY(hg_set_facets(S, z.nf, verts.data(), vel.data()));  // allocates; the records follow
            HG_TRY(upload_all(S, S->facets, z.facets));
            HG_TRY(upload_all(S, S->ss, z.ss));
        }
        if (z.has_neck) {
            std::vector<uint32_t> bj;
            std::vector<double> bs;
            if (max_count(z.neck_count) > S->bcap) return fail(HG_ERR_INVALID, "hg_load: bad neck rows");
            csr_to_rows(z.neck_count, z.neck_j, n, S->bcap, 1, 1, kNoEntry, bj);
            csr_to_rows(z.neck_count, z.neck_str, n, S->bcap, 1, 1, 0.0, bs);
            HG_TRY(upload_all(S, S->bj, bj));
            HG_TRY(upload_all(S, S->bstr, bs));
            HG_TRY(upload_all(S, S->bcnt, z.neck_count));
        }
        if (z.has_legacy) {
            std::vector<uint32_t> bj, sl;
            std::vector<double> lc;
            if (max_count(z.leg_count) > S->bcap) return fail(HG_ERR_INVALID, "hg_load: bad legacy rows");
            csr_to_rows(z.leg_count, z.leg_j, n, S->bcap, 1, 1, kNoEntry, bj);
            csr_to_rows(z.leg_count, z.leg_slot, n, S->bcap, 1, 1, 0u, sl);
            csr_to_rows(z.leg_count, z.leg_conn, n, S->bcap, kConnRec, 15, 0.0, lc);
            HG_TRY(upload_all(S, S->bj, bj));
            HG_TRY(upload_all(S, S->lslot, sl));
            HG_TRY(upload_all(S, S->lconn, lc));
            HG_TRY(upload_all(S, S->bcnt, z.leg_count));
            S->legacy_ready = true;
        }
        S->t = z.t;
        S->step_index = z.step_index;
        S->acc_valid = z.acc_valid != 0;
        return HG_OK;
    };
    const hg_status st = fill();
    if (st != HG_OK) {
        const std::string msg = hg_last_error();
        hg_destroy(S);
        return fail(st, msg);
    }
    *out = S;
    return HG_OK;
}

}  // extern "C"
