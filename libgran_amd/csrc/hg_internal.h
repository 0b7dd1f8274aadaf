// Internal declarations of the hipgran library: the device-side system
// object, kernel launchers and error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/hipgran.h"
#include "dem_math.h"

namespace hg {

// Pair-state table entry tag: bit 31 marks a bonded (alt_sinter) pair; the low
// 31 bits are the partner index j.
constexpr uint32_t kBondBit = 0x80000000u;
constexpr uint32_t kIdxMask = 0x7fffffffu;
constexpr uint32_t kNoEntry = 0xffffffffu;  // (kNoEntry & kIdxMask) exceeds every valid j

// Device flag words (int32 each) written by kernels, read back by the host.
// The MAXCOUNT words are raised only by lists that overflow, so they hold the
// count the overflow needed. FLAG_NUMERIC: a velocity or angular velocity
// left a closing kick non-finite.
enum : int {
    FLAG_NL_OVERFLOW = 0,
    FLAG_PAIR_OVERFLOW = 1,
    FLAG_NL_MAXCOUNT = 2,
    FLAG_PAIR_MAXCOUNT = 3,
    FLAG_NUMERIC = 4,
    FLAG_COUNT = 5
};

constexpr int kMaxFacets = 16;
// Bonded partners per particle in the neck-strength lists (touching equal
// spheres: at most 12).
constexpr uint32_t kMaxBondsPerParticle = 16;
// Pair-state tags loaded up front per particle (the table capacity is never
// below this) and neighbour ids loaded up front per particle. Deeper
// prefetch arrays are indexed at run time and land in scratch memory, which
// costs more than the dependent loads they save.
constexpr uint32_t kTagPrefetch = 1;
constexpr uint32_t kIdPrefetch = 1;
// Doubles per pair-state record: the 9 spring components and one pad word,
// so that records stay 16-byte aligned.
constexpr uint32_t kSpringRec = 10;
// Doubles per legacy connector record: five points and one pad word.
constexpr uint32_t kConnRec = 16;
// New-table tags a particle keeps in registers while it walks its partners
// and writes at the end of the walk, all lanes together.
constexpr uint32_t kNewTagBuf = 2;

// Pair-state table, old (read) and new (written) halves; the tag of slot s of
// particle i at s*stride + i, its springs in the kSpringRec-double record
// starting at (s*stride + i)*kSpringRec. A particle's record is contiguous:
// the walker stores a record when it reaches that pair, at a different
// iteration for each lane, and a record shared with no other lane fills whole
// 32-byte sectors on its own instead of leaving partly written lines in L2.
struct pair_table {
    const uint32_t *pj_old;
    const double *ps_old;
    const uint32_t *pcnt_old;
    uint32_t *pj_new;
    double *ps_new;
    uint32_t *pcnt_new;
    uint32_t cap;
    size_t stride;
    // alt_breaking_sinter neck strengths: particle i's bonded partners at
    // bj[s*stride + i], their strengths at bstr[s*stride + i], s < bcnt[i]
    // (null when bonds do not break)
    const uint32_t *bj;
    const double *bstr;
    const uint32_t *bcnt;
    // legacy sinter_functor (HG_BINARY_SINTER_LEGACY): bonded partners at
    // bj/bcnt as above; the connectors particle i holds for its slot-s bond
    // at lconn[(s*stride + i)*kConnRec + 3q + c] (five points q), and the
    // slot of i in the partner's list at lslot[s*stride + i] (null otherwise)
    const double *lconn;
    const uint32_t *lslot;
};

// Everything a force/integrate kernel needs, passed by value.
struct sys_view {
    uint32_t n;
    const double *x, *v, *w;  // AoS-by-field: 4 doubles per particle (w unused lane)
    double *a, *al;
    // neighbour list (transposed: slot k of particle i at k*n + i)
    const uint32_t *nb;
    const uint32_t *ncnt;
    uint32_t cap_n;
    pair_table pt;
    // facets
    const facet_dev *facets;
    double *ss;  // surface springs: (f*9 + c)*n + i
    int nf;
    int *flags;
    double *diss = nullptr;  // dissipated work, 4 per particle (when tracked)
};

struct hg_system_impl {
    hg_params p{};
    size_t n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // particle state, 4 doubles per particle per field
    double *x = nullptr, *v = nullptr, *th = nullptr, *w = nullptr, *a = nullptr, *al = nullptr;
    bool acc_valid = false;
    bool flags_unchecked = false;  // hg_step_async queued steps whose flags nobody has read yet
    double t = 0.0;
    uint64_t step_index = 0;
    // pair-state table (double buffered)
    uint32_t cap_s = 0;
    uint32_t *pj[2] = {nullptr, nullptr};
    double *ps[2] = {nullptr, nullptr};
    uint32_t *pcnt[2] = {nullptr, nullptr};
    int cur = 0;
    // neck strengths (sinter_breaking): bonded partners and strengths per particle
    uint32_t bcap = 0;
    uint32_t *bj = nullptr, *bcnt = nullptr;
    double *bstr = nullptr;
    // legacy sinter_functor connectors (pair_table::lconn / lslot)
    double *lconn = nullptr;
    uint32_t *lslot = nullptr;
    bool legacy_ready = false;
    // neighbour list
    uint32_t cap_n = 0;
    uint32_t *nb = nullptr, *ncnt = nullptr;
    // cell-list scratch
    uint32_t hash_size = 0;
    uint32_t *cell_of = nullptr, *bucket_count = nullptr, *bucket_start = nullptr, *bucket_fill = nullptr,
             *cell_items = nullptr, *scan_tmp = nullptr;
    double *bbox = nullptr;  // lo[3], hi[3]
    // facets
    int nf = 0;
    facet_dev *facets = nullptr;
    double *ss = nullptr;
    // flags
    int *flags = nullptr;
    int *flags_host = nullptr;  // pinned
    // reductions
    double *red = nullptr;
    double *energy = nullptr;  // per-particle ledger terms, 8 per particle (allocated on first use)
    double *tiny = nullptr;    // per-pair springs of the pair-parallel pass (n*n*kSpringRec, on first use)
    // host copies of the topology
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t bonds_skipped = 0;
    std::vector<void *> allocations;
    // launch profiling (hg_profile_enable)
    bool profiling = false;
    bool per_step = false;  // hg_set_per_step_launches: no one-launch small path
    // dissipated work (hg_set_dissipation_tracking): 4 doubles per particle
    double *diss = nullptr;
    bool track_diss = false;
    struct timed_range {
        int phase;
        hipEvent_t begin, end;
    };
    std::vector<hipEvent_t> event_pool;
    std::vector<timed_range> pending;
    uint64_t prof_launches[4] = {0, 0, 0, 0};
    double prof_ms[4] = {0, 0, 0, 0};
};

enum : int { PHASE_FORCE = 0, PHASE_KICK = 1, PHASE_REBUILD = 2 };

// error helpers (thread-local message)
void set_error(const std::string &msg);
hg_status fail(hg_status st, const std::string &msg);

#define HG_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::hg::fail(HG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// kernel launchers (kernels.hip)
hipError_t launch_force(const hg_params &p, const sys_view &sv, hipStream_t s);
hipError_t launch_kick_drift(double *x, double *v, double *th, double *w, const double *a, const double *al,
                             uint32_t n, bool pending, double half_prev, double half, double dt, facet_dev *facets,
                             int nf, double facet_dt, hipStream_t s);
// Closing kick; raises flags[FLAG_NUMERIC] when a result is not finite.
hipError_t launch_final_kick(double *v, double *w, const double *a, const double *al, uint32_t n, double half,
                             facet_dev *facets, int nf, double facet_dt, int *flags, hipStream_t s);
hipError_t launch_neighbor_build(hg_system_impl &S, hipStream_t s);
hipError_t launch_advance_facet(facet_dev *f, const double *v, double dt, hipStream_t s);
// nsteps steps of a system of at most 256 particles (FP64) in one workgroup
// (see k_small_steps); `even`/`odd` are the views whose pair-state tables the
// even/odd steps read. Systems of at most kTinySystemMax particles (all
// pairs, or lists of capacity cap_n with n*cap_n <= 768) evaluate their pairs
// in parallel (k_tiny_steps) when `scratch` holds n*max(n, cap_n)*kSpringRec
// doubles.
constexpr uint32_t kSmallSystemMax = 256, kTinySystemMax = 27;
hipError_t launch_small_steps(const hg_params &p, const sys_view &even, const sys_view &odd, double *x, double *v,
                              double *th, double *w, facet_dev *facets, uint32_t nsteps, bool pending_first,
                              bool final_kick, double half, double dt, int nf_advance, double *lconn,
                              const uint32_t *bcnt, double *scratch, hipStream_t s);
// sinter_step_handler's connector moves after a drift (x and v already
// advanced): each of particle i's connectors follows dx = v_i dt, then turns
// by dtheta = w_i dt about the new x_i.
hipError_t launch_legacy_move(const double *x, const double *v, const double *w, uint32_t n, double dt,
                              double *lconn, const uint32_t *bcnt, hipStream_t s);
// Exclusive scan of count words; tmp holds scan_tmp_words(count) words.
hipError_t launch_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t count, uint32_t *tmp, hipStream_t s);
uint32_t scan_tmp_words(uint32_t count);
hipError_t launch_zero_u32(uint32_t *p, size_t count, hipStream_t s);
hipError_t launch_zero_f64(double *p, size_t count, hipStream_t s);
hipError_t launch_diagnostics(const double *x, const double *v, const double *w, uint32_t n, double mass,
                              double inertia, double *out, hipStream_t s);

// Inputs of the energy ledger for particles [offset, offset + n) of an arena
// whose per-particle arrays have row stride `stride` (a single system:
// offset 0, stride n). nb/ncnt are null in all-pairs mode; pj/ps/pcnt are the
// current pair-state table (null without a binary functor).
struct energy_view {
    const double *x, *v, *w;
    uint32_t offset, n;
    size_t stride;
    const uint32_t *nb, *ncnt;
    const uint32_t *pj, *pcnt;
    const double *ps;
    // legacy sinter_functor: bonded partners, reverse slots and connectors
    // (pair_table::bj/bcnt/lslot/lconn layout; null otherwise)
    const uint32_t *bj, *bcnt, *lslot;
    const double *lconn;
};
// Eight terms per particle into out[8*k + c] (k = local index): kinetic
// translational and rotational; the particle's half of its pairs' normal
// elastic energy; its half of the tangential, rolling and torsion spring
// energy; its half of the pairs' Hamaker potential; 0.
hipError_t launch_energy(const hg_params &p, const energy_view &E, double mass, double inertia, double *out,
                         hipStream_t s);
// Sums the per-particle ledger terms in particle order: out[0..6] the seven
// components, out[7] their total.
inline void sum_energy(const double *terms, size_t n, double *out) {
    for (int c = 0; c < 8; c++) out[c] = 0.0;
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 7; c++) out[c] += terms[8 * i + c];
    for (int c = 0; c < 7; c++) out[7] += out[c];
}

// host-side set-up (topology.cpp)
// prevent_cycles: alt_sinter's spanning forest; false: alt_breaking_sinter's
// every-touching-pair rule.
// inclusive: gaps of magnitude up to and including d_crit bond (the legacy
// sinter_functor's rule, sinter_bridge.h:155-157) instead of strictly below.
void sinter_topology_host(size_t n, const double *x0, double r_part, double d_crit,
                          std::vector<std::pair<uint32_t, uint32_t>> &bonds, size_t &skipped,
                          bool prevent_cycles = true, bool inclusive = false);

}  // namespace hg

struct hg_system : hg::hg_system_impl {};
