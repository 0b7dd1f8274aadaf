// Pools of independent DEM systems (include/hipgran.h, hg_pool_*): the
// batched and serving path.
//
// All systems of a pool live in one device arena of P particles (P = the
// pool's particle capacity); every per-particle array uses the same transposed
// layout as the single-system path with stride P, so a system is a contiguous
// index range [offset, offset + n) and a workgroup that owns it touches
// consecutive words.
//
// hg_pool_step launches one persistent kernel per rebuild interval. Each
// workgroup repeatedly takes the next system from a largest-first work queue
// and runs all steps of the interval for it: kick/drift of its own particles,
// workgroup barrier, acceleration pass (the same pair fold, walker and facet
// fold as the single-system kernels), barrier. A system's arithmetic does not
// depend on the other systems in the pool, so results are identical alone and
// in any mix. Between intervals a batched hashed cell list rebuilds the
// neighbour lists of the systems that are due, keyed by (system, cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "force_device.h"
#include "hg_internal.h"
#include "pool_internal.h"
#include "snapshot.h"

// Device scratch of hg_pool_energy_noalloc, the entry point that allocates
// nothing: the per-particle ledger terms of one system (8 doubles per
// particle) followed by their 8 sums, for systems of up to
// HG_NOALLOC_MAX_PARTICLES particles (8 MiB). Every GPU that runs the library
// holds its own copy until the process ends. Host code reaches it only through
// the address hipGetSymbolAddress returns for a device (noalloc_scratch),
// never through this name.
constexpr size_t kNoAllocParticles = HG_NOALLOC_MAX_PARTICLES;
static __device__ __align__(256) double scratch[8 * kNoAllocParticles + 8];

namespace hg {

// slot: the system's own index, naming its fault word in pool_view::sflags.
struct pool_meta {
    uint32_t offset, n, cur, acc_valid, nf, rebuild, slot, pad1;
};

struct pool_view {
    double *x, *v, *th, *w, *a, *al;
    const uint32_t *nb;
    const uint32_t *ncnt;
    uint32_t *pj0, *pj1, *pcnt0, *pcnt1;
    double *ps0, *ps1;
    uint32_t cap_s, cap_n;
    const uint32_t *bj;  // neck-strength lists (see pair_table), null unless bonds break
    const double *bstr;
    const uint32_t *bcnt;
    facet_dev *facets;
    uint32_t max_facets;
    double *ss;
    const pool_meta *meta;
    size_t stride;
    int *flags;
    int *sflags;  // per-system fault words (HG_FAULT_* bits), indexed by slot
};

namespace {

// Where a running system's particle state lives during a chunk (chosen per
// system from its size and the workgroup's LDS; local index g - offset):
//   STATE_FULL   all of x, v, omega, theta, a, alpha in LDS (18 doubles per
//                particle), read from global memory at the start of the
//                chunk and written back once at its end;
//   STATE_MIRROR x, v, omega copied into LDS (9 doubles per particle) and
//                refreshed by each owner lane in the kick/drift phase, so
//                partner gathers read LDS; the state itself stays global;
//   STATE_POS    x alone copied into LDS (3 doubles per particle) and
//                refreshed likewise: every partner needs its position, only
//                the bonded and overlapping ones (a few of a 3.2r list) its
//                velocities, which then come from global memory;
//   STATE_GLOBAL everything in global memory.
enum { STATE_GLOBAL = 0, STATE_MIRROR = 1, STATE_FULL = 2, STATE_POS = 3 };
#ifndef HG_CHUNK_EAGER
#define HG_CHUNK_EAGER 1
#endif
#ifndef HG_CHUNK_PREFETCH
#define HG_CHUNK_PREFETCH 1
#endif
#ifndef HG_CHUNK_MIN_BLOCKS
#define HG_CHUNK_MIN_BLOCKS 1
#endif
constexpr bool kChunkEager = HG_CHUNK_EAGER, kChunkPrefetch = HG_CHUNK_PREFETCH;
extern __shared__ double pool_lds[];
constexpr uint32_t kPosStride = 3, kMirrorStride = 9, kFullStride = 18;
template <int MODE>
constexpr uint32_t lds_stride() {
    return MODE == STATE_FULL ? kFullStride : (MODE == STATE_MIRROR ? kMirrorStride : kPosStride);
}

template <int MODE>
__device__ __forceinline__ double *lds_rec(uint32_t offset, uint32_t g) {
    return pool_lds + size_t(g - offset) * lds_stride<MODE>();
}
__device__ __forceinline__ d3 lds3(const double *q) { return d3{q[0], q[1], q[2]}; }
__device__ __forceinline__ void sds3(double *q, d3 a) {
    q[0] = a.x;
    q[1] = a.y;
    q[2] = a.z;
}

template <int MODE>
__device__ __forceinline__ d3 load_x(const pool_view &V, uint32_t offset, uint32_t g) {
    if constexpr (MODE == STATE_GLOBAL) return ld3(V.x, g);
    else return lds3(lds_rec<MODE>(offset, g));
}
template <int MODE>
__device__ __forceinline__ void load_vw(const pool_view &V, uint32_t offset, uint32_t g, d3 &v, d3 &w) {
    if constexpr (MODE == STATE_GLOBAL || MODE == STATE_POS) {
        v = ld3(V.v, g);
        w = ld3(V.w, g);
    } else {
        const double *q = lds_rec<MODE>(offset, g);
        v = lds3(q + 3);
        w = lds3(q + 6);
    }
}

template <int MODE>
__device__ __forceinline__ void load_state(const pool_view &V, uint32_t offset, uint32_t g, d3 &x, d3 &v, d3 &w) {
    if constexpr (MODE == STATE_GLOBAL) {
        x = ld3(V.x, g);
        v = ld3(V.v, g);
        w = ld3(V.w, g);
    } else if constexpr (MODE == STATE_POS) {
        x = lds3(lds_rec<MODE>(offset, g));
        v = ld3(V.v, g);
        w = ld3(V.w, g);
    } else {
        const double *q = lds_rec<MODE>(offset, g);
        x = lds3(q);
        v = lds3(q + 3);
        w = lds3(q + 6);
    }
}

template <int MODE>
__device__ __forceinline__ void store_acc(const pool_view &V, uint32_t offset, uint32_t g, d3 a, d3 al) {
    if constexpr (MODE == STATE_FULL) {
        double *q = lds_rec<MODE>(offset, g);
        sds3(q + 12, a);
        sds3(q + 15, al);
    } else {
        st3(V.a, g, a);
        st3(V.al, g, al);
    }
}

// Fills the LDS records of a system at the start of its chunk.
template <int MODE>
__device__ __forceinline__ void lds_fill(const pool_view &V, const pool_meta &M) {
    for (uint32_t g = M.offset + threadIdx.x; g < M.offset + M.n; g += blockDim.x) {
        double *q = lds_rec<MODE>(M.offset, g);
        sds3(q, ld3(V.x, g));
        if constexpr (MODE == STATE_POS) continue;
        sds3(q + 3, ld3(V.v, g));
        sds3(q + 6, ld3(V.w, g));
        if constexpr (MODE == STATE_FULL) {
            sds3(q + 9, ld3(V.th, g));
            sds3(q + 12, ld3(V.a, g));
            sds3(q + 15, ld3(V.al, g));
        }
    }
}

// LDS copy of a neighbour-list system's partner ids for one rebuild interval:
// kmax slots per particle (kmax = the system's longest list), slot k of local
// particle t at lids[k*n + t] as a halfword offset from the system's first
// particle, then the list counts at lids[kmax*n + t]. A system of at most
// 65,535 particles fits the halfwords.
__host__ __device__ inline size_t lds_ids_bytes(uint32_t n, uint32_t kmax) { return size_t(n) * 2 * (kmax + 1); }

__device__ __forceinline__ void lds_ids_fill(const pool_view &V, const pool_meta &M, uint32_t kmax, uint16_t *lids) {
    for (uint32_t t = threadIdx.x; t < M.n; t += blockDim.x) {
        const uint32_t g = M.offset + t, c = V.ncnt[g];
        lids[kmax * M.n + t] = uint16_t(c);
        for (uint32_t k = 0; k < c; k++) lids[k * M.n + t] = uint16_t(V.nb[size_t(k) * V.stride + g] - M.offset);
    }
}

// Largest value of v over the workgroup (s_red: one word per wavefront).
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t *s_red) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, uint32_t(__shfl_xor(int(v), o)));
    if ((threadIdx.x & 63u) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = 0;
    for (uint32_t w = 0; w < (blockDim.x + 63u) / 64u; w++) r = max(r, s_red[w]);
    __syncthreads();
    return r;
}

// Acceleration pass for particles g = g_first, g_first + g_step, ... < g_end
// of system M (a workgroup strides over the whole system; a wide launch gives
// each thread one particle). With LIDS the partner ids come from the system's
// LDS copy `lids` (lds_ids_fill). after(g, x, v, w, a, alpha) runs once a
// particle's accelerations are stored, with its state and accelerations
// still in registers.
struct no_after {
    __device__ void operator()(uint32_t, d3, d3, d3, d3, d3) const {}
};
// The FP64 pair path's bond constants held in vector registers for a pass
// (as pinned() does for the mixed path): kept as scalars they are spilled and
// re-read lane by lane inside the pair loop.
__device__ __forceinline__ hg_params pinned64(const hg_params &P) {
    hg_params p = P;
    hg_sinter_params &s = p.sinter;
    double *f[] = {&s.k, &s.gamma_n, &s.k_t, &s.gamma_t, &s.k_r, &s.gamma_r,
                   &s.k_o, &s.gamma_o, &s.r_part, &s.mass, &s.inertia, &s.dt};
#pragma unroll
    for (double *q : f) pin_vgpr(*q);
    return p;
}
template <int BIN, int HAM, int PAIRS, int PREC, int LDS, bool LIDS = false, bool FAC = true, class AFTER = no_after>
__device__ __forceinline__ void pool_force_phase(const hg_params &P, const mixed_consts &MC, const pool_view &V,
                                                 const pool_meta &M, uint32_t cur, const facet_dev *F,
                                                 uint32_t g_first, uint32_t g_end, uint32_t g_step,
                                                 const uint16_t *lids = nullptr, uint32_t kmax = 0,
                                                 AFTER after = AFTER{}) {
    pair_table T;
    T.pj_old = cur ? V.pj1 : V.pj0;
    T.ps_old = cur ? V.ps1 : V.ps0;
    T.pcnt_old = cur ? V.pcnt1 : V.pcnt0;
    T.pj_new = cur ? V.pj0 : V.pj1;
    T.ps_new = cur ? V.ps0 : V.ps1;
    T.pcnt_new = cur ? V.pcnt0 : V.pcnt1;
    T.cap = V.cap_s;
    T.stride = V.stride;
    T.bj = V.bj;
    T.bstr = V.bstr;
    T.bcnt = V.bcnt;
    const uint32_t end = M.offset + M.n;
    for (uint32_t g = g_first; g < g_end; g += g_step) {
        d3 xi, vi, wi;
        load_state<LDS>(V, M.offset, g, xi, vi, wi);
        walker_t<PREC == HG_PRECISION_MIXED && kChunkEager> W;
        if constexpr (BIN != HG_BINARY_NONE) W.init(T, g);
        id_list ids;
        if constexpr (PAIRS == HG_PAIRS_NEIGHBOR_LIST && !LIDS) ids.init(V.nb, V.ncnt, V.stride, g, V.cap_n);
        // listed partners of g through fold(j, xj, vw), in list order: vw
        // gives a partner's velocities when its pair needs them, from LDS or
        // from a global gather made a pair ahead with its position
        auto partners = [&](auto &&fold) {
            auto run = [&](const auto &ids_) {
                if constexpr (LDS == STATE_GLOBAL) {
                    nl_partners<kChunkPrefetch>(
                        ids_, [&](uint32_t j) { return pstate{ld3(V.x, j), ld3(V.v, j), ld3(V.w, j)}; },
                        [&](uint32_t j, const pstate &sj) {
                            fold(j, sj.x, [&](d3 &v, d3 &w) { v = sj.v, w = sj.w; });
                        });
                } else {
                    nl_partners<kChunkPrefetch>(
                        ids_, [&](uint32_t j) { return load_x<LDS>(V, M.offset, j); },
                        [&](uint32_t j, d3 xj) {
                            fold(j, xj, [&](d3 &v, d3 &w) { load_vw<LDS>(V, M.offset, j, v, w); });
                        });
                }
            };
            if constexpr (LIDS) {
                const uint32_t t = g - M.offset;
                run(lds_id_list{lids, M.n, t, M.offset, lids[kmax * M.n + t]});
            } else {
                run(ids);
            }
        };
        auto vw_of = [&](uint32_t j) { return [&V, &M, j](d3 &v, d3 &w) { load_vw<LDS>(V, M.offset, j, v, w); }; };
        if constexpr (PREC == HG_PRECISION_FP64) {
            acc6 acc{zero3(), zero3()};
            const hg_params pp = pinned64(P);
            if constexpr (PAIRS == HG_PAIRS_NEIGHBOR_LIST) {
                partners([&](uint32_t j, d3 xj, auto &&vw) { pair_fold_vw<BIN, HAM>(pp, j, xi, vi, wi, xj, vw, W, acc); });
            } else {
                for (uint32_t j = M.offset; j < end; j++) {
                    if (j == g) continue;
                    pair_fold_vw<BIN, HAM>(P, j, xi, vi, wi, load_x<LDS>(V, M.offset, j), vw_of(j), W, acc);
                }
            }
            close_table<BIN>(W, T.pcnt_new, g, V.flags, V.sflags + M.slot);
            if (FAC && M.nf > 0) acc = acc + facet_fold(P, F, int(M.nf), V.ss, V.stride, g, xi, vi, wi);
            store_acc<LDS>(V, M.offset, g, acc.lin, acc.ang);
            after(g, xi, vi, wi, acc.lin, acc.ang);
        } else {
            accf acc{zerof(), zerof()};
            const mixed_consts m = pinned(MC);
            if constexpr (PAIRS == HG_PAIRS_NEIGHBOR_LIST) {
                partners([&](uint32_t j, d3 xj, auto &&vw) { pair_fold_mx<BIN, HAM>(m, j, xi, vi, wi, xj, vw, W, acc); });
            } else {
                for (uint32_t j = M.offset; j < end; j++) {
                    if (j == g) continue;
                    pair_fold_mx<BIN, HAM>(MC, j, xi, vi, wi, load_x<LDS>(V, M.offset, j), vw_of(j), W, acc);
                }
            }
            close_table<BIN>(W, T.pcnt_new, g, V.flags, V.sflags + M.slot);
            if (FAC && M.nf > 0) acc = acc + facet_fold_mx(P, MC, F, int(M.nf), V.ss, V.stride, g, xi, vi, wi);
            const d3 a64 = tod(acc.lin), al64 = tod(acc.ang);
            store_acc<LDS>(V, M.offset, g, a64, al64);
            after(g, xi, vi, wi, a64, al64);
        }
    }
}

// Kick (completing the previous step when `pending`), kick, drift; the same
// operations in the same order as k_kick_drift.
template <int LDS>
__device__ __forceinline__ void pool_kick_drift(const pool_view &V, const pool_meta &M, bool pending, double half,
                                                double dt, uint32_t g_first, uint32_t g_end, uint32_t g_step) {
    for (uint32_t g = g_first; g < g_end; g += g_step) {
        double *q = lds_rec<LDS>(M.offset, g);
        d3 vi, wi, ai, ali, x0, th0;
        if constexpr (LDS == STATE_FULL) {
            vi = lds3(q + 3); wi = lds3(q + 6); ai = lds3(q + 12); ali = lds3(q + 15); x0 = lds3(q); th0 = lds3(q + 9);
        } else {
            vi = ld3(V.v, g); wi = ld3(V.w, g); ai = ld3(V.a, g); ali = ld3(V.al, g); x0 = ld3(V.x, g);
            th0 = ld3(V.th, g);
        }
        if (pending) {
            vi = vi + ai * half;
            wi = wi + ali * half;
        }
        vi = vi + ai * half;
        wi = wi + ali * half;
        const d3 xi = x0 + vi * dt;
        const d3 thi = th0 + wi * dt;
        if constexpr (LDS != STATE_FULL) {
            st3(V.x, g, xi);
            st3(V.th, g, thi);
            st3(V.v, g, vi);
            st3(V.w, g, wi);
        }
        if constexpr (LDS != STATE_GLOBAL) sds3(q, xi);
        if constexpr (LDS == STATE_MIRROR || LDS == STATE_FULL) {
            sds3(q + 3, vi);
            sds3(q + 6, wi);
        }
        if constexpr (LDS == STATE_FULL) sds3(q + 9, thi);
    }
}

// Final kick; a system kept wholly in LDS is written back here. A non-finite
// velocity or angular velocity faults the system (HG_FAULT_NUMERIC).
template <int LDS>
__device__ __forceinline__ void pool_final_kick(const pool_view &V, const pool_meta &M, double half,
                                                uint32_t g_first, uint32_t g_end, uint32_t g_step) {
    bool bad = false;
    for (uint32_t g = g_first; g < g_end; g += g_step) {
        d3 vi, wi;
        if constexpr (LDS == STATE_FULL) {
            const double *q = lds_rec<LDS>(M.offset, g);
            const d3 a = lds3(q + 12), al = lds3(q + 15);
            vi = lds3(q + 3) + a * half;
            wi = lds3(q + 6) + al * half;
            st3(V.x, g, lds3(q));
            st3(V.th, g, lds3(q + 9));
            st3(V.a, g, a);
            st3(V.al, g, al);
        } else {
            vi = ld3(V.v, g) + ld3(V.a, g) * half;
            wi = ld3(V.w, g) + ld3(V.al, g) * half;
        }
        st3(V.v, g, vi);
        st3(V.w, g, wi);
        bad |= nonfinite(vi, wi);
    }
    if (bad) {
        atomicOr(&V.flags[FLAG_NUMERIC], 1);
        atomicOr(&V.sflags[M.slot], HG_FAULT_NUMERIC);
    }
}

// All steps of one rebuild interval for one system, by one workgroup. With
// LIDS the system's partner ids (kmax slots per particle) are copied to LDS
// after its state records.
template <int BIN, int HAM, int PAIRS, int PREC, int LDS, bool LIDS, bool FAC>
__device__ __forceinline__ void run_system(const hg_params &P, const mixed_consts &MC, const pool_view &V,
                                           const pool_meta &M, facet_dev *F, uint32_t nsteps, double dt,
                                           uint32_t kmax) {
    const double half = dt / 2.0;
    uint32_t cur = M.cur;
    const uint32_t g0 = M.offset + threadIdx.x, g_end = M.offset + M.n, gs = blockDim.x;
    if constexpr (LDS != STATE_GLOBAL) lds_fill<LDS>(V, M);
    uint16_t *lids = nullptr;
    if constexpr (LIDS) {
        lids = reinterpret_cast<uint16_t *>(
            pool_lds + (LDS == STATE_GLOBAL ? 0 : size_t(M.n) * lds_stride<LDS>()));
        lds_ids_fill(V, M, kmax, lids);
    }
    if constexpr (LDS != STATE_GLOBAL || LIDS) __syncthreads();
    // step -1 is the initial acceleration pass of a system without valid
    // accelerations; one call site keeps a single copy of the pass in the code
    for (int s = M.acc_valid ? 0 : -1; s < int(nsteps); s++) {
        if (s >= 0) {
            pool_kick_drift<LDS>(V, M, s > 0, half, dt, g0, g_end, gs);
            if (s > 0 && threadIdx.x == 0) advance_facets(F, int(M.nf), dt);
            __syncthreads();
        }
        pool_force_phase<BIN, HAM, PAIRS, PREC, LDS, LIDS, FAC>(P, MC, V, M, cur, F, g0, g_end, gs, lids, kmax);
        cur ^= 1u;
        __syncthreads();
    }
    if (nsteps > 0) {
        pool_final_kick<LDS>(V, M, half, g0, g_end, gs);
        if (threadIdx.x == 0) advance_facets(F, int(M.nf), dt);
    } else if constexpr (LDS == STATE_FULL) {
        // only the initial pass ran: its accelerations go back to global
        for (uint32_t g = M.offset + threadIdx.x; g < M.offset + M.n; g += blockDim.x) {
            const double *q = lds_rec<LDS>(M.offset, g);
            st3(V.a, g, lds3(q + 12));
            st3(V.al, g, lds3(q + 15));
        }
    }
    __syncthreads();
}

// One system's interval by the whole workgroup, in the state tier its size
// allows (s_red: block_max_u32's scratch words).
template <int BIN, int HAM, int PAIRS, int PREC, bool FAC>
__device__ __forceinline__ void run_one(const hg_params &P, const mixed_consts &MC, const pool_view &V,
                                        uint32_t slot, uint32_t nsteps, double dt, uint32_t lds_bytes,
                                        uint32_t *s_red) {
    const pool_meta M = V.meta[slot];
    facet_dev *F = V.facets + size_t(slot) * V.max_facets;
    const size_t need = size_t(M.n) * sizeof(double);
    if constexpr (PAIRS == HG_PAIRS_NEIGHBOR_LIST) {
        // the partner ids go to LDS too wherever they fit beside the state
        uint32_t kmax = 0;
        for (uint32_t t = threadIdx.x; t < M.n; t += blockDim.x) kmax = max(kmax, V.ncnt[M.offset + t]);
        kmax = block_max_u32(kmax, s_red);
        const size_t ids = M.n <= 65535u ? lds_ids_bytes(M.n, kmax) : ~size_t(0) / 2;
        if (need * kFullStride + ids <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_FULL, true, FAC>(P, MC, V, M, F, nsteps, dt, kmax);
        else if (need * kMirrorStride + ids <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_MIRROR, true, FAC>(P, MC, V, M, F, nsteps, dt, kmax);
        else if (need * kPosStride + ids <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_POS, true, FAC>(P, MC, V, M, F, nsteps, dt, kmax);
        else if (ids <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_GLOBAL, true, FAC>(P, MC, V, M, F, nsteps, dt, kmax);
        else
            run_system<BIN, HAM, PAIRS, PREC, STATE_GLOBAL, false, FAC>(P, MC, V, M, F, nsteps, dt, 0);
    } else {
        if (need * kFullStride <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_FULL, false, FAC>(P, MC, V, M, F, nsteps, dt, 0);
        else if (need * kMirrorStride <= lds_bytes)
            run_system<BIN, HAM, PAIRS, PREC, STATE_MIRROR, false, FAC>(P, MC, V, M, F, nsteps, dt, 0);
        else
            run_system<BIN, HAM, PAIRS, PREC, STATE_GLOBAL, false, FAC>(P, MC, V, M, F, nsteps, dt, 0);
    }
}

// Several small systems (a pack: count members, all of them at most one
// particle per thread together) stepped side by side by one workgroup, so
// that a system of 100 particles does not leave 400 lanes idle through every
// step. Neighbour-list mode, all state in LDS: member m's particles take pack
// positions base_m .. base_m + n_m - 1 (thread = position). Each thread steps
// its own particle with the single-system device code, given its member's
// meta with offset_m - base_m in place of the offset: LDS records and partner
// ids are addressed by (arena index - that offset) = pack position (unsigned
// arithmetic), and the ids are stored as pack positions. Every operation of a
// particle is that of run_system, so a system's result does not depend on
// packing. Returns false, having done nothing, when the pack needs more
// threads or LDS than the workgroup has; the caller then steps the members
// one at a time.
constexpr uint32_t kMaxPack = 8;
template <int BIN, int HAM, int PREC, bool FAC>
__device__ __forceinline__ bool run_pack(const hg_params &P, const mixed_consts &MC, const pool_view &V,
                                         const uint32_t *slots, uint32_t cnt, uint32_t nsteps, double dt,
                                         uint32_t lds_bytes, uint32_t *s_red, pool_meta *s_pm, uint32_t *s_base) {
    if (threadIdx.x == 0) {
        uint32_t base = 0;
        for (uint32_t m = 0; m < cnt; m++) {
            s_pm[m] = V.meta[slots[m]];
            s_base[m] = base;
            base += s_pm[m].n;
        }
        s_base[cnt] = base;
    }
    __syncthreads();
    const uint32_t N = s_base[cnt], t = threadIdx.x;
    uint32_t m = 0;
    bool all_valid = true;
    for (uint32_t k = 0; k < cnt; k++) {
        if (k > 0 && t >= s_base[k]) m = k;
        all_valid = all_valid && s_pm[k].acc_valid;
    }
    const bool own = t < N;
    const pool_meta Pm = s_pm[m];
    const uint32_t g = Pm.offset + (t - s_base[m]);
    pool_meta Mv = Pm;
    Mv.offset = Pm.offset - s_base[m];
    Mv.n = N;
    uint32_t kmax = own ? V.ncnt[g] : 0u;
    kmax = block_max_u32(kmax, s_red);
    if (N > blockDim.x || size_t(N) * sizeof(double) * kFullStride + lds_ids_bytes(N, kmax) > lds_bytes) return false;
    facet_dev *F = V.facets + size_t(Pm.slot) * V.max_facets;
    uint16_t *lids = reinterpret_cast<uint16_t *>(pool_lds + size_t(N) * kFullStride);
    if (own) {
        double *q = pool_lds + size_t(t) * kFullStride;
        sds3(q, ld3(V.x, g));
        sds3(q + 3, ld3(V.v, g));
        sds3(q + 6, ld3(V.w, g));
        sds3(q + 9, ld3(V.th, g));
        sds3(q + 12, ld3(V.a, g));
        sds3(q + 15, ld3(V.al, g));
        const uint32_t c = V.ncnt[g];
        lids[kmax * N + t] = uint16_t(c);
        for (uint32_t k = 0; k < c; k++) lids[k * N + t] = uint16_t(V.nb[size_t(k) * V.stride + g] - Mv.offset);
    }
    __syncthreads();
    const double half = dt / 2.0;
    uint32_t cur = Pm.cur;
    auto advance_all = [&]() {
        for (uint32_t k = 0; k < cnt; k++)
            advance_facets(V.facets + size_t(s_pm[k].slot) * V.max_facets, int(s_pm[k].nf), dt);
    };
    for (int s = all_valid ? 0 : -1; s < int(nsteps); s++) {
        if (s >= 0) {
            if (own) pool_kick_drift<STATE_FULL>(V, Mv, s > 0, half, dt, g, g + 1, 1);
            if (s > 0 && t == 0) advance_all();
            __syncthreads();
        }
        if (own && (s >= 0 || !Pm.acc_valid)) {
            pool_force_phase<BIN, HAM, HG_PAIRS_NEIGHBOR_LIST, PREC, STATE_FULL, true, FAC>(P, MC, V, Mv, cur, F, g,
                                                                                           g + 1, 1, lids, kmax);
            cur ^= 1u;
        }
        __syncthreads();
    }
    if (nsteps > 0) {
        if (own) pool_final_kick<STATE_FULL>(V, Mv, half, g, g + 1, 1);
        if (t == 0) advance_all();
    } else if (own) {
        const double *q = pool_lds + size_t(t) * kFullStride;
        st3(V.a, g, lds3(q + 12));
        st3(V.al, g, lds3(q + 15));
    }
    __syncthreads();
    return true;
}

// One rebuild interval of every queued system. The queue holds packs
// (packs[i] = first | count << 24 over the slots in `order`): a pack of one
// is a system stepped by the whole workgroup, a larger one small systems
// stepped side by side (run_pack). Exit: every workgroup leaves once the
// shared counter passes `count`. lds_bytes of dynamic LDS per workgroup
// decide each system's state placement. FAC = false leaves the facet fold
// out (every queued system has nf = 0): it holds the kernel's largest
// register set. TPB threads per workgroup (256 or 512): a 512-thread
// workgroup alone on a CU has the same waves per SIMD as two of 256 and twice
// the LDS for one system.
template <int BIN, int HAM, int PAIRS, int PREC, bool FAC, int TPB>
// The model constants come by value: the scalar registers they spill to are
// re-read with v_readlane, which measured faster than reloading them through
// a pointer inside the pair loop (scalar loads add memory latency there).
__global__ void __launch_bounds__(TPB, HG_CHUNK_MIN_BLOCKS)
    k_pool_chunk(hg_params P, mixed_consts MC, pool_view V, const uint32_t *order, const uint32_t *packs,
                 uint32_t count, uint32_t *counter, uint32_t nsteps, double dt, uint32_t lds_bytes) {
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_red[TPB / 64];
    __shared__ pool_meta s_pm[kMaxPack];
    __shared__ uint32_t s_base[kMaxPack + 1];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= count) break;
        const uint32_t pk = packs[item];
        const uint32_t first = pk & 0xFFFFFFu, cnt = pk >> 24;
        if constexpr (PAIRS == HG_PAIRS_NEIGHBOR_LIST) {
            if (cnt > 1 &&
                run_pack<BIN, HAM, PREC, FAC>(P, MC, V, order + first, cnt, nsteps, dt, lds_bytes, s_red, s_pm, s_base))
                continue;
        }
        for (uint32_t k = 0; k < cnt; k++)
            run_one<BIN, HAM, PAIRS, PREC, FAC>(P, MC, V, order[first + k], nsteps, dt, lds_bytes, s_red);
    }
}

// ------------------------------------------------- wide (large) systems -----
// Systems of at least wide_min particles skip the chunk kernel, where one
// workgroup would carry the whole system and set the interval's critical
// path. They are stepped by per-step launches over all their particles at
// once, one thread per particle: wide_block b covers particles
// [start, start + kWideBlock) of system `slot`. The operations and their order
// per particle are those of run_system, so results are the same either way.
constexpr uint32_t kWideBlock = 128;
constexpr uint32_t kDefaultWideMin = 2048;
constexpr double kDefaultWideCuFraction = 0.25;
enum { TUNE_OFF = 0, TUNE_START = 1, TUNE_RUNNING = 2, TUNE_DONE = 3 };
struct wide_block {
    uint32_t slot, start;
};

// pass < 0: the initial acceleration pass of the systems without valid
// accelerations (others return); pass s >= 0: the pass of step s.
template <int BIN, int HAM, int PAIRS, int PREC>
__global__ void __launch_bounds__(kWideBlock) k_pool_wide_force(hg_params P, mixed_consts MC, pool_view V,
                                                                const wide_block *map, int pass) {
    const wide_block b = map[blockIdx.x];
    const pool_meta M = V.meta[b.slot];
    uint32_t cur = M.cur;
    if (pass < 0) {
        if (M.acc_valid) return;
    } else {
        cur ^= (uint32_t(pass) + (M.acc_valid ? 0u : 1u)) & 1u;
    }
    const uint32_t g = M.offset + b.start + threadIdx.x;
    const uint32_t end = M.offset + M.n;
    if (g >= end) return;
    const facet_dev *F = V.facets + size_t(b.slot) * V.max_facets;
    pool_force_phase<BIN, HAM, PAIRS, PREC, STATE_GLOBAL>(P, MC, V, M, cur, F, g, g + 1, 1);
}

// Kick/drift of step s (pending = s > 0 completes step s-1 and moves the
// facets by one step first, as run_system does).
__global__ void __launch_bounds__(kWideBlock) k_pool_wide_kick_drift(pool_view V, const wide_block *map, int pending,
                                                                     double half, double dt) {
    const wide_block b = map[blockIdx.x];
    const pool_meta M = V.meta[b.slot];
    if (pending && b.start == 0 && threadIdx.x == 0)
        advance_facets(V.facets + size_t(b.slot) * V.max_facets, int(M.nf), dt);
    const uint32_t g = M.offset + b.start + threadIdx.x;
    if (g >= M.offset + M.n) return;
    pool_kick_drift<STATE_GLOBAL>(V, M, pending != 0, half, dt, g, g + 1, 1);
}

__global__ void __launch_bounds__(kWideBlock) k_pool_wide_final_kick(pool_view V, const wide_block *map, double half,
                                                                     double dt) {
    const wide_block b = map[blockIdx.x];
    const pool_meta M = V.meta[b.slot];
    if (b.start == 0 && threadIdx.x == 0) advance_facets(V.facets + size_t(b.slot) * V.max_facets, int(M.nf), dt);
    const uint32_t g = M.offset + b.start + threadIdx.x;
    if (g >= M.offset + M.n) return;
    pool_final_kick<STATE_GLOBAL>(V, M, half, g, g + 1, 1);
}

// Fused wide steps (no wide system has facets, whose one-per-step advance
// other blocks of the same launch would race): positions, velocities and
// angular velocities alternate between two buffers, so that one launch can
// read every particle's step-s state (Vi) while it writes step s+1's (Vo).
// theta, the accelerations and the pair tables stay in place (Vi and Vo share
// them; only a particle's own thread touches its theta and accelerations).
// Kick/drift of the next step for particle g: pool_kick_drift's operations
// with pending = true, on the state and accelerations the pass leaves in
// registers, written into Vo (theta in place).
__device__ __forceinline__ void kick_drift_io(const pool_view &Vi, const pool_view &Vo, double half, double dt,
                                              uint32_t g, d3 x0, d3 vi, d3 wi, d3 ai, d3 ali, d3 th0) {
    vi = vi + ai * half;
    wi = wi + ali * half;
    vi = vi + ai * half;
    wi = wi + ali * half;
    const d3 xi = x0 + vi * dt;
    const d3 thi = th0 + wi * dt;
    st3(Vo.x, g, xi);
    st3(Vi.th, g, thi);
    st3(Vo.v, g, vi);
    st3(Vo.w, g, wi);
}

// The pass of step `pass` from Vi, then the kick/drift of step pass+1 into Vo.
template <int BIN, int HAM, int PAIRS, int PREC>
__global__ void __launch_bounds__(kWideBlock) k_pool_wide_step(hg_params P, mixed_consts MC, pool_view Vi,
                                                               pool_view Vo, const wide_block *map, int pass,
                                                               double half, double dt) {
    const wide_block b = map[blockIdx.x];
    const pool_meta M = Vi.meta[b.slot];
    const uint32_t cur = M.cur ^ ((uint32_t(pass) + (M.acc_valid ? 0u : 1u)) & 1u);
    const uint32_t g = M.offset + b.start + threadIdx.x;
    if (g >= M.offset + M.n) return;
    const facet_dev *F = Vi.facets + size_t(b.slot) * Vi.max_facets;
    const d3 th0 = ld3(Vi.th, g);
    pool_force_phase<BIN, HAM, PAIRS, PREC, STATE_GLOBAL, false, true>(
        P, MC, Vi, M, cur, F, g, g + 1, 1, nullptr, 0, [&](uint32_t gg, d3 x0, d3 v0, d3 w0, d3 a, d3 al) {
            kick_drift_io(Vi, Vo, half, dt, gg, x0, v0, w0, a, al, th0);
        });
}

// Final kick of the fused steps: pool_final_kick's operations from Vi into
// Vo (the pool's own buffers), positions copied across when they differ.
__global__ void __launch_bounds__(kWideBlock) k_pool_wide_final_kick_io(pool_view Vi, pool_view Vo,
                                                                        const wide_block *map, double half) {
    const wide_block b = map[blockIdx.x];
    const pool_meta M = Vi.meta[b.slot];
    const uint32_t g = M.offset + b.start + threadIdx.x;
    if (g >= M.offset + M.n) return;
    const d3 vi = ld3(Vi.v, g) + ld3(Vi.a, g) * half;
    const d3 wi = ld3(Vi.w, g) + ld3(Vi.al, g) * half;
    if (Vi.x != Vo.x) st3(Vo.x, g, ld3(Vi.x, g));
    st3(Vo.v, g, vi);
    st3(Vo.w, g, wi);
    if (nonfinite(vi, wi)) {
        atomicOr(&Vi.flags[FLAG_NUMERIC], 1);
        atomicOr(&Vi.sflags[M.slot], HG_FAULT_NUMERIC);
    }
}

// Sums the per-particle ledger terms in particle order, the order sum_energy
// uses on the host: lane c < 7 adds term c of every particle, then lane 0 adds
// the seven sums; out[0..7].
__global__ void __launch_bounds__(64) k_energy_sum(const double *terms, uint32_t n, double *out) {
    __shared__ double part[7];
    const uint32_t c = threadIdx.x;
    if (c < 7) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; i++) s += terms[8 * size_t(i) + c];
        part[c] = s;
    }
    __syncthreads();
    if (c < 7) out[c] = part[c];
    if (c == 0) {
        double t = 0.0;
        for (int q = 0; q < 7; q++) t += part[q];
        out[7] = t;
    }
}

// ------------------------------------------------------ batched rebuild -----
__device__ __forceinline__ uint32_t pool_hash(uint32_t slot, long long cx, long long cy, long long cz, uint32_t mask) {
    uint64_t h = uint64_t(cx) * 73856093ull ^ uint64_t(cy) * 19349663ull ^ uint64_t(cz) * 83492791ull ^
                 uint64_t(slot) * 2654435761ull;
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 32;
    return uint32_t(h) & mask;
}

__device__ __forceinline__ void pool_cell(d3 q, double cs, long long &cx, long long &cy, long long &cz) {
    cx = (long long)floor(q.x / cs);
    cy = (long long)floor(q.y / cs);
    cz = (long long)floor(q.z / cs);
}

__global__ void __launch_bounds__(256) k_pool_cell_count(const double *x, const uint32_t *owner, const pool_meta *meta,
                                                         uint32_t hw, double cs, uint32_t mask, uint32_t *cell_of,
                                                         uint32_t *bucket_count) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= hw) return;
    const uint32_t slot = owner[g];
    if (slot == kNoEntry || !meta[slot].rebuild) {
        cell_of[g] = kNoEntry;
        return;
    }
    long long cx, cy, cz;
    pool_cell(ld3(x, g), cs, cx, cy, cz);
    const uint32_t h = pool_hash(slot, cx, cy, cz, mask);
    cell_of[g] = h;
    atomicAdd(&bucket_count[h], 1u);
}

// A binned particle as the list fill reads it: its position, index and
// system in one 32-byte record, so that a bucket's candidates are one
// contiguous read each rather than an index followed by two gathers.
struct cell_rec {
    double x, y, z;
    uint32_t j, slot;
};

__global__ void __launch_bounds__(256) k_pool_cell_fill(const double *x, const uint32_t *owner, uint32_t hw,
                                                        const uint32_t *cell_of, const uint32_t *bucket_start,
                                                        uint32_t *bucket_fill, cell_rec *recs) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= hw) return;
    const uint32_t h = cell_of[g];
    if (h == kNoEntry) return;
    const d3 p = ld3(x, g);
    recs[bucket_start[h] + atomicAdd(&bucket_fill[h], 1u)] = cell_rec{p.x, p.y, p.z, g, owner[g]};
}

constexpr uint32_t kListRegs = 16;
constexpr uint32_t kRecBatch = 4;
__global__ void __launch_bounds__(256) k_pool_nl_fill(const double *x, const uint32_t *owner, const uint32_t *cell_of,
                                                      uint32_t hw, double cs, uint32_t mask, double r_verlet,
                                                      const uint32_t *bucket_count, const uint32_t *bucket_start,
                                                      const cell_rec *recs, uint32_t *nb, uint32_t *ncnt,
                                                      size_t stride, uint32_t cap, int *flags, int *sflags) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= hw || cell_of[g] == kNoEntry) return;
    const uint32_t slot = owner[g];
    const d3 xi = ld3(x, g);
    long long cx, cy, cz;
    pool_cell(xi, cs, cx, cy, cz);
    // squared-distance prefilter: a correctly rounded sqrt(d2) < r_verlet
    // implies d2 < r_verlet^2 (1 + 2^-52), so this bound drops no pair the
    // exact test keeps
    const double d2_max = r_verlet * r_verlet * (1.0 + 1e-12);
    // The list is kept sorted in registers (indices fixed at compile time,
    // so no scratch memory) while it has at most kListRegs entries, and in
    // nb itself by insertion once it grows past that.
    uint32_t r[kListRegs];
#pragma unroll
    for (uint32_t s = 0; s < kListRegs; s++) r[s] = 0;
    uint32_t cnt = 0;
    bool spilled = false;
    auto insert = [&](uint32_t j) {
        if (!spilled && cnt == kListRegs) {
#pragma unroll
            for (uint32_t s = 0; s < kListRegs; s++)
                if (s < cap) nb[size_t(s) * stride + g] = r[s];
            spilled = true;
        }
        if (spilled) {
            if (cnt < cap) {
                uint32_t pos = cnt;
                while (pos > 0) {
                    const uint32_t prev = nb[size_t(pos - 1) * stride + g];
                    if (prev < j) break;
                    nb[size_t(pos) * stride + g] = prev;
                    pos--;
                }
                nb[size_t(pos) * stride + g] = j;
            }
        } else {
            uint32_t pos = 0;
#pragma unroll
            for (uint32_t s = 0; s < kListRegs; s++) pos += (s < cnt && r[s] < j) ? 1u : 0u;
#pragma unroll
            for (uint32_t s = kListRegs - 1; s >= 1; s--) r[s] = s > pos ? r[s - 1] : (s == pos ? j : r[s]);
            r[0] = pos == 0 ? j : r[0];
        }
        cnt++;
    };
    // Cells are twice the list radius wide, so a partner lies in the
    // particle's own cell or the neighbour on the side of the half the
    // particle sits in, along each axis: 2 x 2 x 2 cells. (In cell units a
    // partner is less than 0.5 (1 - 1e-9) away; a particle in the lower half,
    // u - floor(u) < 0.5, has its partners in floor(u) - 1 .. floor(u), one in
    // the upper half in floor(u) .. floor(u) + 1.) The cells are unrolled, so
    // the hashes already visited (a bucket shared by two cells is scanned
    // once) sit in registers too.
    const long long bx = cx - (xi.x / cs - double(cx) < 0.5 ? 1 : 0);
    const long long by = cy - (xi.y / cs - double(cy) < 0.5 ? 1 : 0);
    const long long bz = cz - (xi.z / cs - double(cz) < 0.5 ? 1 : 0);
    uint32_t seen[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t h = pool_hash(slot, bx + (c & 1), by + (c >> 1 & 1), bz + (c >> 2), mask);
        seen[c] = h;
        bool dup = false;
#pragma unroll
        for (int s = 0; s < c; s++) dup |= seen[s] == h;
        if (dup) continue;
        const uint32_t b = bucket_start[h], e = b + bucket_count[h];
        // kRecBatch records loaded together (the last one again past the
        // end), then tested in bucket order: the loads overlap one another
        // instead of each waiting for the previous test
        for (uint32_t q = b; q < e; q += kRecBatch) {
            cell_rec c[kRecBatch];
#pragma unroll
            for (uint32_t k = 0; k < kRecBatch; k++) c[k] = recs[q + k < e ? q + k : e - 1];
#pragma unroll
            for (uint32_t k = 0; k < kRecBatch; k++) {
                const uint32_t j = c[k].j;
                if (q + k >= e || j == g || c[k].slot != slot) continue;
                const d3 d = xi - d3{c[k].x, c[k].y, c[k].z};
                if (!(dot(d, d) < d2_max)) continue;
                if (norm(d) < r_verlet) insert(j);
            }
        }
    }
    if (!spilled) {
#pragma unroll
        for (uint32_t s = 0; s < kListRegs; s++)
            if (s < cnt && s < cap) nb[size_t(s) * stride + g] = r[s];
    }
    ncnt[g] = cnt < cap ? cnt : cap;
    if (cnt > cap) {
        atomicOr(&flags[FLAG_NL_OVERFLOW], 1);
        atomicMax(&flags[FLAG_NL_MAXCOUNT], int(cnt));
        atomicOr(&sflags[slot], HG_FAULT_NEIGHBOR_CAPACITY);
    }
}

// Zeroes `rows` strided rows [offset, offset+n) of an array with row stride.
__global__ void k_zero_rows_f64(double *p, size_t stride, size_t offset, size_t n, uint32_t rows) {
    const size_t total = rows * n;
    for (size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x; k < total; k += size_t(gridDim.x) * blockDim.x)
        p[(k / n) * stride + offset + (k % n)] = 0.0;
}

inline unsigned nblk(size_t n, unsigned bs) { return unsigned((n + bs - 1) / bs); }

template <int BIN, int HAM, int PAIRS, int PREC, bool FAC, int TPB>
hipError_t launch_chunk_k(const hg_params &p, const pool_view &V, const uint32_t *order, const uint32_t *packs,
                          uint32_t count,
                          uint32_t *counter, uint32_t nsteps, double dt, unsigned grid, uint32_t lds_cap,
                          hipStream_t s) {
    const size_t lds_bytes = lds_cap;
    if (lds_bytes > 0) {
        hipError_t e =
            hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pool_chunk<BIN, HAM, PAIRS, PREC, FAC, TPB>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
        if (e != hipSuccess) return e;
    }
    k_pool_chunk<BIN, HAM, PAIRS, PREC, FAC, TPB><<<grid, TPB, lds_bytes, s>>>(p, make_mixed_consts(p), V, order,
                                                                                packs, count, counter, nsteps, dt,
                                                                                lds_cap);
    return hipGetLastError();
}

template <int BIN, int HAM, int PAIRS>
hipError_t launch_chunk_t(const hg_params &p, const pool_view &V, const uint32_t *order, const uint32_t *packs,
                          uint32_t count,
                          uint32_t *counter, uint32_t nsteps, double dt, unsigned grid, uint32_t lds_cap,
                          bool facets, int tpb, hipStream_t s) {
#define HG_KCASE(PREC, FAC, TPB)                                                                                \
    if (p.precision == PREC && facets == FAC && tpb == TPB)                                                    \
        return launch_chunk_k<BIN, HAM, PAIRS, PREC, FAC, TPB>(p, V, order, packs, count, counter, nsteps, dt, \
                                                                grid, lds_cap, s);
    HG_KCASE(HG_PRECISION_MIXED, true, 256)
    HG_KCASE(HG_PRECISION_MIXED, false, 256)
    HG_KCASE(HG_PRECISION_FP64, true, 256)
    HG_KCASE(HG_PRECISION_FP64, false, 256)
    HG_KCASE(HG_PRECISION_MIXED, true, 512)
    HG_KCASE(HG_PRECISION_MIXED, false, 512)
    HG_KCASE(HG_PRECISION_FP64, true, 512)
    HG_KCASE(HG_PRECISION_FP64, false, 512)
#undef HG_KCASE
    return hipErrorInvalidValue;
}

// facets: whether any queued system has facets (else the kernel without the
// facet fold runs); tpb: threads per workgroup (256 or 512).
template <int PAIRS>
hipError_t launch_chunk_p(const hg_params &p, const pool_view &V, const uint32_t *order, const uint32_t *packs,
                          uint32_t count,
                          uint32_t *counter, uint32_t nsteps, double dt, unsigned grid, uint32_t lds_cap,
                          bool facets, int tpb, hipStream_t s) {
#define HG_PCASE(B, H) \
    if (p.binary_model == B && p.hamaker_model == H) \
        return launch_chunk_t<B, H, PAIRS>(p, V, order, packs, count, counter, nsteps, dt, grid, lds_cap, facets, \
                                           tpb, s);
    HG_PCASE(HG_BINARY_NONE, HG_HAMAKER_NONE)
    HG_PCASE(HG_BINARY_NONE, HG_HAMAKER_FULL)
    HG_PCASE(HG_BINARY_NONE, HG_HAMAKER_LIMITED)
    HG_PCASE(HG_BINARY_NONE, HG_HAMAKER_SIMPLIFIED)
    HG_PCASE(HG_BINARY_CONTACT, HG_HAMAKER_NONE)
    HG_PCASE(HG_BINARY_CONTACT, HG_HAMAKER_FULL)
    HG_PCASE(HG_BINARY_CONTACT, HG_HAMAKER_LIMITED)
    HG_PCASE(HG_BINARY_CONTACT, HG_HAMAKER_SIMPLIFIED)
    HG_PCASE(HG_BINARY_SINTER, HG_HAMAKER_NONE)
    HG_PCASE(HG_BINARY_SINTER, HG_HAMAKER_FULL)
    HG_PCASE(HG_BINARY_SINTER, HG_HAMAKER_LIMITED)
    HG_PCASE(HG_BINARY_SINTER, HG_HAMAKER_SIMPLIFIED)
#undef HG_PCASE
    return hipErrorInvalidValue;
}

// Vo null: the pass alone (k_pool_wide_force); else the fused step into Vo.
template <int BIN, int HAM, int PAIRS>
hipError_t launch_wide_force_t(const hg_params &p, const pool_view &V, const wide_block *map, unsigned blocks,
                               int pass, hipStream_t s, const pool_view *Vo, double half, double dt) {
    const mixed_consts mc = make_mixed_consts(p);
    if (p.precision == HG_PRECISION_MIXED) {
        if (Vo)
            k_pool_wide_step<BIN, HAM, PAIRS, HG_PRECISION_MIXED>
                <<<blocks, kWideBlock, 0, s>>>(p, mc, V, *Vo, map, pass, half, dt);
        else
            k_pool_wide_force<BIN, HAM, PAIRS, HG_PRECISION_MIXED><<<blocks, kWideBlock, 0, s>>>(p, mc, V, map, pass);
    } else {
        if (Vo)
            k_pool_wide_step<BIN, HAM, PAIRS, HG_PRECISION_FP64>
                <<<blocks, kWideBlock, 0, s>>>(p, mc, V, *Vo, map, pass, half, dt);
        else
            k_pool_wide_force<BIN, HAM, PAIRS, HG_PRECISION_FP64><<<blocks, kWideBlock, 0, s>>>(p, mc, V, map, pass);
    }
    return hipGetLastError();
}

template <int PAIRS>
hipError_t launch_wide_force_p(const hg_params &p, const pool_view &V, const wide_block *map, unsigned blocks,
                               int pass, hipStream_t s, const pool_view *Vo = nullptr, double half = 0.0,
                               double dt = 0.0) {
#define HG_WCASE(B, H)                                                                                 \
    if (p.binary_model == B && p.hamaker_model == H)                                                   \
        return launch_wide_force_t<B, H, PAIRS>(p, V, map, blocks, pass, s, Vo, half, dt);
    HG_WCASE(HG_BINARY_NONE, HG_HAMAKER_NONE)
    HG_WCASE(HG_BINARY_NONE, HG_HAMAKER_FULL)
    HG_WCASE(HG_BINARY_NONE, HG_HAMAKER_LIMITED)
    HG_WCASE(HG_BINARY_NONE, HG_HAMAKER_SIMPLIFIED)
    HG_WCASE(HG_BINARY_CONTACT, HG_HAMAKER_NONE)
    HG_WCASE(HG_BINARY_CONTACT, HG_HAMAKER_FULL)
    HG_WCASE(HG_BINARY_CONTACT, HG_HAMAKER_LIMITED)
    HG_WCASE(HG_BINARY_CONTACT, HG_HAMAKER_SIMPLIFIED)
    HG_WCASE(HG_BINARY_SINTER, HG_HAMAKER_NONE)
    HG_WCASE(HG_BINARY_SINTER, HG_HAMAKER_FULL)
    HG_WCASE(HG_BINARY_SINTER, HG_HAMAKER_LIMITED)
    HG_WCASE(HG_BINARY_SINTER, HG_HAMAKER_SIMPLIFIED)
#undef HG_WCASE
    return hipErrorInvalidValue;
}

}  // namespace
}  // namespace hg

using namespace hg;

struct hg_pool {
    hg_params p{};
    int device = 0;  // the device current at hg_pool_create
    size_t P = 0, S = 0, MF = 0;
    // per-particle energy-ledger terms (hg_pool_energy), grown to the largest
    // system asked for
    double *energy = nullptr;
    size_t energy_cap = 0;
    hipStream_t stream = nullptr;
    double *x = nullptr, *v = nullptr, *th = nullptr, *w = nullptr, *a = nullptr, *al = nullptr;
    double *xb = nullptr, *vb = nullptr, *wb = nullptr;  // second x, v, omega for the fused wide steps
    uint32_t cap_n = 0, *nb = nullptr, *ncnt = nullptr;
    uint32_t cap_s = 0, *pj[2] = {nullptr, nullptr}, *pcnt[2] = {nullptr, nullptr};
    uint32_t *bj = nullptr, *bcnt = nullptr;  // neck strengths (sinter_breaking)
    double *bstr = nullptr;
    double *ps[2] = {nullptr, nullptr};
    uint32_t *owner = nullptr;
    facet_dev *facets = nullptr;
    double *ss = nullptr;
    pool_meta *meta_d = nullptr;
    std::vector<pool_meta> meta;
    std::vector<uint64_t> step_index;
    std::vector<double> t;
    std::vector<char> active;
    std::vector<char> parked;  // out of the launch order while its state is read back
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> bonds;
    std::vector<size_t> skipped;
    uint32_t *order_d = nullptr, *counter_d = nullptr;
    std::vector<uint32_t> order;   // every active system, largest first
    uint32_t narrow_count = 0;     // systems below wide_min, uploaded to order_d
    uint32_t *packs_d = nullptr;   // the chunk kernel's queue: first | count << 24 over order_d
    uint32_t pack_count = 0;
    bool packing = true;           // small systems share a workgroup (neighbour-list mode)
    uint32_t wide_min = 0;         // systems with n >= wide_min use the wide launches (0: none)
    wide_block *wide_map_d = nullptr;
    size_t wide_map_cap = 0;
    unsigned wide_blocks = 0;
    bool wide_pending_initial = false;
    // with wide systems present the wide launches run on stream_w beside the
    // chunk kernel, whose grid then leaves room worth wide_cus CUs
    hipStream_t stream_w = nullptr;
    hipStream_t stream_c = nullptr;  // copies that overlap an interval (pool_commit, pool_read_state)
    int n_cus = 0, wide_cus = 0;
    bool wide_graph = true;  // replay the wide launches of an interval as a HIP graph
    int tune_phase = 0;      // TUNE_*
    struct tune_cand {
        uint32_t wide_min;
        int wide_cus, tpb, wgs;  // launch split, chunk shape
    };
    std::vector<tune_cand> tune_cands;
    int tune_stage = 0;  // 0: chunk shapes, 1: launch splits
    std::vector<double> tune_scores;                   // seconds per step
    size_t tune_pos = 0;
    hipGraphExec_t wg_exec = nullptr;
    uint32_t wg_len = 0;
    bool wg_fused = false;
    unsigned wg_blocks = 0;
    double wg_dt = 0.0;
    hipEvent_t wev[3] = {nullptr, nullptr, nullptr};
    double last_wide_ms = 0.0;
    uint64_t last_wide_launches = 0;
    bool order_dirty = true;
    bool meta_dirty = true;
    uint32_t hash_size = 0, *cell_of = nullptr, *bucket_count = nullptr, *bucket_start = nullptr,
             *bucket_fill = nullptr, *scan_tmp = nullptr;
    cell_rec *cell_recs = nullptr;
    int *flags = nullptr, *flags_host = nullptr;
    // per-system fault words set by the kernels (device, cleared after every
    // read) and the faults recorded so far (host, per slot)
    int *sflags = nullptr, *sflags_host = nullptr;
    std::vector<int> fault;
    int fault_nl_need = 0, fault_pair_need = 0;  // counts the last overflows needed
    std::map<uint32_t, uint32_t> free_ranges;  // offset -> length
    uint32_t high_water = 0;
    std::vector<void *> allocations;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t evb[2] = {nullptr, nullptr};  // start and end of an interval's device work
    pool_meta *meta_pinned = nullptr;        // staging of the per-interval meta upload
    double last_busy_ms = 0.0;
    // the interval queued by pool_begin and not yet closed by pool_end
    struct {
        bool active = false;
        uint32_t len = 0, count = 0;
        bool rebuilt = false, wide = false;
        double dt = 0.0;
        std::chrono::steady_clock::time_point t0;
    } fl;
    double last_chunk_ms = 0.0, last_rebuild_ms = 0.0;
    uint64_t last_chunks = 0, last_chunk_steps = 0;
    int grid = 0;
    uint32_t lds_cap = 0;
    size_t lds_bytes_per_wg = 0;
    size_t lds_per_cu = 0, lds_per_block = 0;  // device limits
    int wg_per_cu = 2;                         // chunk-kernel workgroups per CU
    int chunk_tpb = 256;                       // and their threads
};

namespace {

template <typename T>
hg_status palloc(hg_pool *Q, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HG_ERR_OOM : HG_ERR_HIP,
                    "pool: hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes: " + hipGetErrorString(e));
    Q->allocations.push_back(*p);
    return HG_OK;
}

#define PG_TRY(expr)                  \
    do {                              \
        hg_status st_ = (expr);       \
        if (st_ != HG_OK) return st_; \
    } while (0)

uint32_t pool_cap_n(const hg_params &p) {
    if (p.pair_mode != HG_PAIRS_NEIGHBOR_LIST) return 0;
    if (p.neighbor_capacity) return p.neighbor_capacity;
    const double r = p.contact.r_part > 0 ? p.contact.r_part : (p.sinter.r_part > 0 ? p.sinter.r_part : p.hamaker.r_part);
    double ratio = r > 0 ? p.r_verlet / r : 64.0;
    double bound = 0.7405 * (std::pow(ratio + 1.0, 3.0) - 1.0) * 1.3;
    return uint32_t(std::ceil(std::max(bound, 8.0) / 8.0) * 8.0);
}

uint32_t pool_cap_s(const hg_params &p) {
    if (p.binary_model == HG_BINARY_NONE) return 0;
    return std::max<uint32_t>(p.pair_capacity ? p.pair_capacity : 24u, kTagPrefetch);
}

// Reads the flag words of the launches queued so far. When one is raised, the
// per-system fault words say which systems tripped it: they are added to
// Q->fault (those systems are no longer stepped) and pushed onto `faulted`,
// and both the flags and the fault words are cleared on the device.
hg_status collect_faults(hg_pool *Q, std::vector<int32_t> &faulted) {
    HG_HIP(hipMemcpyAsync(Q->flags_host, Q->flags, FLAG_COUNT * sizeof(int), hipMemcpyDeviceToHost, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    const int *f = Q->flags_host;
    if (!f[FLAG_NL_OVERFLOW] && !f[FLAG_PAIR_OVERFLOW] && !f[FLAG_NUMERIC]) return HG_OK;
    if (f[FLAG_NL_OVERFLOW]) Q->fault_nl_need = std::max(Q->fault_nl_need, f[FLAG_NL_MAXCOUNT]);
    if (f[FLAG_PAIR_OVERFLOW]) Q->fault_pair_need = std::max(Q->fault_pair_need, f[FLAG_PAIR_MAXCOUNT]);
    HG_HIP(hipMemcpyAsync(Q->sflags_host, Q->sflags, Q->S * sizeof(int), hipMemcpyDeviceToHost, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    for (size_t s = 0; s < Q->S; s++) {
        if (!Q->sflags_host[s] || !Q->active[s]) continue;
        if (!Q->fault[s]) faulted.push_back(int32_t(s));
        Q->fault[s] |= Q->sflags_host[s];
        Q->order_dirty = true;
    }
    HG_HIP(launch_zero_u32(reinterpret_cast<uint32_t *>(Q->flags), FLAG_COUNT, Q->stream));
    HG_HIP(launch_zero_u32(reinterpret_cast<uint32_t *>(Q->sflags), Q->S, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    return HG_OK;
}

// k chunk-kernel workgroups per CU (the grid), each with 1/k of the CU's
// LDS (less the static words) for its system's state and partner ids.
void set_chunk_wgs(hg_pool *Q, int k) {
    Q->wg_per_cu = k;
    Q->grid = Q->n_cus * k;
    size_t per_wg = std::min<size_t>(Q->lds_per_block, Q->lds_per_cu / size_t(k));
    per_wg = per_wg > 1024 ? per_wg - 1024 : 0;
    Q->lds_cap = uint32_t(per_wg);
    Q->lds_bytes_per_wg = per_wg;
}

pool_view make_pool_view(hg_pool *Q) {
    pool_view V;
    V.x = Q->x; V.v = Q->v; V.th = Q->th; V.w = Q->w; V.a = Q->a; V.al = Q->al;
    V.nb = Q->nb;
    V.ncnt = Q->ncnt;
    V.pj0 = Q->pj[0]; V.pj1 = Q->pj[1];
    V.pcnt0 = Q->pcnt[0]; V.pcnt1 = Q->pcnt[1];
    V.ps0 = Q->ps[0]; V.ps1 = Q->ps[1];
    V.cap_s = Q->cap_s;
    V.bj = Q->bj;
    V.bstr = Q->bstr;
    V.bcnt = Q->bcnt;
    V.cap_n = Q->cap_n;
    V.facets = Q->facets;
    V.max_facets = uint32_t(Q->MF);
    V.ss = Q->ss;
    V.meta = Q->meta_d;
    V.stride = Q->P;
    V.flags = Q->flags;
    V.sflags = Q->sflags;
    return V;
}

hg_status valid_slot(hg_pool *Q, int32_t slot) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (slot < 0 || size_t(slot) >= Q->S || !Q->active[slot]) return fail(HG_ERR_INVALID, "no such system in the pool");
    return HG_OK;
}

hg_status upload_meta(hg_pool *Q) {
    HG_HIP(hipMemcpyAsync(Q->meta_d, Q->meta.data(), Q->S * sizeof(pool_meta), hipMemcpyHostToDevice, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    Q->meta_dirty = false;
    return HG_OK;
}

// Copies strided rows [offset, offset+n) of a device array into a host array
// with row stride n; `width` elements per particle (a row of the device array
// holds P*width elements).
template <typename T>
hg_status download_rows(hg_pool *Q, const T *dev, uint32_t offset, uint32_t n, uint32_t rows, T *host,
                        size_t width = 1) {
    if (rows == 0 || n == 0) return HG_OK;
    const size_t row = n * width * sizeof(T);
    HG_HIP(hipMemcpy2DAsync(host, row, dev + offset * width, Q->P * width * sizeof(T), row, rows,
                            hipMemcpyDeviceToHost, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    return HG_OK;
}

hg_status rebuild_due(hg_pool *Q) {
    const uint32_t hw = Q->high_water;
    if (hw == 0) return HG_OK;
    const double cs = 2.0 * Q->p.r_verlet * (1.0 + 1e-9);  // two list radii: 2 x 2 x 2 cells per particle
    const uint32_t mask = Q->hash_size - 1;
    HG_HIP(launch_zero_u32(Q->bucket_count, Q->hash_size, Q->stream));
    HG_HIP(launch_zero_u32(Q->bucket_fill, Q->hash_size, Q->stream));
    k_pool_cell_count<<<nblk(hw, 256), 256, 0, Q->stream>>>(Q->x, Q->owner, Q->meta_d, hw, cs, mask, Q->cell_of,
                                                           Q->bucket_count);
    HG_HIP(hipGetLastError());
    HG_HIP(launch_exclusive_scan(Q->bucket_count, Q->bucket_start, Q->hash_size, Q->scan_tmp, Q->stream));
    k_pool_cell_fill<<<nblk(hw, 256), 256, 0, Q->stream>>>(Q->x, Q->owner, hw, Q->cell_of, Q->bucket_start,
                                                          Q->bucket_fill, Q->cell_recs);
    HG_HIP(hipGetLastError());
    k_pool_nl_fill<<<nblk(hw, 64), 64, 0, Q->stream>>>(Q->x, Q->owner, Q->cell_of, hw, cs, mask, Q->p.r_verlet,
                                                      Q->bucket_count, Q->bucket_start, Q->cell_recs, Q->nb,
                                                      Q->ncnt, Q->P, Q->cap_n, Q->flags, Q->sflags);
    HG_HIP(hipGetLastError());
    return HG_OK;
}

// Inputs of the energy ledger of one system of the pool.
energy_view pool_energy_view(hg_pool *Q, const pool_meta &M) {
    const bool nl = Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST, tab = Q->cap_s != 0;
    return energy_view{Q->x,
                       Q->v,
                       Q->w,
                       M.offset,
                       M.n,
                       Q->P,
                       nl ? Q->nb : nullptr,
                       nl ? Q->ncnt : nullptr,
                       tab ? Q->pj[M.cur] : nullptr,
                       tab ? Q->pcnt[M.cur] : nullptr,
                       tab ? Q->ps[M.cur] : nullptr,
                       nullptr,
                       nullptr,
                       nullptr,
                       nullptr};
}

// The device address of `scratch` on device dev: looked up once per device,
// with that device current, the first time a call reaches it; later calls
// reuse it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

hg_status noalloc_scratch(int dev, double **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(HG_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        HG_HIP(hipGetDevice(&cur));
        if (cur != dev) HG_HIP(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        HG_HIP(e);
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return HG_OK;
}

}  // namespace

extern "C" {

size_t hg_pool_bytes_per_particle(const hg_params *p, size_t max_facets) {
    if (!p) return 0;
    const size_t cap_n = pool_cap_n(*p), cap_s = pool_cap_s(*p);
    size_t b = 9 * 32;                   // x, v, theta, omega, a, alpha; second x, v, omega
    b += cap_n * 4 + 4;                  // neighbour slots + count
    b += 2 * (cap_s * (4 + 8 * kSpringRec) + 4);  // pair-state table, double buffered
    b += max_facets * 72;                // facet springs
    b += 4;                              // owner
    if (p->binary_model == HG_BINARY_SINTER && p->sinter_breaking)
        b += kMaxBondsPerParticle * 12 + 4;  // neck-strength lists
    if (p->pair_mode == HG_PAIRS_NEIGHBOR_LIST) b += 4 + 32 + 3 * 2 * 4 + 4;  // cell list scratch
    return b;
}

hg_status hg_pool_create(const hg_params *params, size_t max_particles, size_t max_systems, size_t max_facets,
                         hg_pool **out) {
    if (!out || !params) return fail(HG_ERR_INVALID, "hg_pool_create: params and out are required");
    *out = nullptr;
    if (max_particles == 0 || max_particles >= (size_t(1) << 31) || max_systems == 0 ||
        max_systems >= (size_t(1) << 24))
        return fail(HG_ERR_INVALID, "hg_pool_create: need 0 < max_particles < 2^31 and 0 < max_systems < 2^24");
    if (max_facets > size_t(kMaxFacets)) return fail(HG_ERR_INVALID, "hg_pool_create: at most 16 facets per system");
    if (params->precision != HG_PRECISION_FP64 && params->precision != HG_PRECISION_MIXED)
        return fail(HG_ERR_INVALID, "hg_pool_create: unknown precision");
    if (params->binary_model < HG_BINARY_NONE || params->binary_model > HG_BINARY_SINTER)
        return fail(HG_ERR_INVALID, "hg_pool_create: pools run no binary functor, contact_force, alt_sinter or "
                                    "alt_breaking_sinter; the legacy sinter_functor runs on single systems");
    if (params->hamaker_model < HG_HAMAKER_NONE || params->hamaker_model > HG_HAMAKER_SIMPLIFIED)
        return fail(HG_ERR_INVALID, "hg_pool_create: unknown hamaker_model");
    if (params->sinter_breaking != 0 && (params->sinter_breaking != 1 || params->binary_model != HG_BINARY_SINTER))
        return fail(HG_ERR_INVALID, "hg_pool_create: sinter_breaking is 0, or 1 with HG_BINARY_SINTER");
    if (params->pair_mode == HG_PAIRS_NEIGHBOR_LIST && !(params->r_verlet > 0.0))
        return fail(HG_ERR_INVALID, "hg_pool_create: neighbour-list mode needs r_verlet > 0");
    if (params->facet_advance != 0)
        return fail(HG_ERR_INVALID, "hg_pool_create: pools advance facets after every step (facet_advance = 0)");
    int dev = 0;
    HG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HG_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(HG_ERR_NO_DEVICE, std::string("hipgran is built for gfx950; current device is ") + prop.gcnArchName);

    hg_pool *Q = new hg_pool();
    auto bail = [&](hg_status st) {
        std::string msg = hg_last_error();
        hg_pool_destroy(Q);
        fail(st, msg);
        return st;
    };
#define PG_TRYC(expr)                           \
    do {                                        \
        hg_status st_ = (expr);                 \
        if (st_ != HG_OK) return bail(st_);     \
    } while (0)
#define PG_HIPC(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bail(fail(HG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)
    Q->p = *params;
    Q->device = dev;
    Q->P = max_particles;
    Q->S = max_systems;
    Q->MF = max_facets;
    Q->n_cus = prop.multiProcessorCount;
    Q->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    Q->lds_per_block = prop.sharedMemPerBlock;
    set_chunk_wgs(Q, 2);
    PG_HIPC(hipStreamCreateWithFlags(&Q->stream, hipStreamNonBlocking));
    for (auto &e : Q->ev) PG_HIPC(hipEventCreate(&e));
    for (auto &e : Q->wev) PG_HIPC(hipEventCreate(&e));
    for (auto &e : Q->evb) PG_HIPC(hipEventCreate(&e));
    Q->wide_min = kDefaultWideMin;
    Q->wide_cus = int(std::lround(kDefaultWideCuFraction * Q->n_cus));
    PG_HIPC(hipStreamCreateWithFlags(&Q->stream_w, hipStreamNonBlocking));
    PG_HIPC(hipStreamCreateWithFlags(&Q->stream_c, hipStreamNonBlocking));
    const size_t P = Q->P;
    double **fields[9] = {&Q->x, &Q->v, &Q->th, &Q->w, &Q->a, &Q->al, &Q->xb, &Q->vb, &Q->wb};
    for (auto f : fields) PG_TRYC(palloc(Q, f, 4 * P));
    Q->cap_n = pool_cap_n(Q->p);
    Q->cap_s = pool_cap_s(Q->p);
    if (Q->cap_n) {
        PG_TRYC(palloc(Q, &Q->nb, size_t(Q->cap_n) * P));
        PG_TRYC(palloc(Q, &Q->ncnt, P));
        PG_HIPC(launch_zero_u32(Q->ncnt, P, Q->stream));
        uint32_t hs = 1024;
        while (hs < 2 * P && hs < (1u << 30)) hs <<= 1;
        Q->hash_size = hs;
        PG_TRYC(palloc(Q, &Q->cell_of, P));
        PG_TRYC(palloc(Q, &Q->cell_recs, P));
        PG_TRYC(palloc(Q, &Q->bucket_count, hs));
        PG_TRYC(palloc(Q, &Q->bucket_start, hs));
        PG_TRYC(palloc(Q, &Q->bucket_fill, hs));
        PG_TRYC(palloc(Q, &Q->scan_tmp, scan_tmp_words(hs)));
    }
    if (Q->cap_s) {
        for (int b = 0; b < 2; b++) {
            PG_TRYC(palloc(Q, &Q->pj[b], size_t(Q->cap_s) * P));
            PG_TRYC(palloc(Q, &Q->ps[b], size_t(Q->cap_s) * kSpringRec * P));
            PG_TRYC(palloc(Q, &Q->pcnt[b], P));
        }
    }
    if (Q->p.binary_model == HG_BINARY_SINTER && Q->p.sinter_breaking) {
        PG_TRYC(palloc(Q, &Q->bj, size_t(kMaxBondsPerParticle) * P));
        PG_TRYC(palloc(Q, &Q->bstr, size_t(kMaxBondsPerParticle) * P));
        PG_TRYC(palloc(Q, &Q->bcnt, P));
        PG_HIPC(launch_zero_u32(Q->bcnt, P, Q->stream));
    }
    PG_TRYC(palloc(Q, &Q->owner, P));
    {
        std::vector<uint32_t> none(P, kNoEntry);
        PG_HIPC(hipMemcpy(Q->owner, none.data(), P * 4, hipMemcpyHostToDevice));
    }
    PG_TRYC(palloc(Q, &Q->facets, std::max<size_t>(1, Q->S * std::max<size_t>(Q->MF, 1))));
    if (Q->MF) PG_TRYC(palloc(Q, &Q->ss, Q->MF * 9 * P));
    PG_TRYC(palloc(Q, &Q->meta_d, Q->S));
    PG_TRYC(palloc(Q, &Q->order_d, Q->S));
    PG_TRYC(palloc(Q, &Q->packs_d, Q->S));
    Q->wide_map_cap = Q->P / kWideBlock + Q->S + 1;
    PG_TRYC(palloc(Q, &Q->wide_map_d, Q->wide_map_cap));
    PG_TRYC(palloc(Q, &Q->counter_d, 1));
    PG_TRYC(palloc(Q, &Q->flags, FLAG_COUNT));
    PG_TRYC(palloc(Q, &Q->sflags, Q->S));
    PG_HIPC(hipHostMalloc(reinterpret_cast<void **>(&Q->flags_host), FLAG_COUNT * sizeof(int)));
    PG_HIPC(hipHostMalloc(reinterpret_cast<void **>(&Q->sflags_host), Q->S * sizeof(int)));
    PG_HIPC(hipHostMalloc(reinterpret_cast<void **>(&Q->meta_pinned), Q->S * sizeof(pool_meta)));
    PG_HIPC(launch_zero_u32(reinterpret_cast<uint32_t *>(Q->flags), FLAG_COUNT, Q->stream));
    PG_HIPC(launch_zero_u32(reinterpret_cast<uint32_t *>(Q->sflags), Q->S, Q->stream));
    Q->meta.assign(Q->S, pool_meta{0, 0, 0, 0, 0, 0, 0, 0});
    Q->step_index.assign(Q->S, 0);
    Q->t.assign(Q->S, 0.0);
    Q->active.assign(Q->S, 0);
    Q->parked.assign(Q->S, 0);
    Q->fault.assign(Q->S, 0);
    Q->bonds.assign(Q->S, {});
    Q->skipped.assign(Q->S, 0);
    Q->free_ranges[0] = uint32_t(P);
    PG_TRYC(upload_meta(Q));
    *out = Q;
    return HG_OK;
#undef PG_TRYC
#undef PG_HIPC
}

hg_status hg_pool_destroy(hg_pool *Q) {
    if (!Q) return HG_OK;
    if (Q->stream) (void)hipStreamSynchronize(Q->stream);
    if (Q->stream_w) (void)hipStreamSynchronize(Q->stream_w);
    if (Q->stream_c) (void)hipStreamSynchronize(Q->stream_c);
    for (void *p : Q->allocations) (void)hipFree(p);
    for (auto e : Q->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : Q->wev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : Q->evb)
        if (e) (void)hipEventDestroy(e);
    if (Q->meta_pinned) (void)hipHostFree(Q->meta_pinned);
    if (Q->wg_exec) (void)hipGraphExecDestroy(Q->wg_exec);
    if (Q->stream_w) (void)hipStreamDestroy(Q->stream_w);
    if (Q->stream_c) (void)hipStreamDestroy(Q->stream_c);
    if (Q->flags_host) (void)hipHostFree(Q->flags_host);
    if (Q->sflags_host) (void)hipHostFree(Q->sflags_host);
    if (Q->stream) (void)hipStreamDestroy(Q->stream);
    delete Q;
    return HG_OK;
}

}  // extern "C"

namespace hg {

hg_status pool_prepare(const hg_params &p, size_t n, const double *x0, const double *v0, const double *theta0,
                       const double *omega0, const std::vector<std::pair<uint32_t, uint32_t>> *given_bonds,
                       size_t given_skipped, pool_prep &out) {
    if (!x0) return fail(HG_ERR_INVALID, "hg_pool_add: x0 is required");
    if (n == 0) return fail(HG_ERR_INVALID, "hg_pool_add: empty system");
    out.n = n;
    out.bonds.clear();
    out.skipped = 0;
    out.rows.assign(p.binary_model == HG_BINARY_SINTER ? n : 0, {});
    if (p.binary_model == HG_BINARY_SINTER) {
        if (given_bonds) {
            out.bonds = *given_bonds;
            out.skipped = given_skipped;
        } else {
            sinter_topology_host(n, x0, p.sinter.r_part, p.sinter.d_crit, out.bonds, out.skipped, !p.sinter_breaking);
        }
        for (auto &b : out.bonds) {
            out.rows[b.first].push_back(b.second | kBondBit);
            out.rows[b.second].push_back(b.first | kBondBit);
        }
        for (auto &r : out.rows)
            std::sort(r.begin(), r.end(), [](uint32_t a, uint32_t b) { return (a & kIdxMask) < (b & kIdxMask); });
    }
    const double *src[4] = {x0, v0, theta0, omega0};
    for (int f = 0; f < 4; f++) {
        out.q4[f].assign(4 * n, 0.0);
        if (src[f])
            for (size_t i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) out.q4[f][4 * i + c] = src[f][3 * i + c];
    }
    return HG_OK;
}

hg_status pool_commit(hg_pool *Q, pool_prep &prep, int32_t *slot_out, hipStream_t cs) {
    if (!cs && Q) cs = Q->stream;
    if (!Q || !slot_out) return fail(HG_ERR_INVALID, "hg_pool_add: pool and slot are required");
    const size_t n = prep.n;
    for (auto &r : prep.rows)
        if (r.size() > Q->cap_s || (Q->bcnt && r.size() > kMaxBondsPerParticle))
            return fail(HG_ERR_CAPACITY, "hg_pool_add: a particle has " + std::to_string(r.size()) +
                                             " bonds, pair capacity is " + std::to_string(Q->cap_s));
    int32_t slot = -1;
    for (size_t s = 0; s < Q->S; s++)
        if (!Q->active[s]) {
            slot = int32_t(s);
            break;
        }
    if (slot < 0) return fail(HG_ERR_CAPACITY, "hg_pool_add: every system slot is in use");
    auto it = Q->free_ranges.begin();
    for (; it != Q->free_ranges.end(); ++it)
        if (it->second >= n) break;
    if (it == Q->free_ranges.end()) return fail(HG_ERR_CAPACITY, "hg_pool_add: no free particle range of that size");
    const uint32_t offset = it->first;
    const uint32_t len = it->second;
    Q->free_ranges.erase(it);
    if (len > n) Q->free_ranges[offset + uint32_t(n)] = len - uint32_t(n);
    Q->high_water = std::max<uint32_t>(Q->high_water, offset + uint32_t(n));

    double *dst[4] = {Q->x, Q->v, Q->th, Q->w};
    for (int f = 0; f < 4; f++)
        HG_HIP(hipMemcpyAsync(dst[f] + 4 * size_t(offset), prep.q4[f].data(), prep.q4[f].size() * 8,
                              hipMemcpyHostToDevice, cs));
    HG_HIP(launch_zero_f64(Q->a + 4 * size_t(offset), 4 * n, cs));
    HG_HIP(launch_zero_f64(Q->al + 4 * size_t(offset), 4 * n, cs));
    std::vector<uint32_t> own(n, uint32_t(slot));
    HG_HIP(hipMemcpyAsync(Q->owner + offset, own.data(), n * 4, hipMemcpyHostToDevice, cs));
    std::vector<uint32_t> pj, cnt;
    if (Q->cap_s) {
        pj.assign(size_t(Q->cap_s) * n, kNoEntry);
        cnt.assign(n, 0u);
        for (size_t i = 0; i < prep.rows.size(); i++) {
            cnt[i] = uint32_t(prep.rows[i].size());
            for (size_t s = 0; s < prep.rows[i].size(); s++)
                pj[s * n + i] = (prep.rows[i][s] & kBondBit) | (offset + (prep.rows[i][s] & kIdxMask));
        }
        HG_HIP(hipMemcpy2DAsync(Q->pj[0] + offset, Q->P * 4, pj.data(), n * 4, n * 4, Q->cap_s,
                                hipMemcpyHostToDevice, cs));
        HG_HIP(hipMemcpyAsync(Q->pcnt[0] + offset, cnt.data(), n * 4, hipMemcpyHostToDevice, cs));
        k_zero_rows_f64<<<256, 256, 0, cs>>>(Q->ps[0], size_t(Q->P) * kSpringRec, size_t(offset) * kSpringRec,
                                                     n * kSpringRec, Q->cap_s);
        HG_HIP(hipGetLastError());
    }
    if (Q->cap_n) HG_HIP(launch_zero_u32(Q->ncnt + offset, n, cs));
    // the host buffers above are read by the copies: wait before they go
    HG_HIP(hipStreamSynchronize(cs));
    Q->meta[slot] = pool_meta{offset, uint32_t(n), 0, 0, 0, 1, uint32_t(slot), 0};
    Q->step_index[slot] = 0;
    Q->t[slot] = 0.0;
    Q->active[slot] = 1;
    Q->fault[slot] = 0;
    Q->bonds[slot] = std::move(prep.bonds);
    Q->skipped[slot] = prep.skipped;
    Q->order_dirty = true;
    Q->meta_dirty = true;
    *slot_out = slot;
    if (Q->bcnt) {  // neck strengths start infinite
        std::vector<double> inf(Q->bonds[slot].size(), std::numeric_limits<double>::infinity());
        PG_TRY(pool_set_necks(Q, slot, inf.data(), cs));
    }
    return HG_OK;
}

}  // namespace hg

namespace {
// hg_pool_add; `given_bonds` (with `given_skipped`) replaces the bond search
// on x0 (a system reloaded from a snapshot keeps the bonds it was built with).
hg_status pool_add_impl(hg_pool *Q, size_t n, const double *x0, const double *v0, const double *theta0,
                        const double *omega0, const std::vector<std::pair<uint32_t, uint32_t>> *given_bonds,
                        size_t given_skipped, int32_t *slot_out) {
    if (!Q || !x0 || !slot_out) return fail(HG_ERR_INVALID, "hg_pool_add: pool, x0 and slot are required");
    pool_prep prep;
    PG_TRY(pool_prepare(Q->p, n, x0, v0, theta0, omega0, given_bonds, given_skipped, prep));
    return pool_commit(Q, prep, slot_out, Q->stream);
}
}  // namespace

extern "C" {

hg_status hg_pool_add(hg_pool *Q, size_t n, const double *x0, const double *v0, const double *theta0,
                      const double *omega0, int32_t *slot_out) {
    return pool_add_impl(Q, n, x0, v0, theta0, omega0, nullptr, 0, slot_out);
}

hg_status hg_pool_set_neck_strengths(hg_pool *Q, int32_t slot, const double *strengths) {
    return pool_set_necks(Q, slot, strengths, Q ? Q->stream : nullptr);
}

}  // extern "C"

namespace hg {

// hg_pool_set_neck_strengths with the copies on stream cs (the copy stream
// while an interval runs: the slot is not in the launch order).
hg_status pool_set_necks(hg_pool *Q, int32_t slot, const double *strengths, hipStream_t cs) {
    PG_TRY(valid_slot(Q, slot));
    if (!strengths) return fail(HG_ERR_INVALID, "strengths is null");
    if (!Q->bcnt) return fail(HG_ERR_INVALID, "hg_pool_set_neck_strengths: the pool's bonds do not break "
                                              "(params.sinter_breaking)");
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    std::vector<uint32_t> bj(size_t(kMaxBondsPerParticle) * n, kNoEntry), cnt(n, 0);
    std::vector<double> bs(size_t(kMaxBondsPerParticle) * n, 0.0);
    for (size_t b = 0; b < Q->bonds[slot].size(); b++) {
        for (int end = 0; end < 2; end++) {
            const uint32_t i = end ? Q->bonds[slot][b].second : Q->bonds[slot][b].first;
            const uint32_t j = end ? Q->bonds[slot][b].first : Q->bonds[slot][b].second;
            bj[size_t(cnt[i]) * n + i] = off + j;  // arena index, as the pair-state tags
            bs[size_t(cnt[i]) * n + i] = strengths[b];
            cnt[i]++;
        }
    }
    HG_HIP(hipMemcpy2DAsync(Q->bj + off, Q->P * 4, bj.data(), n * 4, n * 4, kMaxBondsPerParticle,
                            hipMemcpyHostToDevice, cs));
    HG_HIP(hipMemcpy2DAsync(Q->bstr + off, Q->P * 8, bs.data(), n * 8, n * 8, kMaxBondsPerParticle,
                            hipMemcpyHostToDevice, cs));
    HG_HIP(hipMemcpyAsync(Q->bcnt + off, cnt.data(), n * 4, hipMemcpyHostToDevice, cs));
    HG_HIP(hipStreamSynchronize(cs));
    return HG_OK;
}

}  // namespace hg

namespace hg {

hg_status pool_release(hg_pool *Q, int32_t slot, hipStream_t cs) {
    PG_TRY(valid_slot(Q, slot));
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    {
        std::vector<uint32_t> none(n, kNoEntry);
        HG_HIP(hipMemcpyAsync(Q->owner + off, none.data(), n * 4, hipMemcpyHostToDevice, cs));
        HG_HIP(hipStreamSynchronize(cs));
    }
    Q->active[slot] = 0;
    Q->parked[slot] = 0;
    Q->fault[slot] = 0;
    Q->meta[slot] = pool_meta{0, 0, 0, 0, 0, 0, 0, 0};
    Q->bonds[slot].clear();
    Q->order_dirty = true;
    // free the range and coalesce with neighbours
    auto it = Q->free_ranges.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != Q->free_ranges.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        Q->free_ranges.erase(nx);
    }
    if (it != Q->free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            Q->free_ranges.erase(it);
        }
    }
    // lower the high-water mark when the top range is free
    auto last = std::prev(Q->free_ranges.end());
    if (last->first + last->second == Q->P) Q->high_water = std::min<uint32_t>(Q->high_water, last->first);
    Q->meta_dirty = true;
    return HG_OK;
}

hg_status pool_read_state(hg_pool *Q, int32_t slot, double *x, double *v, double *theta, double *omega,
                          hipStream_t cs) {
    PG_TRY(valid_slot(Q, slot));
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    std::vector<double> buf(4 * size_t(n));
    const double *src[4] = {Q->x, Q->v, Q->th, Q->w};
    double *dst[4] = {x, v, theta, omega};
    for (int f = 0; f < 4; f++) {
        if (!dst[f]) continue;
        HG_HIP(hipMemcpyAsync(buf.data(), src[f] + 4 * size_t(off), buf.size() * 8, hipMemcpyDeviceToHost, cs));
        HG_HIP(hipStreamSynchronize(cs));
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) dst[f][3 * i + c] = buf[4 * i + c];
    }
    return HG_OK;
}

hipStream_t pool_copy_stream(hg_pool *Q) { return Q->stream_c; }

}  // namespace hg

extern "C" {

hg_status hg_pool_remove(hg_pool *Q, int32_t slot) {
    PG_TRY(valid_slot(Q, slot));
    HG_HIP(hipStreamSynchronize(Q->stream));
    return pool_release(Q, slot, Q->stream);
}

}  // extern "C"

namespace hg {

hg_status pool_set_facets(hg_pool *Q, int32_t slot, size_t nf, const double *verts, const double *v_facet,
                          hipStream_t cs) {
    PG_TRY(valid_slot(Q, slot));
    if (nf > Q->MF) return fail(HG_ERR_INVALID, "hg_pool_set_facets: more facets than the pool's max_facets");
    if (nf > 0 && (!verts || !v_facet)) return fail(HG_ERR_INVALID, "hg_pool_set_facets: verts and v_facet required");
    std::vector<facet_dev> F(std::max<size_t>(nf, 1));
    for (size_t f = 0; f < nf; f++) {
        const double *q = verts + 9 * f;
        for (int k = 0; k < 3; k++) F[f].vert[k] = d3{q[3 * k], q[3 * k + 1], q[3 * k + 2]};
        F[f].v_facet = d3{v_facet[3 * f], v_facet[3 * f + 1], v_facet[3 * f + 2]};
        d3 u = F[f].vert[1] - F[f].vert[0], w = F[f].vert[2] - F[f].vert[0];
        F[f].area = 1.0 / 2.0 * norm(cross(u, w));
        F[f].normal = normalized(cross(u, w));
        F[f].pad[0] = F[f].pad[1] = 0.0;
    }
    if (nf > 0) {
        HG_HIP(hipMemcpyAsync(Q->facets + size_t(slot) * Q->MF, F.data(), nf * sizeof(facet_dev),
                              hipMemcpyHostToDevice, cs));
        k_zero_rows_f64<<<256, 256, 0, cs>>>(Q->ss, Q->P, Q->meta[slot].offset, Q->meta[slot].n,
                                                    uint32_t(nf * 9));
        HG_HIP(hipGetLastError());
    }
    HG_HIP(hipStreamSynchronize(cs));
    Q->meta[slot].nf = uint32_t(nf);
    Q->meta[slot].acc_valid = 0;
    Q->meta_dirty = true;
    return HG_OK;
}

}  // namespace hg

extern "C" {

hg_status hg_pool_set_facets(hg_pool *Q, int32_t slot, size_t nf, const double *verts, const double *v_facet) {
    return pool_set_facets(Q, slot, nf, verts, v_facet, Q ? Q->stream : nullptr);
}

namespace {

// Sorts the active systems largest first, splits off those that take the
// wide launches and uploads the chunk kernel's queue and the wide block map.
hg_status refresh_order(hg_pool *Q) {
    Q->order.clear();
    for (size_t s = 0; s < Q->S; s++)
        if (Q->active[s] && !Q->fault[s] && !Q->parked[s]) Q->order.push_back(uint32_t(s));
    std::stable_sort(Q->order.begin(), Q->order.end(),
                     [&](uint32_t a, uint32_t b) { return Q->meta[a].n > Q->meta[b].n; });
    std::vector<wide_block> map;
    size_t first_narrow = 0;
    while (first_narrow < Q->order.size() && Q->wide_min && Q->meta[Q->order[first_narrow]].n >= Q->wide_min) {
        const uint32_t s = Q->order[first_narrow++];
        for (uint32_t st = 0; st < Q->meta[s].n; st += kWideBlock) map.push_back(wide_block{s, st});
    }
    if (map.size() > Q->wide_map_cap) return fail(HG_ERR_INVALID, "hg_pool_step: wide map overflow");
    Q->narrow_count = uint32_t(Q->order.size() - first_narrow);
    Q->wide_blocks = unsigned(map.size());
    // The queue, largest first: each system of at least a workgroup's
    // threads alone, then (neighbour-list mode) the smaller ones packed best
    // fit decreasing: each, in descending size, joins the fullest pack it
    // fits (one particle per thread, at most kMaxPack members), or opens one.
    const bool pack = Q->packing && Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST;
    const uint32_t tpb = uint32_t(Q->chunk_tpb);
    std::vector<uint32_t> slots, packs;
    slots.reserve(Q->narrow_count);
    size_t k = first_narrow;
    for (; k < Q->order.size() && (!pack || Q->meta[Q->order[k]].n >= tpb); k++) {
        packs.push_back(uint32_t(slots.size()) | (1u << 24));
        slots.push_back(Q->order[k]);
    }
    if (k < Q->order.size()) {
        std::vector<std::vector<uint32_t>> bins;
        std::vector<uint32_t> fill;
        std::vector<std::vector<uint32_t>> open(tpb + 1);  // bins with room for a member, by fill
        for (; k < Q->order.size(); k++) {
            const uint32_t s = Q->order[k], n = Q->meta[s].n;
            uint32_t b = ~0u;
            for (uint32_t f = tpb - n; f >= 1 && b == ~0u; f--)
                if (!open[f].empty()) {
                    b = open[f].back();
                    open[f].pop_back();
                }
            if (b == ~0u) {
                b = uint32_t(bins.size());
                bins.emplace_back();
                fill.push_back(0);
            }
            bins[b].push_back(s);
            fill[b] += n;
            if (bins[b].size() < kMaxPack) open[fill[b]].push_back(b);
        }
        std::vector<uint32_t> idx(bins.size());
        for (uint32_t b = 0; b < idx.size(); b++) idx[b] = b;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return fill[a] > fill[b]; });
        for (uint32_t b : idx) {
            packs.push_back(uint32_t(slots.size()) | (uint32_t(bins[b].size()) << 24));
            slots.insert(slots.end(), bins[b].begin(), bins[b].end());
        }
    }
    Q->pack_count = uint32_t(packs.size());
    if (Q->narrow_count) {
        HG_HIP(hipMemcpyAsync(Q->order_d, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, Q->stream));
        HG_HIP(hipMemcpyAsync(Q->packs_d, packs.data(), packs.size() * 4, hipMemcpyHostToDevice, Q->stream));
    }
    if (!map.empty())
        HG_HIP(hipMemcpyAsync(Q->wide_map_d, map.data(), map.size() * sizeof(wide_block), hipMemcpyHostToDevice,
                              Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    Q->order_dirty = false;
    return HG_OK;
}

// Online tuner, in two stages: the chunk kernel's shape (256 threads, two
// workgroups per CU; 512 threads, one: the same waves per SIMD with twice the
// LDS per system) at the current launch split, then the launch split (wide
// threshold and the CUs left for the wide launches) at the better shape.
// Every candidate runs two intervals (the first may record a graph); the
// second one's host-measured time per step is its score. Neither choice
// changes any result, so tuning happens on live steps.
void tune_splits(hg_pool *Q) {
    Q->tune_cands.clear();
    size_t largest = 0;
    for (size_t s = 0; s < Q->S; s++)
        if (Q->active[s]) largest = std::max<size_t>(largest, Q->meta[s].n);
    Q->tune_cands.push_back({0u, Q->wide_cus, Q->chunk_tpb, Q->wg_per_cu});
    for (uint32_t wm : {1024u, 2048u, 4096u, 8192u}) {
        if (wm > largest) break;  // no system would go wide
        Q->tune_cands.push_back({wm, 0, Q->chunk_tpb, Q->wg_per_cu});  // one after the other
        for (int frac : {32, 16, 8, 4})
            Q->tune_cands.push_back({wm, std::max(1, Q->n_cus / frac), Q->chunk_tpb, Q->wg_per_cu});
    }
    Q->tune_stage = 1;
    Q->tune_scores.assign(Q->tune_cands.size(), 0.0);
    Q->tune_pos = 0;
    Q->tune_phase = Q->tune_cands.size() > 1 ? TUNE_RUNNING : TUNE_DONE;
    if (Q->tune_phase == TUNE_DONE) {
        Q->wide_min = 0;
        Q->order_dirty = true;
    }
}

void tune_begin(hg_pool *Q) {
    const uint32_t wm = uint32_t(Q->wide_min);
    Q->tune_cands = {{wm, Q->wide_cus, 256, 2}, {wm, Q->wide_cus, 512, 1}};
    Q->tune_stage = 0;
    Q->tune_scores.assign(Q->tune_cands.size(), 0.0);
    Q->tune_pos = 0;
    Q->tune_phase = TUNE_RUNNING;
}

void tune_apply(hg_pool *Q, const hg_pool::tune_cand &c) {
    if (Q->wide_min != c.wide_min) {
        Q->wide_min = c.wide_min;
        Q->order_dirty = true;
    }
    Q->wide_cus = c.wide_cus;
    if (Q->chunk_tpb != c.tpb) Q->order_dirty = true;  // the packs follow the threads per workgroup
    Q->chunk_tpb = c.tpb;
    set_chunk_wgs(Q, c.wgs);
}

void tune_select(hg_pool *Q) { tune_apply(Q, Q->tune_cands[Q->tune_pos / 2]); }

void tune_record(hg_pool *Q, double seconds_per_step) {
    if (Q->tune_pos % 2 == 1) Q->tune_scores[Q->tune_pos / 2] = seconds_per_step;
    if (++Q->tune_pos < 2 * Q->tune_cands.size()) return;
    size_t best = 0;
    for (size_t k = 1; k < Q->tune_cands.size(); k++)
        if (Q->tune_scores[k] < Q->tune_scores[best]) best = k;
    tune_apply(Q, Q->tune_cands[best]);
    Q->order_dirty = true;
    if (Q->tune_stage == 0)
        tune_splits(Q);
    else
        Q->tune_phase = TUNE_DONE;
}
}  // namespace

}  // extern "C"

namespace hg {

hg_status pool_begin(hg_pool *Q, double dt, uint64_t max_len, uint32_t rebuild_every, uint32_t *len_out) {
    *len_out = 0;
    if (Q->fl.active) return fail(HG_ERR_INVALID, "pool: an interval is already queued");
    if (Q->tune_phase == TUNE_START) tune_begin(Q);
    if (Q->tune_phase == TUNE_RUNNING) tune_select(Q);
    if (Q->order_dirty) PG_TRY(refresh_order(Q));
    if (Q->order.empty() || max_len == 0) return HG_OK;
    const bool nl = Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST;
    const bool cadence = nl && rebuild_every;
    const pool_view V = make_pool_view(Q);
    const double half = dt / 2.0;
    // one launch runs every system up to the nearest step at which any of
    // them is due for a rebuild, as hg_step does per system
    uint64_t len64 = std::min<uint64_t>(cadence ? rebuild_every : 2000, max_len);
    if (cadence)
        for (uint32_t s : Q->order) len64 = std::min<uint64_t>(len64, rebuild_every - Q->step_index[s] % rebuild_every);
    const uint32_t len = uint32_t(len64);
    const uint32_t count = Q->narrow_count;
    const unsigned grid = unsigned(std::min<uint64_t>(Q->pack_count, uint64_t(Q->grid)));
    bool any_wide_initial = false;
    for (size_t k = 0; k < Q->order.size() - count; k++) any_wide_initial |= !Q->meta[Q->order[k]].acc_valid;
    const auto t_interval = std::chrono::steady_clock::now();
    // The interval's device work is queued without waiting in between: the
    // meta (with the rebuild flags of the systems due) goes up from pinned
    // memory, the rebuild follows on the pool's stream, and the wide launches
    // on the second stream wait for it by an event.
    HG_HIP(hipEventRecord(Q->evb[0], Q->stream));
    bool rebuilt = false;
    if (nl)
        for (uint32_t s : Q->order) {
            const bool due = Q->meta[s].rebuild || (cadence && Q->step_index[s] % rebuild_every == 0);
            Q->meta[s].rebuild = due ? 1u : 0u;
            rebuilt |= due;
        }
    std::memcpy(Q->meta_pinned, Q->meta.data(), Q->S * sizeof(pool_meta));
    HG_HIP(hipMemcpyAsync(Q->meta_d, Q->meta_pinned, Q->S * sizeof(pool_meta), hipMemcpyHostToDevice, Q->stream));
    Q->meta_dirty = false;
    if (rebuilt) {
        HG_HIP(hipEventRecord(Q->ev[0], Q->stream));
        PG_TRY(rebuild_due(Q));
        for (uint32_t s : Q->order) Q->meta[s].rebuild = 0;
    }
    HG_HIP(hipEventRecord(Q->ev[1], Q->stream));
    // wide systems: per-step launches on their own CUs, beside the chunk
    // kernel on the others (all CUs go to whichever kind is alone)
    // wide_cus = 0: no split; the wide steps are queued on the pool's stream
    // ahead of the chunk kernel and each kind has the whole GPU in turn
    const bool split = Q->wide_blocks && count && Q->wide_cus > 0;
    hipStream_t sc = Q->stream;
    if (Q->wide_blocks) {
        hipStream_t s2 = split ? Q->stream_w : Q->stream;
        if (split) HG_HIP(hipStreamWaitEvent(Q->stream_w, Q->ev[1], 0));
        HG_HIP(hipEventRecord(Q->wev[0], s2));
        auto force = [&](int pass) {
            return nl ? launch_wide_force_p<HG_PAIRS_NEIGHBOR_LIST>(Q->p, V, Q->wide_map_d, Q->wide_blocks, pass, s2)
                      : launch_wide_force_p<HG_PAIRS_ALL>(Q->p, V, Q->wide_map_d, Q->wide_blocks, pass, s2);
        };
        if (any_wide_initial) {
            HG_HIP(force(-1));
            Q->last_wide_launches++;
        }
        // Without wide facets the steps are fused: the first kick/drift, then
        // one launch per step doing its pass and the next step's kick/drift,
        // x, v and omega alternating between the pool's buffers (A) and the
        // second set (B), the last pass alone and a final kick that leaves
        // the state in A: len + 2 launches instead of 2 len + 1.
        bool wide_facets = false;
        for (size_t k = 0; k < Q->order.size() - count; k++) wide_facets |= Q->meta[Q->order[k]].nf > 0;
        const bool fused = !wide_facets;
        pool_view VB = V;
        VB.x = Q->xb;
        VB.v = Q->vb;
        VB.w = Q->wb;
        auto enqueue_fused = [&](hipStream_t st_) -> hipError_t {
            k_pool_wide_kick_drift<<<Q->wide_blocks, kWideBlock, 0, st_>>>(V, Q->wide_map_d, 0, half, dt);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            for (uint32_t st = 0; st + 1 < len; st++) {
                const pool_view &Vi = st % 2 ? VB : V, &Vo = st % 2 ? V : VB;
                e = nl ? launch_wide_force_p<HG_PAIRS_NEIGHBOR_LIST>(Q->p, Vi, Q->wide_map_d, Q->wide_blocks,
                                                                     int(st), st_, &Vo, half, dt)
                       : launch_wide_force_p<HG_PAIRS_ALL>(Q->p, Vi, Q->wide_map_d, Q->wide_blocks, int(st), st_,
                                                           &Vo, half, dt);
                if (e != hipSuccess) return e;
            }
            const pool_view &Vl = (len - 1) % 2 ? VB : V;
            e = nl ? launch_wide_force_p<HG_PAIRS_NEIGHBOR_LIST>(Q->p, Vl, Q->wide_map_d, Q->wide_blocks,
                                                                 int(len - 1), st_)
                   : launch_wide_force_p<HG_PAIRS_ALL>(Q->p, Vl, Q->wide_map_d, Q->wide_blocks, int(len - 1), st_);
            if (e != hipSuccess) return e;
            k_pool_wide_final_kick_io<<<Q->wide_blocks, kWideBlock, 0, st_>>>(Vl, V, Q->wide_map_d, half);
            return hipGetLastError();
        };
        // the interval's launches: recorded once into a HIP graph per
        // (interval length, dt, wide set, fused or not) and replayed;
        // launched one by one when graphs are off
        auto enqueue = [&](hipStream_t st_) -> hipError_t {
            if (fused) return enqueue_fused(st_);
            for (uint32_t st = 0; st < len; st++) {
                k_pool_wide_kick_drift<<<Q->wide_blocks, kWideBlock, 0, st_>>>(V, Q->wide_map_d, st > 0 ? 1 : 0,
                                                                                half, dt);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
                e = nl ? launch_wide_force_p<HG_PAIRS_NEIGHBOR_LIST>(Q->p, V, Q->wide_map_d, Q->wide_blocks,
                                                                     int(st), st_)
                       : launch_wide_force_p<HG_PAIRS_ALL>(Q->p, V, Q->wide_map_d, Q->wide_blocks, int(st), st_);
                if (e != hipSuccess) return e;
            }
            k_pool_wide_final_kick<<<Q->wide_blocks, kWideBlock, 0, st_>>>(V, Q->wide_map_d, half, dt);
            return hipGetLastError();
        };
        if (Q->wide_graph) {
            if (!Q->wg_exec || Q->wg_len != len || Q->wg_dt != dt || Q->wg_blocks != Q->wide_blocks ||
                Q->wg_fused != fused) {
                if (Q->wg_exec) HG_HIP(hipGraphExecDestroy(Q->wg_exec));
                Q->wg_exec = nullptr;
                hipGraph_t graph = nullptr;
                HG_HIP(hipStreamBeginCapture(Q->stream_w, hipStreamCaptureModeThreadLocal));
                const hipError_t ec = enqueue(Q->stream_w);
                const hipError_t ee = hipStreamEndCapture(Q->stream_w, &graph);
                HG_HIP(ec);
                HG_HIP(ee);
                const hipError_t ei = hipGraphInstantiate(&Q->wg_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HG_HIP(ei);
                Q->wg_len = len;
                Q->wg_fused = fused;
                Q->wg_dt = dt;
                Q->wg_blocks = Q->wide_blocks;
            }
            HG_HIP(hipGraphLaunch(Q->wg_exec, s2));
        } else {
            HG_HIP(enqueue(s2));
        }
        Q->last_wide_launches += fused ? len + 2 : 2 * len + 1;
        HG_HIP(hipEventRecord(Q->wev[1], s2));
    }
    if (count) {
        const unsigned g = split ? std::min(grid, unsigned(Q->wg_per_cu * (Q->n_cus - Q->wide_cus))) : grid;
        HG_HIP(launch_zero_u32(Q->counter_d, 1, sc));
        HG_HIP(hipEventRecord(Q->ev[2], sc));
        bool facets = false;
        for (size_t k = Q->order.size() - count; k < Q->order.size(); k++) facets |= Q->meta[Q->order[k]].nf > 0;
        hipError_t e = nl ? launch_chunk_p<HG_PAIRS_NEIGHBOR_LIST>(Q->p, V, Q->order_d, Q->packs_d, Q->pack_count,
                                                                   Q->counter_d, len, dt, g, Q->lds_cap, facets,
                                                                   Q->chunk_tpb, sc)
                          : launch_chunk_p<HG_PAIRS_ALL>(Q->p, V, Q->order_d, Q->packs_d, Q->pack_count, Q->counter_d,
                                                         len, dt, g, Q->lds_cap, facets, Q->chunk_tpb, sc);
        HG_HIP(e);
        HG_HIP(hipEventRecord(Q->ev[3], sc));
    }
    if (split) HG_HIP(hipStreamWaitEvent(Q->stream, Q->wev[1], 0));
    HG_HIP(hipEventRecord(Q->evb[1], Q->stream));
    Q->fl.active = true;
    Q->fl.len = len;
    Q->fl.count = count;
    Q->fl.rebuilt = rebuilt;
    Q->fl.wide = Q->wide_blocks != 0;
    Q->fl.dt = dt;
    Q->fl.t0 = t_interval;
    *len_out = len;
    return HG_OK;
}

hg_status pool_end(hg_pool *Q, std::vector<int32_t> &faulted) {
    if (!Q->fl.active) return HG_OK;
    Q->fl.active = false;
    const uint32_t len = Q->fl.len;
    const double dt = Q->fl.dt;
    HG_HIP(hipStreamSynchronize(Q->stream));
    // host mirror of what the kernels did to each system, before any fault is
    // read, so that the systems that go on keep a true mirror
    for (uint32_t s : Q->order) {
        pool_meta &M = Q->meta[s];
        if (Q->cap_s) {
            if (!M.acc_valid) M.cur ^= 1u;
            M.cur ^= (len & 1u);
        }
        M.acc_valid = 1;
        Q->step_index[s] += len;
        for (uint32_t k = 0; k < len; k++) Q->t[s] = Q->t[s] + dt;
    }
    Q->meta_dirty = true;
    PG_TRY(collect_faults(Q, faulted));
    float ms = 0.f;
    if (Q->fl.rebuilt) {
        HG_HIP(hipEventElapsedTime(&ms, Q->ev[0], Q->ev[1]));
        Q->last_rebuild_ms += ms;
    }
    HG_HIP(hipEventElapsedTime(&ms, Q->evb[0], Q->evb[1]));
    Q->last_busy_ms += ms;
    if (Q->fl.count) {
        HG_HIP(hipEventElapsedTime(&ms, Q->ev[2], Q->ev[3]));
        Q->last_chunk_ms += ms;
        Q->last_chunks++;
        Q->last_chunk_steps += len;
    }
    if (Q->fl.wide) {
        HG_HIP(hipEventElapsedTime(&ms, Q->wev[0], Q->wev[1]));
        Q->last_wide_ms += ms;
    }
    if (Q->tune_phase == TUNE_RUNNING)
        tune_record(Q, std::chrono::duration<double>(std::chrono::steady_clock::now() - Q->fl.t0).count() / len);
    return HG_OK;
}

hg_status pool_fault_status(hg_pool *Q, const std::vector<int32_t> &faulted, const char *where) {
    if (faulted.empty()) return HG_OK;
    bool numeric = false;
    std::string msg = std::string(where) + ": stopped system";
    msg += faulted.size() > 1 ? "s" : "";
    for (int32_t s : faulted) {
        const int f = Q->fault[s];
        numeric |= (f & HG_FAULT_NUMERIC) != 0;
        msg += " " + std::to_string(s) + " (";
        std::string why;
        if (f & HG_FAULT_NEIGHBOR_CAPACITY)
            why += "neighbour list needs " + std::to_string(Q->fault_nl_need) + " slots per particle, capacity is " +
                   std::to_string(Q->cap_n);
        if (f & HG_FAULT_PAIR_CAPACITY)
            why += std::string(why.empty() ? "" : "; ") + "pair-state table needs " +
                   std::to_string(Q->fault_pair_need) + " slots per particle, capacity is " + std::to_string(Q->cap_s);
        if (f & HG_FAULT_NUMERIC) why += std::string(why.empty() ? "" : "; ") + "velocity became NaN or infinite";
        msg += why + ")";
    }
    return fail(numeric ? HG_ERR_NUMERIC : HG_ERR_CAPACITY, msg + "; the other systems completed the steps");
}

void pool_reset_profile(hg_pool *Q) {
    Q->last_chunk_ms = Q->last_rebuild_ms = Q->last_wide_ms = Q->last_busy_ms = 0.0;
    Q->last_chunks = Q->last_wide_launches = Q->last_chunk_steps = 0;
}

void pool_park(hg_pool *Q, int32_t slot) {
    Q->parked[slot] = 1;
    Q->order_dirty = true;
}

}  // namespace hg

extern "C" {

hg_status hg_pool_step(hg_pool *Q, double dt, uint64_t nsteps, uint32_t rebuild_every) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    pool_reset_profile(Q);
    if (nsteps == 0) return HG_OK;
    std::vector<int32_t> faulted;  // systems that stopped during this call
    for (uint64_t done = 0; done < nsteps;) {
        uint32_t len = 0;
        PG_TRY(pool_begin(Q, dt, nsteps - done, rebuild_every, &len));
        if (len == 0) break;
        PG_TRY(pool_end(Q, faulted));
        done += len;
    }
    PG_TRY(upload_meta(Q));
    return pool_fault_status(Q, faulted, "hg_pool_step");
}

hg_status hg_pool_fault(hg_pool *Q, int32_t slot, int *fault) {
    PG_TRY(valid_slot(Q, slot));
    if (!fault) return fail(HG_ERR_INVALID, "hg_pool_fault: fault is required");
    *fault = Q->fault[slot];
    return HG_OK;
}

hg_status hg_pool_set_wide_threshold(hg_pool *Q, size_t wide_min) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (wide_min > 0xffffffffu) wide_min = 0;
    Q->wide_min = uint32_t(wide_min);
    Q->order_dirty = true;
    return HG_OK;
}

hg_status hg_pool_autotune(hg_pool *Q, int enable) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    Q->tune_phase = enable ? TUNE_START : TUNE_OFF;
    return HG_OK;
}

hg_status hg_pool_get_tuning(hg_pool *Q, size_t *wide_min, int *wide_cus, int *tuned) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (wide_min) *wide_min = Q->wide_min;
    if (wide_cus) *wide_cus = Q->wide_cus;
    if (tuned) *tuned = Q->tune_phase == TUNE_DONE ? 1 : 0;
    return HG_OK;
}

hg_status hg_pool_get_chunk_shape(hg_pool *Q, int *threads, int *wg_per_cu) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (threads) *threads = Q->chunk_tpb;
    if (wg_per_cu) *wg_per_cu = Q->wg_per_cu;
    return HG_OK;
}

hg_status hg_pool_set_wide_graph(hg_pool *Q, int enable) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    Q->wide_graph = enable != 0;
    return HG_OK;
}

hg_status hg_pool_set_packing(hg_pool *Q, int enable) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (Q->packing != (enable != 0)) Q->order_dirty = true;
    Q->packing = enable != 0;
    return HG_OK;
}

hg_status hg_pool_set_chunk_shape(hg_pool *Q, int threads, int wg_per_cu) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (threads != 256 && threads != 512) return fail(HG_ERR_INVALID, "hg_pool_set_chunk_shape: threads 256 or 512");
    if (wg_per_cu < 1 || wg_per_cu > 2048 / threads)
        return fail(HG_ERR_INVALID, "hg_pool_set_chunk_shape: wg_per_cu 1.." + std::to_string(2048 / threads));
    if (Q->chunk_tpb != threads) Q->order_dirty = true;
    Q->chunk_tpb = threads;
    set_chunk_wgs(Q, wg_per_cu);
    return HG_OK;
}

hg_status hg_pool_set_wide_cus(hg_pool *Q, int wide_cus) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (wide_cus < 0 || wide_cus >= Q->n_cus)
        return fail(HG_ERR_INVALID, "hg_pool_set_wide_cus: need 0 <= wide_cus < " + std::to_string(Q->n_cus));
    Q->wide_cus = wide_cus;
    return HG_OK;
}

hg_status hg_pool_last_busy(hg_pool *Q, double *busy_ms) {
    if (!Q || !busy_ms) return fail(HG_ERR_INVALID, "pool and busy_ms are required");
    *busy_ms = Q->last_busy_ms;
    return HG_OK;
}

hg_status hg_pool_last_wide_profile(hg_pool *Q, double *wide_ms, uint64_t *launches) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (wide_ms) *wide_ms = Q->last_wide_ms;
    if (launches) *launches = Q->last_wide_launches;
    return HG_OK;
}

hg_status hg_pool_get_state(hg_pool *Q, int32_t slot, double *x, double *v, double *theta, double *omega) {
    PG_TRY(valid_slot(Q, slot));
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    std::vector<double> buf(4 * size_t(n));
    const double *src[4] = {Q->x, Q->v, Q->th, Q->w};
    double *dst[4] = {x, v, theta, omega};
    for (int f = 0; f < 4; f++) {
        if (!dst[f]) continue;
        HG_HIP(hipMemcpyAsync(buf.data(), src[f] + 4 * size_t(off), buf.size() * 8, hipMemcpyDeviceToHost, Q->stream));
        HG_HIP(hipStreamSynchronize(Q->stream));
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) dst[f][3 * i + c] = buf[4 * i + c];
    }
    return HG_OK;
}

hg_status hg_pool_energy(hg_pool *Q, int32_t slot, double mass, double inertia, double *out) {
    PG_TRY(valid_slot(Q, slot));
    if (!out) return fail(HG_ERR_INVALID, "hg_pool_energy: out is required");
    const pool_meta &M = Q->meta[slot];
    const bool nl = Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST;
    if (nl && M.rebuild)
        return fail(HG_ERR_INVALID, "hg_pool_energy: the system is listed by its first step; step it first");
    if (M.n > Q->energy_cap) {
        if (Q->energy) {
            auto it = std::find(Q->allocations.begin(), Q->allocations.end(), static_cast<void *>(Q->energy));
            if (it != Q->allocations.end()) Q->allocations.erase(it);
            HG_HIP(hipFree(Q->energy));
            Q->energy = nullptr;
            Q->energy_cap = 0;
        }
        PG_TRY(palloc(Q, &Q->energy, size_t(8) * M.n));
        Q->energy_cap = M.n;
    }
    double *terms = Q->energy;
    const energy_view E = pool_energy_view(Q, M);
    std::vector<double> h(size_t(8) * M.n);
    hipError_t e = launch_energy(Q->p, E, mass, inertia, terms, Q->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), terms, h.size() * sizeof(double), hipMemcpyDeviceToHost, Q->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(Q->stream);
    HG_HIP(e);
    sum_energy(h.data(), M.n, out);
    return HG_OK;
}

hg_status hg_pool_energy_noalloc(hg_pool *Q, int32_t slot, double mass, double inertia, double *out) {
    PG_TRY(valid_slot(Q, slot));
    if (!out) return fail(HG_ERR_INVALID, "hg_pool_energy_noalloc: out is required");
    const pool_meta &M = Q->meta[slot];
    if (M.n > kNoAllocParticles)
        return fail(HG_ERR_CAPACITY, "hg_pool_energy_noalloc: " + std::to_string(M.n) + " particles exceed its " +
                                         std::to_string(kNoAllocParticles) + "; use hg_pool_energy");
    if (Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST && M.rebuild)
        return fail(HG_ERR_INVALID, "hg_pool_energy_noalloc: the system is listed by its first step; step it first");
    double *scr = nullptr;
    PG_TRY(noalloc_scratch(Q->device, &scr));
    double *sums = scr + 8 * kNoAllocParticles;
    hipError_t e = launch_energy(Q->p, pool_energy_view(Q, M), mass, inertia, scr, Q->stream);
    if (e == hipSuccess) {
        k_energy_sum<<<1, 64, 0, Q->stream>>>(scr, M.n, sums);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, sums, 8 * sizeof(double), hipMemcpyDeviceToHost, Q->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(Q->stream);
    HG_HIP(e);
    return HG_OK;
}

hg_status hg_pool_get_accelerations(hg_pool *Q, int32_t slot, double *a, double *alpha) {
    PG_TRY(valid_slot(Q, slot));
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    std::vector<double> buf(4 * size_t(n));
    const double *src[2] = {Q->a, Q->al};
    double *dst[2] = {a, alpha};
    for (int f = 0; f < 2; f++) {
        if (!dst[f]) continue;
        HG_HIP(hipMemcpyAsync(buf.data(), src[f] + 4 * size_t(off), buf.size() * 8, hipMemcpyDeviceToHost, Q->stream));
        HG_HIP(hipStreamSynchronize(Q->stream));
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) dst[f][3 * i + c] = buf[4 * i + c];
    }
    return HG_OK;
}

hg_status hg_pool_get_neighbors(hg_pool *Q, int32_t slot, uint64_t *offsets, uint32_t *idx, size_t cap,
                                size_t *total) {
    PG_TRY(valid_slot(Q, slot));
    if (!offsets) return fail(HG_ERR_INVALID, "offsets is required");
    if (!Q->cap_n) return fail(HG_ERR_INVALID, "pool has no neighbour lists");
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n;
    std::vector<uint32_t> nb(size_t(Q->cap_n) * n), cnt(n);
    PG_TRY(download_rows(Q, Q->nb, off, n, Q->cap_n, nb.data()));
    PG_TRY(download_rows(Q, Q->ncnt, off, n, 1, cnt.data()));
    size_t tot = 0;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = tot;
        for (uint32_t k = 0; k < cnt[i]; k++) {
            if (idx && tot < cap) idx[tot] = nb[size_t(k) * n + i] - off;
            tot++;
        }
    }
    offsets[n] = tot;
    if (total) *total = tot;
    return HG_OK;
}

hg_status hg_pool_get_springs(hg_pool *Q, int32_t slot, int which, uint32_t *ij, double *springs, size_t cap,
                              size_t *total) {
    PG_TRY(valid_slot(Q, slot));
    if (total) *total = 0;
    if (!Q->cap_s) return HG_OK;
    const uint32_t off = Q->meta[slot].offset, n = Q->meta[slot].n, cur = Q->meta[slot].cur;
    std::vector<uint32_t> pj(size_t(Q->cap_s) * n), cnt(n);
    std::vector<double> ps(size_t(Q->cap_s) * kSpringRec * n);
    PG_TRY(download_rows(Q, Q->pj[cur], off, n, Q->cap_s, pj.data()));
    PG_TRY(download_rows(Q, Q->ps[cur], off, n, Q->cap_s, ps.data(), kSpringRec));
    PG_TRY(download_rows(Q, Q->pcnt[cur], off, n, 1, cnt.data()));
    size_t tot = 0;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t s = 0; s < cnt[i]; s++) {
            const uint32_t tag = pj[size_t(s) * n + i];
            if (((tag & kBondBit) != 0) != (which == 1)) continue;
            if (tot < cap) {
                if (ij) {
                    ij[2 * tot] = i;
                    ij[2 * tot + 1] = (tag & kIdxMask) - off;
                }
                if (springs)
                    for (int c = 0; c < 9; c++) springs[9 * tot + c] = ps[(size_t(s) * n + i) * kSpringRec + c];
            }
            tot++;
        }
    if (total) *total = tot;
    return HG_OK;
}

hg_status hg_pool_get_bonds(hg_pool *Q, int32_t slot, uint32_t *pairs, size_t cap, size_t *nbonds,
                            size_t *nskipped) {
    PG_TRY(valid_slot(Q, slot));
    const auto &B = Q->bonds[slot];
    for (size_t b = 0; pairs && b < B.size() && b < cap; b++) {
        pairs[2 * b] = B[b].first;
        pairs[2 * b + 1] = B[b].second;
    }
    if (nbonds) *nbonds = B.size();
    if (nskipped) *nskipped = Q->skipped[slot];
    return HG_OK;
}

hg_status hg_pool_info(hg_pool *Q, int32_t slot, size_t *n, double *t, uint64_t *step_index) {
    PG_TRY(valid_slot(Q, slot));
    if (n) *n = Q->meta[slot].n;
    if (t) *t = Q->t[slot];
    if (step_index) *step_index = Q->step_index[slot];
    return HG_OK;
}

hg_status hg_pool_last_step_profile(hg_pool *Q, double *chunk_ms, double *rebuild_ms, uint64_t *chunks,
                                    uint64_t *chunk_steps) {
    if (!Q) return fail(HG_ERR_INVALID, "pool is null");
    if (chunk_ms) *chunk_ms = Q->last_chunk_ms;
    if (rebuild_ms) *rebuild_ms = Q->last_rebuild_ms;
    if (chunks) *chunks = Q->last_chunks;
    if (chunk_steps) *chunk_steps = Q->last_chunk_steps;
    return HG_OK;
}

}  // extern "C"

namespace {
template <typename T>
hg_status upload_rows(hg_pool *Q, T *dev, uint32_t offset, uint32_t n, uint32_t rows, const std::vector<T> &host,
                      size_t width = 1) {
    if (rows == 0 || n == 0) return HG_OK;
    const size_t row = n * width * sizeof(T);
    HG_HIP(hipMemcpy2DAsync(dev + offset * width, Q->P * width * sizeof(T), host.data(), row, row, rows,
                            hipMemcpyHostToDevice, Q->stream));
    HG_HIP(hipStreamSynchronize(Q->stream));
    return HG_OK;
}
}  // namespace

extern "C" {

hg_status hg_pool_save(hg_pool *Q, int32_t slot, void *buf, size_t cap, size_t *size) {
    PG_TRY(valid_slot(Q, slot));
    if (!size) return fail(HG_ERR_INVALID, "hg_pool_save: size is required");
    const pool_meta &M = Q->meta[slot];
    const uint32_t off = M.offset, n = M.n, cur = M.cur;
    snapshot z;
    z.p = Q->p;
    z.n = n;
    z.t = Q->t[slot];
    z.step_index = Q->step_index[slot];
    z.acc_valid = M.acc_valid;
    const bool nl = Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST;
    z.needs_rebuild = nl && M.rebuild ? 1u : 0u;
    const double *fields[6] = {Q->x, Q->v, Q->th, Q->w, Q->a, Q->al};
    std::vector<double> *dst[6] = {&z.x, &z.v, &z.th, &z.w, &z.a, &z.al};
    for (int f = 0; f < 6; f++) {
        std::vector<double> q4(4 * size_t(n));
        PG_TRY(download_rows(Q, fields[f] + 4 * size_t(off), 0, n, 1, q4.data(), 4));
        *dst[f] = pack3(q4, n);
    }
    if (nl && !M.rebuild) {
        std::vector<uint32_t> nb(size_t(Q->cap_n) * n);
        z.has_nl = true;
        z.nl_count.resize(n);
        PG_TRY(download_rows(Q, Q->nb, off, n, Q->cap_n, nb.data()));
        PG_TRY(download_rows(Q, Q->ncnt, off, n, 1, z.nl_count.data()));
        rows_to_csr(nb.data(), z.nl_count.data(), n, 1, 1, z.nl_ids);
        shift_ids(z.nl_ids, off);
    }
    if (Q->cap_s) {
        std::vector<uint32_t> pj(size_t(Q->cap_s) * n);
        std::vector<double> ps(size_t(Q->cap_s) * kSpringRec * n);
        z.has_table = true;
        z.pt_count.resize(n);
        PG_TRY(download_rows(Q, Q->pj[cur], off, n, Q->cap_s, pj.data()));
        PG_TRY(download_rows(Q, Q->ps[cur], off, n, Q->cap_s, ps.data(), kSpringRec));
        PG_TRY(download_rows(Q, Q->pcnt[cur], off, n, 1, z.pt_count.data()));
        rows_to_csr(pj.data(), z.pt_count.data(), n, 1, 1, z.pt_tags);
        rows_to_csr(ps.data(), z.pt_count.data(), n, kSpringRec, 9, z.pt_springs);
        shift_ids(z.pt_tags, off);
    }
    for (auto &b : Q->bonds[slot]) {
        z.bonds.push_back(b.first);
        z.bonds.push_back(b.second);
    }
    z.skipped = Q->skipped[slot];
    z.nf = M.nf;
    if (M.nf > 0) {
        z.facets.resize(M.nf);
        z.ss.resize(size_t(M.nf) * 9 * n);
        HG_HIP(hipMemcpyAsync(z.facets.data(), Q->facets + size_t(slot) * Q->MF, M.nf * sizeof(facet_dev),
                              hipMemcpyDeviceToHost, Q->stream));
        HG_HIP(hipStreamSynchronize(Q->stream));
        PG_TRY(download_rows(Q, Q->ss, off, n, M.nf * 9, z.ss.data()));
    }
    if (Q->bstr) {
        std::vector<uint32_t> bj(size_t(kMaxBondsPerParticle) * n);
        std::vector<double> bs(size_t(kMaxBondsPerParticle) * n);
        z.has_neck = true;
        z.neck_count.resize(n);
        PG_TRY(download_rows(Q, Q->bj, off, n, kMaxBondsPerParticle, bj.data()));
        PG_TRY(download_rows(Q, Q->bstr, off, n, kMaxBondsPerParticle, bs.data()));
        PG_TRY(download_rows(Q, Q->bcnt, off, n, 1, z.neck_count.data()));
        rows_to_csr(bj.data(), z.neck_count.data(), n, 1, 1, z.neck_j);
        rows_to_csr(bs.data(), z.neck_count.data(), n, 1, 1, z.neck_str);
        shift_ids(z.neck_j, off);
    }
    const std::vector<uint8_t> bytes = encode_snapshot(z);
    *size = bytes.size();
    if (!buf) return HG_OK;
    if (cap < bytes.size())
        return fail(HG_ERR_CAPACITY, "hg_pool_save: the snapshot needs " + std::to_string(bytes.size()) + " bytes");
    std::memcpy(buf, bytes.data(), bytes.size());
    return HG_OK;
}

hg_status hg_pool_load(hg_pool *Q, const void *buf, size_t size, int32_t *slot_out) {
    if (!Q || !slot_out) return fail(HG_ERR_INVALID, "hg_pool_load: pool and slot are required");
    snapshot z;
    PG_TRY(decode_snapshot(buf, size, z));
    if (std::memcmp(&z.p, &Q->p, sizeof(hg_params)) != 0)
        return fail(HG_ERR_INVALID, "hg_pool_load: the snapshot was taken with other parameters than the pool's");
    if (z.has_legacy) return fail(HG_ERR_INVALID, "hg_pool_load: pools do not run the legacy sinter_functor");
    if (z.nf > Q->MF) return fail(HG_ERR_INVALID, "hg_pool_load: more facets than the pool's max_facets");
    if ((z.has_nl && max_count(z.nl_count) > Q->cap_n) || (z.has_table && max_count(z.pt_count) > Q->cap_s))
        return fail(HG_ERR_CAPACITY, "hg_pool_load: a row of the snapshot exceeds the pool's capacity");
    const uint32_t n = uint32_t(z.n);
    std::vector<std::pair<uint32_t, uint32_t>> bonds(z.bonds.size() / 2);
    for (size_t b = 0; b < bonds.size(); b++) bonds[b] = {z.bonds[2 * b], z.bonds[2 * b + 1]};
    int32_t slot = -1;
    PG_TRY(pool_add_impl(Q, n, z.x.data(), z.v.data(), z.th.data(), z.w.data(), &bonds, z.skipped, &slot));
    const uint32_t off = Q->meta[slot].offset;
    auto fill = [&]() -> hg_status {
        for (auto [dev, host] : {std::pair<double *, const std::vector<double> *>{Q->a, &z.a}, {Q->al, &z.al}}) {
            std::vector<double> q4(4 * size_t(n), 0.0);
            for (uint32_t i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) q4[4 * size_t(i) + c] = (*host)[3 * size_t(i) + c];
            HG_HIP(hipMemcpyAsync(dev + 4 * size_t(off), q4.data(), q4.size() * 8, hipMemcpyHostToDevice, Q->stream));
            HG_HIP(hipStreamSynchronize(Q->stream));
        }
        pool_meta &M = Q->meta[slot];
        if (z.has_table && Q->cap_s) {
            std::vector<uint32_t> tags = z.pt_tags, pj;
            std::vector<double> ps;
            shift_ids(tags, -int64_t(off));
            csr_to_rows(z.pt_count, tags, n, Q->cap_s, 1, 1, kNoEntry, pj);
            csr_to_rows(z.pt_count, z.pt_springs, n, Q->cap_s, kSpringRec, 9, 0.0, ps);
            M.cur = 0;
            PG_TRY(upload_rows(Q, Q->pj[0], off, n, Q->cap_s, pj));
            PG_TRY(upload_rows(Q, Q->ps[0], off, n, Q->cap_s, ps, kSpringRec));
            PG_TRY(upload_rows(Q, Q->pcnt[0], off, n, 1, z.pt_count));
        }
        if (Q->p.pair_mode == HG_PAIRS_NEIGHBOR_LIST && z.has_nl) {
            std::vector<uint32_t> ids = z.nl_ids, nb;
            shift_ids(ids, -int64_t(off));
            csr_to_rows(z.nl_count, ids, n, Q->cap_n, 1, 1, 0u, nb);
            PG_TRY(upload_rows(Q, Q->nb, off, n, Q->cap_n, nb));
            PG_TRY(upload_rows(Q, Q->ncnt, off, n, 1, z.nl_count));
            M.rebuild = 0;
        }
        if (z.nf > 0) {
            HG_HIP(hipMemcpyAsync(Q->facets + size_t(slot) * Q->MF, z.facets.data(), z.nf * sizeof(facet_dev),
                                  hipMemcpyHostToDevice, Q->stream));
            HG_HIP(hipStreamSynchronize(Q->stream));
            PG_TRY(upload_rows(Q, Q->ss, off, n, z.nf * 9, z.ss));
            M.nf = z.nf;
        }
        if (z.has_neck && Q->bcnt) {
            if (max_count(z.neck_count) > kMaxBondsPerParticle) return fail(HG_ERR_INVALID, "hg_pool_load: bad neck rows");
            std::vector<uint32_t> ids = z.neck_j, bj;
            std::vector<double> bs;
            shift_ids(ids, -int64_t(off));
            csr_to_rows(z.neck_count, ids, n, kMaxBondsPerParticle, 1, 1, kNoEntry, bj);
            csr_to_rows(z.neck_count, z.neck_str, n, kMaxBondsPerParticle, 1, 1, 0.0, bs);
            PG_TRY(upload_rows(Q, Q->bj, off, n, kMaxBondsPerParticle, bj));
            PG_TRY(upload_rows(Q, Q->bstr, off, n, kMaxBondsPerParticle, bs));
            PG_TRY(upload_rows(Q, Q->bcnt, off, n, 1, z.neck_count));
        }
        M.acc_valid = z.acc_valid;
        Q->t[slot] = z.t;
        Q->step_index[slot] = z.step_index;
        Q->meta_dirty = true;
        Q->order_dirty = true;
        return upload_meta(Q);
    };
    const hg_status st = fill();
    if (st != HG_OK) {
        const std::string msg = hg_last_error();
        (void)hg_pool_remove(Q, slot);
        return fail(st, msg);
    }
    *slot_out = slot;
    return HG_OK;
}

}  // extern "C"
