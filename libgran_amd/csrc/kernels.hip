// gfx950 kernels of the FP64 DEM step: acceleration pass (pair fold + facets),
// velocity-Verlet kick/drift, cell-list neighbour rebuild and diagnostics.
//
// Data layout in HBM
//   particle fields x, v, theta, omega, a, alpha: one array per field, 4
//     doubles (32 B) per particle, so a particle's vector is one aligned
//     32-byte load and a wavefront's loads are contiguous.
//   neighbour list: transposed fixed-capacity slots, nb[k*n + i], ascending j
//     per particle; consecutive lanes (particles) read consecutive words.
//   pair-state table: per particle a j-sorted list of (tag, spring9) where
//     tag = j | bonded-bit; tags transposed as pj[s*n + i], springs in one
//     contiguous 80-byte record per entry at ps[(s*n + i)*10].
//     It holds every bonded pair and every pair whose contact spring the
//     reference would hold non-zero. Each acceleration pass reads the old
//     table and writes the new one while walking the partners in ascending j,
//     so spring state survives neighbour rebuilds with no separate merge.
#include <hip/hip_runtime.h>

#include "force_device.h"
#include "hg_internal.h"

namespace hg {

namespace {


// Neighbour-list acceleration pass: one lane per particle i, partners in the
// list's ascending order, so the per-particle sum has the reference's order.
// TRACK: the particle's share of its pairs' dissipated work this step is
// added to S.diss (4 doubles per particle: normal, tangential, rolling,
// torsion).
template <int BIN, int HAM, bool TRACK = false>
__device__ __forceinline__ void force_nl_one(const hg_params &P, const sys_view &S, uint32_t i) {
    const d3 xi = ld3(S.x, i), vi = ld3(S.v, i), wi = ld3(S.w, i);
    acc6 acc{zero3(), zero3()};
    walker_t<false> W;
    if constexpr (BIN != HG_BINARY_NONE) W.init(S.pt, i);
    if constexpr (BIN == HG_BINARY_SINTER_LEGACY) W.init_legacy(S.pt);
    id_list ids;
    ids.init(S.nb, S.ncnt, S.n, i, S.cap_n);
    double d4[4] = {0.0, 0.0, 0.0, 0.0};
    nl_partners(ids, [&](uint32_t j) { return pstate{ld3(S.x, j), ld3(S.v, j), ld3(S.w, j)}; },
                [&](uint32_t j, const pstate &sj) {
                    pair_fold_vw<BIN, HAM, TRACK>(P, j, xi, vi, wi, sj.x, [&](d3 &v, d3 &w) { v = sj.v, w = sj.w; },
                                                  W, acc, d4);
                });
    close_table<BIN>(W, S.pt.pcnt_new, i, S.flags);
    if (S.nf > 0) acc = acc + facet_fold(P, S.facets, S.nf, S.ss, S.n, i, xi, vi, wi);
    st3(S.a, i, acc.lin);
    st3(S.al, i, acc.ang);
    if constexpr (TRACK)
        for (int c = 0; c < 4; c++) S.diss[4 * size_t(i) + c] = S.diss[4 * size_t(i) + c] + d4[c];
}

template <int BIN, int HAM, bool TRACK>
__global__ void __launch_bounds__(256) k_force_nl(hg_params P, sys_view S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    force_nl_one<BIN, HAM, TRACK>(P, S, i);
}

// Mixed-precision neighbour-list pass (HG_PRECISION_MIXED, dem_mixed.h).
template <int BIN, int HAM>
__global__ void __launch_bounds__(256) k_force_nl_mx(hg_params P, mixed_consts MC, sys_view S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    const d3 xi = ld3(S.x, i), vi = ld3(S.v, i), wi = ld3(S.w, i);
    accf acc{zerof(), zerof()};
    walker_t<true> W;
    if constexpr (BIN != HG_BINARY_NONE) W.init(S.pt, i);
    id_list ids;
    ids.init(S.nb, S.ncnt, S.n, i, S.cap_n);
    nl_partners(ids, [&](uint32_t j) { return pstate{ld3(S.x, j), ld3(S.v, j), ld3(S.w, j)}; },
                [&](uint32_t j, const pstate &sj) {
                    pair_fold_mx<BIN, HAM>(MC, j, xi, vi, wi, sj.x, [&](d3 &v, d3 &w) { v = sj.v, w = sj.w; }, W,
                                           acc);
                });
    close_table<BIN>(W, S.pt.pcnt_new, i, S.flags);
    if (S.nf > 0) acc = acc + facet_fold_mx(P, MC, S.facets, S.nf, S.ss, S.n, i, xi, vi, wi);
    st3(S.a, i, tod(acc.lin));
    st3(S.al, i, tod(acc.ang));
}

// All-pairs acceleration pass (granular_system): every j != i in ascending
// order. Partner state is staged through LDS one tile per block-wide step;
// all lanes read the same LDS word, a broadcast with no bank conflict.
constexpr int kTile = 256;
struct tile_lds {
    double sx[3][kTile], sv[3][kTile], sw[3][kTile];
};
// The pass for particle i = blockIdx.x * blockDim.x + threadIdx.x; every
// thread of the block takes part (it stages tiles), live or not.
template <int BIN, int HAM, bool TRACK = false>
__device__ __forceinline__ void force_all_block(const hg_params &P, const sys_view &S, tile_lds &T) {
    auto &sx = T.sx;
    auto &sv = T.sv;
    auto &sw = T.sw;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < S.n;
    d3 xi = zero3(), vi = zero3(), wi = zero3();
    if (live) {
        xi = ld3(S.x, i);
        vi = ld3(S.v, i);
        wi = ld3(S.w, i);
    }
    acc6 acc{zero3(), zero3()};
    double d4[4] = {0.0, 0.0, 0.0, 0.0};
    walker_t<false> W;
    if constexpr (BIN != HG_BINARY_NONE) {
        if (live) W.init(S.pt, i);
    }
    if constexpr (BIN == HG_BINARY_SINTER_LEGACY) {
        if (live) W.init_legacy(S.pt);
    }
    for (uint32_t base = 0; base < S.n; base += kTile) {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < kTile && base + t < S.n; t += blockDim.x) {
            d3 q = ld3(S.x, base + t), u = ld3(S.v, base + t), r = ld3(S.w, base + t);
            sx[0][t] = q.x; sx[1][t] = q.y; sx[2][t] = q.z;
            sv[0][t] = u.x; sv[1][t] = u.y; sv[2][t] = u.z;
            sw[0][t] = r.x; sw[1][t] = r.y; sw[2][t] = r.z;
        }
        __syncthreads();
        if (live) {
            const uint32_t m = (S.n - base) < uint32_t(kTile) ? (S.n - base) : uint32_t(kTile);
            for (uint32_t jj = 0; jj < m; jj++) {
                const uint32_t j = base + jj;
                if (j == i) continue;
                pair_fold_vw<BIN, HAM, TRACK>(
                    P, j, xi, vi, wi, d3{sx[0][jj], sx[1][jj], sx[2][jj]},
                    [&](d3 &v, d3 &w) { v = d3{sv[0][jj], sv[1][jj], sv[2][jj]}, w = d3{sw[0][jj], sw[1][jj], sw[2][jj]}; },
                    W, acc, d4);
            }
        }
    }
    if (!live) return;
    close_table<BIN>(W, S.pt.pcnt_new, i, S.flags);
    if (S.nf > 0) acc = acc + facet_fold(P, S.facets, S.nf, S.ss, S.n, i, xi, vi, wi);
    st3(S.a, i, acc.lin);
    st3(S.al, i, acc.ang);
    if constexpr (TRACK)
        for (int c = 0; c < 4; c++) S.diss[4 * size_t(i) + c] = S.diss[4 * size_t(i) + c] + d4[c];
}

template <int BIN, int HAM, bool TRACK>
__global__ void __launch_bounds__(kTile) k_force_all(hg_params P, sys_view S) {
    __shared__ tile_lds T;
    force_all_block<BIN, HAM, TRACK>(P, S, T);
}

// Kick (completing the previous step when `pending`), kick, drift:
//   v += a*(dt/2); w += alpha*(dt/2); x += v*dt; theta += w*dt.
// Lane 0 of block 0 advances the facets when a previous step completes here.
__device__ __forceinline__ void kick_drift_one(double *x, double *v, double *th, double *w, const double *a,
                                               const double *al, uint32_t i, int pending, double half_prev,
                                               double half, double dt) {
    d3 vi = ld3(v, i), wi = ld3(w, i);
    const d3 ai = ld3(a, i), ali = ld3(al, i);
    if (pending) {
        vi = vi + ai * half_prev;
        wi = wi + ali * half_prev;
    }
    vi = vi + ai * half;
    wi = wi + ali * half;
    st3(x, i, ld3(x, i) + vi * dt);
    st3(th, i, ld3(th, i) + wi * dt);
    st3(v, i, vi);
    st3(w, i, wi);
}

__global__ void __launch_bounds__(256) k_kick_drift(double *x, double *v, double *th, double *w, const double *a,
                                                    const double *al, uint32_t n, int pending, double half_prev,
                                                    double half, double dt, facet_dev *facets, int nf,
                                                    double facet_dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (pending && i == 0) advance_facets(facets, nf, facet_dt);
    if (i >= n) return;
    kick_drift_one(x, v, th, w, a, al, i, pending, half_prev, half, dt);
}

// sinter_step_handler (sinter_bridge.h:24-42, 55-75) after a drift: each of
// particle i's connectors moves by dx = v_i dt (increment_x), then turns by
// dtheta = w_i dt about the new centre (increment_theta). v, w and x are the
// values the drift just used and wrote, so dx and dtheta are its increments.
__device__ __forceinline__ void legacy_move_one(const double *x, const double *v, const double *w, uint32_t n,
                                                uint32_t i, double dt, double *lconn, const uint32_t *bcnt) {
    const d3 dx = ld3(v, i) * dt, dth = ld3(w, i) * dt, xi = ld3(x, i);
    const uint32_t c = bcnt[i];
    for (uint32_t s = 0; s < c; s++) {
        double *r = lconn + (size_t(s) * n + i) * kConnRec;
        for (int q = 0; q < 5; q++) {
            d3 p = d3{r[3 * q], r[3 * q + 1], r[3 * q + 2]} + dx;
            const d3 spring_vector = p - xi;
            p = p + cross(dth, spring_vector);
            r[3 * q] = p.x;
            r[3 * q + 1] = p.y;
            r[3 * q + 2] = p.z;
        }
    }
}

__global__ void __launch_bounds__(256) k_legacy_move(const double *x, const double *v, const double *w, uint32_t n,
                                                     double dt, double *lconn, const uint32_t *bcnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    legacy_move_one(x, v, w, n, i, dt, lconn, bcnt);
}

// Closing kick of particle i; a non-finite result raises FLAG_NUMERIC.
__device__ __forceinline__ void final_kick_one(double *v, double *w, const double *a, const double *al, uint32_t i,
                                               double half, int *flags) {
    const d3 vi = ld3(v, i) + ld3(a, i) * half, wi = ld3(w, i) + ld3(al, i) * half;
    st3(v, i, vi);
    st3(w, i, wi);
    if (nonfinite(vi, wi)) atomicOr(&flags[FLAG_NUMERIC], 1);
}

__global__ void __launch_bounds__(256) k_final_kick(double *v, double *w, const double *a, const double *al,
                                                    uint32_t n, double half, facet_dev *facets, int nf,
                                                    double facet_dt, int *flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) advance_facets(facets, nf, facet_dt);
    if (i >= n) return;
    final_kick_one(v, w, a, al, i, half, flags);
}

// nsteps steps of a system of at most kTile particles in one workgroup, one
// particle per thread: per step the kick/drift (with the legacy connector
// moves), a barrier, the acceleration pass, a barrier, as the per-step
// launches of k_kick_drift, k_legacy_move and k_force_nl / k_force_all do it,
// through the same device code. The pass reads pair-state table `even` on
// even steps and `odd` on odd ones (the two halves swap every pass).
// pending_first: the first kick also completes a previous step; final_kick:
// the last step's closing kick is applied (otherwise it is left pending, as
// between the per-step launches).
template <int BIN, int HAM>
__global__ void __launch_bounds__(kTile) k_small_steps(hg_params P, sys_view even, sys_view odd, double *x, double *v,
                                                       double *th, double *w, facet_dev *facets, uint32_t nsteps,
                                                       int pending_first, int final_kick, double half, double dt,
                                                       int nf_advance, double *lconn, const uint32_t *bcnt) {
    __shared__ tile_lds T;
    const uint32_t i = threadIdx.x;
    const uint32_t n = even.n;
    const bool live = i < n;
    for (uint32_t s = 0; s < nsteps; s++) {
        const int pending = (s > 0 || pending_first) ? 1 : 0;
        if (pending && i == 0) advance_facets(facets, nf_advance, dt);
        if (live) {
            kick_drift_one(x, v, th, w, even.a, even.al, i, pending, half, half, dt);
            if (lconn) legacy_move_one(x, v, w, n, i, dt, lconn, bcnt);
        }
        __syncthreads();
        if (P.pair_mode == HG_PAIRS_NEIGHBOR_LIST) {
            if (live) force_nl_one<BIN, HAM>(P, (s & 1u) ? odd : even, i);
        } else {
            force_all_block<BIN, HAM>(P, (s & 1u) ? odd : even, T);
        }
        __syncthreads();
    }
    if (final_kick) {
        if (i == 0) advance_facets(facets, nf_advance, dt);
        if (live) final_kick_one(v, w, even.a, even.al, i, half, even.flags);
    }
}

// triangular_facet::update_positions with the driver's velocity (one lane).
__global__ void k_advance_facet(facet_dev *f, double vx, double vy, double vz, double dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f->v_facet = d3{vx, vy, vz};
        for (int k = 0; k < 3; k++) f->vert[k] = f->vert[k] + f->v_facet * dt;
    }
}

__global__ void k_zero_u32(uint32_t *p, size_t count) {
    for (size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x; k < count; k += size_t(gridDim.x) * blockDim.x)
        p[k] = 0;
}
__global__ void k_zero_f64(double *p, size_t count) {
    for (size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x; k < count; k += size_t(gridDim.x) * blockDim.x)
        p[k] = 0.0;
}

// ---------------------------------------------------- neighbour rebuild -----
// Single-block bounding box of the positions: bbox = lo[3], hi[3].
__global__ void __launch_bounds__(1024) k_bbox(const double *x, uint32_t n, double *bbox) {
    __shared__ double red[6][1024];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        d3 q = ld3(x, i);
        lo[0] = fmin(lo[0], q.x); lo[1] = fmin(lo[1], q.y); lo[2] = fmin(lo[2], q.z);
        hi[0] = fmax(hi[0], q.x); hi[1] = fmax(hi[1], q.y); hi[2] = fmax(hi[2], q.z);
    }
    for (int c = 0; c < 3; c++) {
        red[c][threadIdx.x] = lo[c];
        red[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + s]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) bbox[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t cell_hash(long long cx, long long cy, long long cz, uint32_t mask) {
    uint64_t h = uint64_t(cx) * 73856093ull ^ uint64_t(cy) * 19349663ull ^ uint64_t(cz) * 83492791ull;
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 32;
    return uint32_t(h) & mask;
}

__device__ __forceinline__ void cell_coords(d3 q, const double *bbox, double cs, long long &cx, long long &cy,
                                            long long &cz) {
    cx = (long long)((q.x - bbox[0]) / cs);
    cy = (long long)((q.y - bbox[1]) / cs);
    cz = (long long)((q.z - bbox[2]) / cs);
}

__global__ void __launch_bounds__(256) k_cell_count(const double *x, uint32_t n, const double *bbox, double cs,
                                                    uint32_t mask, uint32_t *cell_of, uint32_t *bucket_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long cx, cy, cz;
    cell_coords(ld3(x, i), bbox, cs, cx, cy, cz);
    const uint32_t h = cell_hash(cx, cy, cz, mask);
    cell_of[i] = h;
    atomicAdd(&bucket_count[h], 1u);
}

// Exclusive scan in three launches: per-block scan of kScanChunk words with
// block totals, a single-block scan of the totals, then the offsets added.
constexpr uint32_t kScanThreads = 256, kScanPer = 8, kScanChunk = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        uint32_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    total = sh[blockDim.x - 1];
    uint32_t r = sh[threadIdx.x] - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_blocks(const uint32_t *in, uint32_t *out, uint32_t count,
                                                               uint32_t *block_sums) {
    __shared__ uint32_t sh[kScanThreads];
    const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
    uint32_t vals[kScanPer];
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanPer; k++) {
        vals[k] = base + k < count ? in[base + k] : 0u;
        s += vals[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, sh, total);
    for (uint32_t k = 0; k < kScanPer; k++) {
        if (base + k < count) out[base + k] = run;
        run += vals[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_top(uint32_t *block_sums, uint32_t nblocks) {
    __shared__ uint32_t sh[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblocks; base += kScanThreads) {
        const uint32_t idx = base + threadIdx.x;
        const uint32_t v = idx < nblocks ? block_sums[idx] : 0u;
        uint32_t total;
        const uint32_t r = block_exclusive_scan(v, sh, total);
        if (idx < nblocks) block_sums[idx] = r + carry;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t *out, uint32_t count, const uint32_t *block_sums) {
    const uint32_t add = block_sums[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanChunk;
    for (uint32_t k = threadIdx.x; k < kScanChunk && base + k < count; k += blockDim.x) out[base + k] += add;
}

__global__ void __launch_bounds__(256) k_cell_fill(uint32_t n, const uint32_t *cell_of, const uint32_t *bucket_start,
                                                   uint32_t *bucket_fill, uint32_t *cell_items) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = cell_of[i];
    cell_items[bucket_start[h] + atomicAdd(&bucket_fill[h], 1u)] = i;
}

// Per particle: the 27 surrounding cells' buckets (deduplicated, since two
// cells may share a bucket), candidates filtered by the reference predicate
// norm(x_i - x_j) < r_verlet, inserted in ascending j.
__global__ void __launch_bounds__(256) k_nl_fill(const double *x, uint32_t n, const double *bbox, double cs,
                                                 uint32_t mask, double r_verlet, const uint32_t *bucket_count,
                                                 const uint32_t *bucket_start, const uint32_t *cell_items,
                                                 uint32_t *nb, uint32_t *ncnt, uint32_t cap, int *flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d3 xi = ld3(x, i);
    long long cx, cy, cz;
    cell_coords(xi, bbox, cs, cx, cy, cz);
    uint32_t seen[27];
    int nseen = 0;
    uint32_t cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const uint32_t h = cell_hash(cx + dx, cy + dy, cz + dz, mask);
                bool dup = false;
                for (int s = 0; s < nseen; s++) dup |= seen[s] == h;
                if (dup) continue;
                seen[nseen++] = h;
                const uint32_t b = bucket_start[h], e = b + bucket_count[h];
                for (uint32_t q = b; q < e; q++) {
                    const uint32_t j = cell_items[q];
                    if (j == i) continue;
                    if (norm(xi - ld3(x, j)) < r_verlet) {
                        if (cnt < cap) {
                            // insertion into the ascending slot list
                            uint32_t pos = cnt;
                            while (pos > 0) {
                                const uint32_t prev = nb[size_t(pos - 1) * n + i];
                                if (prev < j) break;
                                nb[size_t(pos) * n + i] = prev;
                                pos--;
                            }
                            nb[size_t(pos) * n + i] = j;
                        }
                        cnt++;
                    }
                }
            }
    ncnt[i] = cnt < cap ? cnt : cap;
    if (cnt > cap) {
        atomicOr(&flags[FLAG_NL_OVERFLOW], 1);
        atomicMax(&flags[FLAG_NL_MAXCOUNT], int(cnt));
    }
}

// ---------------------------------------------------------- diagnostics -----
// Single block: out = KE_total, KE_trans, KE_rot, |P|, |L|, Rg, cm(3).
__global__ void __launch_bounds__(1024) k_diag(const double *x, const double *v, const double *w, uint32_t n,
                                               double mass, double inertia, double *out) {
    __shared__ double red[8][1024];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // 2*KE_trans, 2*KE_rot, sum v (3), L (3)
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        d3 vi = ld3(v, i), wi = ld3(w, i), xi = ld3(x, i);
        acc[0] += mass * dot(vi, vi);
        acc[1] += inertia * dot(wi, wi);
        acc[2] += vi.x; acc[3] += vi.y; acc[4] += vi.z;
        d3 L = inertia * wi + cross(xi, mass * vi);
        acc[5] += L.x; acc[6] += L.y; acc[7] += L.z;
    }
    for (int c = 0; c < 8; c++) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 8; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    d3 cmsum = zero3();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) cmsum = cmsum + ld3(x, i);
    __syncthreads();
    double k0 = red[0][0], k1 = red[1][0];
    d3 P{red[2][0], red[3][0], red[4][0]}, L{red[5][0], red[6][0], red[7][0]};
    __syncthreads();
    red[0][threadIdx.x] = cmsum.x; red[1][threadIdx.x] = cmsum.y; red[2][threadIdx.x] = cmsum.z;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    const d3 cm = d3{red[0][0], red[1][0], red[2][0]} / double(n);
    __syncthreads();
    double rg = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        d3 dd = ld3(x, i) - cm;
        rg += dot(dd, dd);
    }
    red[0][threadIdx.x] = rg;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[0][threadIdx.x] += red[0][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = k0 / 2.0 + k1 / 2.0;
        out[1] = k0 / 2.0;
        out[2] = k1 / 2.0;
        out[3] = mass * norm(P);
        out[4] = norm(L);
        out[5] = sqrt(red[0][0] / double(n));
        out[6] = cm.x;
        out[7] = cm.y;
        out[8] = cm.z;
    }
}

// Energy ledger terms of one particle (see launch_energy). Pair terms are
// split evenly between the two particles of a pair and each ordered pair's
// springs carry half of that pair's spring energy, so the sums over all
// particles count every pair once.
__global__ void __launch_bounds__(256) k_energy(hg_params P, energy_view E, double mass, double inertia,
                                                double *out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E.n) return;
    const uint32_t g = E.offset + k;
    const d3 xi = ld3(E.x, g), vi = ld3(E.v, g), wi = ld3(E.w, g);
    double e[8] = {0.5 * mass * dot(vi, vi), 0.5 * inertia * dot(wi, wi), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t cnt = E.pcnt ? E.pcnt[g] : 0u;
    for (uint32_t s = 0; s < cnt; s++) {
        const bool bonded = (E.pj[size_t(s) * E.stride + g] & kBondBit) != 0;
        const spring9 sp = ld_rec(E.ps + (size_t(s) * E.stride + g) * kSpringRec);
        const double kt = bonded ? P.sinter.k_t : P.contact.k_t, kr = bonded ? P.sinter.k_r : P.contact.k_r,
                     ko = bonded ? P.sinter.k_o : P.contact.k_o;
        e[3] += 0.25 * kt * dot(sp.s[0], sp.s[0]);
        e[4] += 0.25 * kr * dot(sp.s[1], sp.s[1]);
        e[5] += 0.25 * ko * dot(sp.s[2], sp.s[2]);
    }
    auto pair = [&](uint32_t j) {
        const double dist = norm(xi - ld3(E.x, j));
        uint32_t ls = kNoEntry;
        if (E.lconn)
            for (uint32_t s = 0; s < E.bcnt[g]; s++)
                if (E.bj[size_t(s) * E.stride + g] == j) ls = s;
        if (ls != kNoEntry) {
            // the bond's five connector springs, k_d/2 (|s| - L0)^2 each, half
            // to each ordered pair
            const double *ri = E.lconn + (size_t(ls) * E.stride + g) * kConnRec;
            const double *rj = E.lconn + (size_t(E.lslot[size_t(ls) * E.stride + g]) * E.stride + j) * kConnRec;
            for (int q = 0; q < 5; q++) {
                const d3 sq = d3{rj[3 * q], rj[3 * q + 1], rj[3 * q + 2]} - d3{ri[3 * q], ri[3 * q + 1], ri[3 * q + 2]};
                const double d = (q < 4 ? 2.0 * P.contact.r_part : 0.0) - norm(sq);
                e[2] += 0.25 * P.sinter.k * d * d;
            }
        } else if (P.binary_model != HG_BINARY_NONE) {
            bool bonded = false;
            for (uint32_t s = 0; s < cnt; s++) {
                const uint32_t tag = E.pj[size_t(s) * E.stride + g];
                if ((tag & kIdxMask) == j) bonded = (tag & kBondBit) != 0;
            }
            if (bonded) {
                const double ov = 2.0 * P.sinter.r_part - dist;
                e[2] += 0.25 * P.sinter.k * ov * ov;
            } else {
                const double ov = 2.0 * P.contact.r_part - dist;
                if (ov > 0.0) e[2] += 0.25 * P.contact.k * ov * ov;
            }
        }
        if (P.hamaker_model != HG_HAMAKER_NONE) e[6] += 0.5 * hamaker_potential(P.hamaker_model, P.hamaker, dist);
    };
    if (E.nb) {
        const uint32_t m = E.ncnt[g];
        for (uint32_t q = 0; q < m; q++) pair(E.nb[size_t(q) * E.stride + g]);
    } else {
        for (uint32_t j = E.offset; j < E.offset + E.n; j++)
            if (j != g) pair(j);
    }
    for (int c = 0; c < 8; c++) out[8 * size_t(k) + c] = e[c];
}

inline unsigned blocks_for(size_t n, unsigned bs) { return unsigned((n + bs - 1) / bs); }

template <int BIN, int HAM>
hipError_t launch_force_t(const hg_params &p, const sys_view &sv, hipStream_t s) {
    if (sv.n == 0) return hipSuccess;
    if (p.pair_mode == HG_PAIRS_NEIGHBOR_LIST && p.precision == HG_PRECISION_MIXED) {
        if constexpr (BIN == HG_BINARY_SINTER_LEGACY) return hipErrorInvalidValue;  // FP64 only
        else k_force_nl_mx<BIN, HAM><<<blocks_for(sv.n, 64), 64, 0, s>>>(p, make_mixed_consts(p), sv);
    } else if (sv.diss) {
        // dissipation tracking (FP64 contact and alt_sinter bonds only)
        if constexpr (BIN == HG_BINARY_CONTACT || BIN == HG_BINARY_SINTER) {
            if (p.pair_mode == HG_PAIRS_NEIGHBOR_LIST)
                k_force_nl<BIN, HAM, true><<<blocks_for(sv.n, 64), 64, 0, s>>>(p, sv);
            else
                k_force_all<BIN, HAM, true><<<blocks_for(sv.n, kTile), kTile, 0, s>>>(p, sv);
        } else {
            return hipErrorInvalidValue;
        }
    } else if (p.pair_mode == HG_PAIRS_NEIGHBOR_LIST) {
        // 64-lane blocks spread a modest particle count over more CUs.
        k_force_nl<BIN, HAM, false><<<blocks_for(sv.n, 64), 64, 0, s>>>(p, sv);
    } else {
        k_force_all<BIN, HAM, false><<<blocks_for(sv.n, kTile), kTile, 0, s>>>(p, sv);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_force(const hg_params &p, const sys_view &sv, hipStream_t s) {
#define HG_CASE(B, H)                                                       \
    if (p.binary_model == B && p.hamaker_model == H) return launch_force_t<B, H>(p, sv, s);
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_SIMPLIFIED)
#undef HG_CASE
    return hipErrorInvalidValue;
}

// nsteps steps of a system of at most kTinyMax particles in one workgroup of
// kTinyThreads, one thread per ordered pair (all pairs, or particle i's
// listed partner slot k): each pair's functors evaluated at once (its old
// springs looked up in the pair-state row of i), its term kept in LDS and its
// new springs in `scratch` (kSpringRec doubles per pair: the 9 components,
// then keep | bonded << 1). Then one thread per particle adds its terms in
// partner order and writes its new row as the walker of the per-particle
// pass does: old entries of partners not evaluated are carried in place, an
// evaluated partner's entry is replaced by its kept springs. So the sums and
// tables are those of k_force_all / k_force_nl. Contact, alt_sinter (bonds
// broken past their neck strength drop their entry), the legacy
// sinter_functor (its connectors moved after each drift, as k_legacy_move
// does) or no binary functor.
constexpr uint32_t kTinyMax = 27, kTinyThreads = 768;

template <int BIN, int HAM, bool NL>
__global__ void __launch_bounds__(kTinyThreads)
    k_tiny_steps(hg_params P, sys_view even, sys_view odd, double *x, double *v, double *th, double *w,
                 facet_dev *facets, uint32_t nsteps, int pending_first, int final_kick, double half, double dt,
                 int nf_advance, double *scratch, double *lconn, const uint32_t *bcnt) {
    __shared__ double terms[kTinyThreads * 6];
    const uint32_t t = threadIdx.x;
    const uint32_t n = even.n;
    constexpr bool nl = NL;
    const uint32_t slots = nl ? even.cap_n : n;  // pair threads per particle
    for (uint32_t s = 0; s < nsteps; s++) {
        const int pending = (s > 0 || pending_first) ? 1 : 0;
        if (pending && t == 0) advance_facets(facets, nf_advance, dt);
        if (t < n) {
            kick_drift_one(x, v, th, w, even.a, even.al, t, pending, half, half, dt);
            if constexpr (BIN == HG_BINARY_SINTER_LEGACY) legacy_move_one(x, v, w, n, t, dt, lconn, bcnt);
        }
        __syncthreads();
        const sys_view &S = (s & 1u) ? odd : even;
        if (t < n * slots) {
            const uint32_t i = t / slots, k = t % slots;
            bool eval;
            uint32_t j;
            if (nl) {
                eval = k < S.ncnt[i];
                j = eval ? S.nb[size_t(k) * n + i] : 0u;
            } else {
                eval = k != i;
                j = k;
            }
            if (eval) {
                const d3 xi = ld3(S.x, i), vi = ld3(S.v, i), wi = ld3(S.w, i);
                const d3 xj = ld3(S.x, j), vj = ld3(S.v, j), wj = ld3(S.w, j);
                acc6 b{zero3(), zero3()};
                spring9 sp = zero_spring();
                bool bonded = false, keep = false;
                uint32_t ls = kNoEntry;  // the legacy bond's slot in i's list
                if constexpr (BIN == HG_BINARY_SINTER_LEGACY)
                    for (uint32_t q = 0; q < S.pt.bcnt[i]; q++)
                        if (S.pt.bj[size_t(q) * n + i] == j) ls = q;
                if (ls != kNoEntry) {
                    // sinter_bridge.h:222-223: its five connector springs; the
                    // pair has no table entry (keep stays false)
                    const double *ri = S.pt.lconn + (size_t(ls) * n + i) * kConnRec;
                    const double *rj = S.pt.lconn + (size_t(S.pt.lslot[size_t(ls) * n + i]) * n + j) * kConnRec;
                    d3 ci[5], cj[5];
                    for (int q = 0; q < 5; q++) {
                        ci[q] = d3{ri[3 * q], ri[3 * q + 1], ri[3 * q + 2]};
                        cj[q] = d3{rj[3 * q], rj[3 * q + 1], rj[3 * q + 2]};
                    }
                    b = legacy_bonded_eval(P.sinter.k, P.sinter.gamma_n, P.contact.r_part, P.contact.mass,
                                           P.contact.inertia, ci, cj, xi, xj, vi, vj, wi, wj);
                } else if constexpr (BIN != HG_BINARY_NONE) {
                    const uint32_t cnt = S.pt.pcnt_old[i];
                    for (uint32_t q = 0; q < cnt; q++) {
                        const uint32_t tag = S.pt.pj_old[size_t(q) * n + i];
                        if ((tag & kIdxMask) == j) {
                            sp = ld_rec(S.pt.ps_old + (size_t(q) * n + i) * kSpringRec);
                            bonded = (tag & kBondBit) != 0;
                        }
                    }
                    // the bond's neck strength, as walker_t::neck finds it
                    auto neck = [&] {
                        double r = __builtin_huge_val();
                        if (S.pt.bcnt)
                            for (uint32_t q = 0; q < S.pt.bcnt[i]; q++)
                                if (S.pt.bj[size_t(q) * n + i] == j) r = S.pt.bstr[size_t(q) * n + i];
                        return r;
                    };
                    b = pair_binary<BIN>(P, xi, xj, vi, wi, [&](d3 &v, d3 &w) { v = vj, w = wj; }, sp, bonded,
                                         keep, neck);
                }
                const acc6 e = pair_term<BIN, HAM>(b, pair_hamaker<HAM>(P, xi, xj));
                double *tm = terms + size_t(t) * 6;
                tm[0] = e.lin.x; tm[1] = e.lin.y; tm[2] = e.lin.z;
                tm[3] = e.ang.x; tm[4] = e.ang.y; tm[5] = e.ang.z;
                if constexpr (BIN != HG_BINARY_NONE) {
                    double *r = scratch + size_t(t) * kSpringRec;
                    st_rec(r, sp);
                    r[9] = double((keep ? 1 : 0) | (bonded ? 2 : 0));
                }
            }
        }
        __syncthreads();
        if (t < n) {
            const uint32_t i = t;
            acc6 acc{zero3(), zero3()};
            uint32_t q = 0, p = 0;
            const uint32_t cnt_old = BIN != HG_BINARY_NONE ? S.pt.pcnt_old[i] : 0u;
            auto put = [&](uint32_t tag, const spring9 &sp) {
                if (q < S.pt.cap) {
                    S.pt.pj_new[size_t(q) * n + i] = tag;
                    st_rec(S.pt.ps_new + (size_t(q) * n + i) * kSpringRec, sp);
                }
                q++;
            };
            auto carry_below = [&](uint32_t j) {  // old entries of partners before j
                while (p < cnt_old && (S.pt.pj_old[size_t(p) * n + i] & kIdxMask) < j) {
                    put(S.pt.pj_old[size_t(p) * n + i], ld_rec(S.pt.ps_old + (size_t(p) * n + i) * kSpringRec));
                    p++;
                }
            };
            const uint32_t m = nl ? S.ncnt[i] : n;
            for (uint32_t k = 0; k < m; k++) {
                const uint32_t j = nl ? S.nb[size_t(k) * n + i] : k;
                if (!nl && j == i) continue;
                const double *tm = terms + size_t(i * slots + k) * 6;
                acc.lin = acc.lin + d3{tm[0], tm[1], tm[2]};
                acc.ang = acc.ang + d3{tm[3], tm[4], tm[5]};
                if constexpr (BIN != HG_BINARY_NONE) {
                    if constexpr (NL) {  // with all pairs every partner is evaluated: nothing to carry
                        carry_below(j);
                        if (p < cnt_old && (S.pt.pj_old[size_t(p) * n + i] & kIdxMask) == j) p++;  // replaced
                    }
                    const double *r = scratch + size_t(i * slots + k) * kSpringRec;
                    const int flags = int(r[9]);
                    if (flags & 1) put(j | ((flags & 2) ? kBondBit : 0u), ld_rec(r));
                }
            }
            if constexpr (BIN != HG_BINARY_NONE) {
                if constexpr (NL) carry_below(kNoEntry);
                S.pt.pcnt_new[i] = q < S.pt.cap ? q : S.pt.cap;
                if (q > S.pt.cap) {
                    atomicOr(&S.flags[FLAG_PAIR_OVERFLOW], 1);
                    atomicMax(&S.flags[FLAG_PAIR_MAXCOUNT], int(q));
                }
            }
            if (S.nf > 0) acc = acc + facet_fold(P, S.facets, S.nf, S.ss, S.n, i, ld3(S.x, i), ld3(S.v, i), ld3(S.w, i));
            st3(S.a, i, acc.lin);
            st3(S.al, i, acc.ang);
        }
        __syncthreads();
    }
    if (final_kick) {
        if (t == 0) advance_facets(facets, nf_advance, dt);
        if (t < n) final_kick_one(v, w, even.a, even.al, t, half, even.flags);
    }
}

template <int BIN, int HAM>
hipError_t launch_small_steps_t(const hg_params &p, const sys_view &even, const sys_view &odd, double *x, double *v,
                                double *th, double *w, facet_dev *facets, uint32_t nsteps, bool pending_first,
                                bool final_kick, double half, double dt, int nf_advance, double *lconn,
                                const uint32_t *bcnt, double *scratch, hipStream_t s) {
    const uint32_t slots = p.pair_mode == HG_PAIRS_NEIGHBOR_LIST ? even.cap_n : even.n;
    if (scratch && even.n <= kTinyMax && even.n * slots <= kTinyThreads) {
        if (p.pair_mode == HG_PAIRS_NEIGHBOR_LIST)
            k_tiny_steps<BIN, HAM, true><<<1, kTinyThreads, 0, s>>>(p, even, odd, x, v, th, w, facets, nsteps,
                                                                    pending_first ? 1 : 0, final_kick ? 1 : 0, half,
                                                                    dt, nf_advance, scratch, lconn, bcnt);
        else
            k_tiny_steps<BIN, HAM, false><<<1, kTinyThreads, 0, s>>>(p, even, odd, x, v, th, w, facets, nsteps,
                                                                     pending_first ? 1 : 0, final_kick ? 1 : 0, half,
                                                                     dt, nf_advance, scratch, lconn, bcnt);
        return hipGetLastError();
    }
    k_small_steps<BIN, HAM><<<1, kTile, 0, s>>>(p, even, odd, x, v, th, w, facets, nsteps, pending_first ? 1 : 0,
                                                final_kick ? 1 : 0, half, dt, nf_advance, lconn, bcnt);
    return hipGetLastError();
}

hipError_t launch_small_steps(const hg_params &p, const sys_view &even, const sys_view &odd, double *x, double *v,
                              double *th, double *w, facet_dev *facets, uint32_t nsteps, bool pending_first,
                              bool final_kick, double half, double dt, int nf_advance, double *lconn,
                              const uint32_t *bcnt, double *scratch, hipStream_t s) {
    if (even.n == 0 || even.n > uint32_t(kTile) || p.precision != HG_PRECISION_FP64) return hipErrorInvalidValue;
#define HG_CASE(B, H)                                                                                            \
    if (p.binary_model == B && p.hamaker_model == H)                                                             \
        return launch_small_steps_t<B, H>(p, even, odd, x, v, th, w, facets, nsteps, pending_first, final_kick, \
                                          half, dt, nf_advance, lconn, bcnt, scratch, s);
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_NONE, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_CONTACT, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_SINTER, HG_HAMAKER_SIMPLIFIED)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_NONE)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_FULL)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_LIMITED)
    HG_CASE(HG_BINARY_SINTER_LEGACY, HG_HAMAKER_SIMPLIFIED)
#undef HG_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_kick_drift(double *x, double *v, double *th, double *w, const double *a, const double *al,
                             uint32_t n, bool pending, double half_prev, double half, double dt, facet_dev *facets,
                             int nf, double facet_dt, hipStream_t s) {
    unsigned nb = blocks_for(n > 0 ? n : 1, 256);
    k_kick_drift<<<nb, 256, 0, s>>>(x, v, th, w, a, al, n, pending ? 1 : 0, half_prev, half, dt, facets, nf,
                                    facet_dt);
    return hipGetLastError();
}

hipError_t launch_legacy_move(const double *x, const double *v, const double *w, uint32_t n, double dt,
                              double *lconn, const uint32_t *bcnt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_legacy_move<<<blocks_for(n, 256), 256, 0, s>>>(x, v, w, n, dt, lconn, bcnt);
    return hipGetLastError();
}

hipError_t launch_final_kick(double *v, double *w, const double *a, const double *al, uint32_t n, double half,
                             facet_dev *facets, int nf, double facet_dt, int *flags, hipStream_t s) {
    unsigned nb = blocks_for(n > 0 ? n : 1, 256);
    k_final_kick<<<nb, 256, 0, s>>>(v, w, a, al, n, half, facets, nf, facet_dt, flags);
    return hipGetLastError();
}

hipError_t launch_advance_facet(facet_dev *f, const double *v, double dt, hipStream_t s) {
    k_advance_facet<<<1, 64, 0, s>>>(f, v[0], v[1], v[2], dt);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned nb = unsigned(std::min<size_t>(1024, (count + 255) / 256));
    k_zero_u32<<<nb, 256, 0, s>>>(p, count);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    unsigned nb = unsigned(std::min<size_t>(1024, (count + 255) / 256));
    k_zero_f64<<<nb, 256, 0, s>>>(p, count);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t count, uint32_t *tmp, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const uint32_t nblocks = (count + kScanChunk - 1) / kScanChunk;
    k_scan_blocks<<<nblocks, kScanThreads, 0, s>>>(in, out, count, tmp);
    k_scan_top<<<1, kScanThreads, 0, s>>>(tmp, nblocks);
    k_scan_add<<<nblocks, kScanThreads, 0, s>>>(out, count, tmp);
    return hipGetLastError();
}

uint32_t scan_tmp_words(uint32_t count) { return (count + kScanChunk - 1) / kScanChunk + 1; }

hipError_t launch_neighbor_build(hg_system_impl &S, hipStream_t s) {
    const uint32_t n = uint32_t(S.n);
    if (n == 0) return hipSuccess;
    const double cs = S.p.r_verlet * (1.0 + 1e-9);
    const uint32_t mask = S.hash_size - 1;
    hipError_t e;
    if ((e = launch_zero_u32(S.bucket_count, S.hash_size, s)) != hipSuccess) return e;
    if ((e = launch_zero_u32(S.bucket_fill, S.hash_size, s)) != hipSuccess) return e;
    k_bbox<<<1, 1024, 0, s>>>(S.x, n, S.bbox);
    k_cell_count<<<blocks_for(n, 256), 256, 0, s>>>(S.x, n, S.bbox, cs, mask, S.cell_of, S.bucket_count);
    if ((e = launch_exclusive_scan(S.bucket_count, S.bucket_start, S.hash_size, S.scan_tmp, s)) != hipSuccess)
        return e;
    k_cell_fill<<<blocks_for(n, 256), 256, 0, s>>>(n, S.cell_of, S.bucket_start, S.bucket_fill, S.cell_items);
    k_nl_fill<<<blocks_for(n, 64), 64, 0, s>>>(S.x, n, S.bbox, cs, mask, S.p.r_verlet, S.bucket_count,
                                               S.bucket_start, S.cell_items, S.nb, S.ncnt, S.cap_n, S.flags);
    return hipGetLastError();
}

hipError_t launch_diagnostics(const double *x, const double *v, const double *w, uint32_t n, double mass,
                              double inertia, double *out, hipStream_t s) {
    k_diag<<<1, 1024, 0, s>>>(x, v, w, n, mass, inertia, out);
    return hipGetLastError();
}

hipError_t launch_energy(const hg_params &p, const energy_view &E, double mass, double inertia, double *out,
                         hipStream_t s) {
    if (E.n == 0) return hipSuccess;
    k_energy<<<blocks_for(E.n, 256), 256, 0, s>>>(p, E, mass, inertia, out);
    return hipGetLastError();
}

}  // namespace hg
