// Device building blocks of the acceleration pass shared by the single-system
// kernels (kernels.hip) and the pool kernels (pool.hip): particle loads, the
// pair-state table walker, the binary functor fold and the facet fold.
//
// All arrays are "transposed" with an explicit stride: slot s of particle i
// lives at s*stride + i, so consecutive lanes (consecutive i) touch
// consecutive words.
#pragma once

#include <hip/hip_runtime.h>

#include "dem_mixed.h"
#include "hg_internal.h"

namespace hg {

__device__ __forceinline__ d3 ld3(const double *base, uint32_t i) {
    const double4 q = reinterpret_cast<const double4 *>(base)[i];
    return d3{q.x, q.y, q.z};
}
__device__ __forceinline__ void st3(double *base, uint32_t i, d3 a) {
    reinterpret_cast<double4 *>(base)[i] = double4{a.x, a.y, a.z, 0.0};
}

__device__ __forceinline__ spring9 ld_spring(const double *b, size_t stride) {
    return spring9{{d3{b[0], b[stride], b[2 * stride]}, d3{b[3 * stride], b[4 * stride], b[5 * stride]},
                    d3{b[6 * stride], b[7 * stride], b[8 * stride]}}};
}
__device__ __forceinline__ void st_spring(double *b, size_t stride, const spring9 &s) {
    b[0] = s.s[0].x; b[stride] = s.s[0].y; b[2 * stride] = s.s[0].z;
    b[3 * stride] = s.s[1].x; b[4 * stride] = s.s[1].y; b[5 * stride] = s.s[1].z;
    b[6 * stride] = s.s[2].x; b[7 * stride] = s.s[2].y; b[8 * stride] = s.s[2].z;
}

// A pair-state record (kSpringRec doubles, 16-byte aligned) in five 16-byte
// accesses; the pad word is written as zero.
__device__ __forceinline__ spring9 ld_rec(const double *r) {
    const double2 *q = reinterpret_cast<const double2 *>(r);
    const double2 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4];
    return spring9{{d3{a.x, a.y, b.x}, d3{b.y, c.x, c.y}, d3{d.x, d.y, e.x}}};
}
__device__ __forceinline__ void st_rec(double *r, const spring9 &s) {
    double2 *q = reinterpret_cast<double2 *>(r);
    q[0] = double2{s.s[0].x, s.s[0].y};
    q[1] = double2{s.s[0].z, s.s[1].x};
    q[2] = double2{s.s[1].y, s.s[1].z};
    q[3] = double2{s.s[2].x, s.s[2].y};
    q[4] = double2{s.s[2].z, 0.0};
}
static_assert(kSpringRec == 10, "ld_rec/st_rec move 10 doubles");

// Walks one particle's old pair-state table in ascending j alongside the
// partner loop, writing the new table. Entries not visited this pass (pairs
// that are not listed) are copied unchanged: the reference leaves the dense
// spring table untouched for pairs it does not evaluate.
// The first kTagPrefetch tags of the old row are loaded together with the
// count (the table capacity is at least that many slots), so walking the
// usual one to four entries needs no dependent load. The first kNewTagBuf
// tags of the new row are held in registers and written by finish(), when
// every lane writes its tag row s in the same instruction.
// EAGER loads the springs of old entry p together with its tag when the walk
// reaches p (every old entry is either taken or copied, so that load is never
// wasted) and keeps them in 18 registers until used, a whole pair evaluation
// later; otherwise they are loaded when used. The mixed-precision kernels
// walk eagerly; the FP64 ones have no registers to spare for it.
template <bool EAGER>
struct walker_t {
    const uint32_t *pj_old;
    const double *ps_old;
    uint32_t *pj_new;
    double *ps_new;
    const uint32_t *bj;
    const double *bstr;
    const uint32_t *bcnt;
    size_t stride;
    uint32_t i, cnt_old, p, q, cap, ej;
    uint32_t pre[kTagPrefetch];
    uint32_t nt[kNewTagBuf];
    spring9 rec;  // springs of old entry p (EAGER)
    // legacy sinter_functor bonds: cursor lb over i's lbc bonded partners
    // (ascending), lbn the partner at the cursor (kNoEntry past the end)
    const double *lconn;
    const uint32_t *lslot;
    uint32_t lb, lbc, lbn;

    __device__ void init(const pair_table &T, uint32_t i_) {
        pj_old = T.pj_old;
        ps_old = T.ps_old;
        pj_new = T.pj_new;
        ps_new = T.ps_new;
        bj = T.bj;
        bstr = T.bstr;
        bcnt = T.bcnt;
        stride = T.stride;
        i = i_;
        cap = T.cap;
        cnt_old = T.pcnt_old[i];
#pragma unroll
        for (uint32_t s = 0; s < kTagPrefetch; s++) pre[s] = pj_old[size_t(s) * stride + i];
        p = 0;
        q = 0;
        // ej and rec describe old entry p only while p < cnt_old; they are
        // loaded into the same registers whatever the count (slot 0 always
        // exists), so no register copy has to wait for the loads
        ej = pre[0];
        if constexpr (EAGER) rec = ld_rec(ps_old + size_t(i) * kSpringRec);
#pragma unroll
        for (uint32_t s = 0; s < kNewTagBuf; s++) nt[s] = kNoEntry;
    }
    __device__ void init_legacy(const pair_table &T) {
        lconn = T.lconn;
        lslot = T.lslot;
        lb = 0;
        lbc = bcnt[i];
        lbn = lbc > 0 ? bj[i] : kNoEntry;
    }
    // Partners arrive in ascending order, so the cursor only moves forward.
    __device__ bool legacy_bonded(uint32_t j, uint32_t &slot) {
        while (lbn < j) {
            lb++;
            lbn = lb < lbc ? bj[size_t(lb) * stride + i] : kNoEntry;
        }
        slot = lb;
        return lbn == j;
    }
    __device__ uint32_t tag_at(uint32_t s) const {
        uint32_t t = pre[0];
#pragma unroll
        for (uint32_t k = 1; k < kTagPrefetch; k++) t = s == k ? pre[k] : t;
        return s < kTagPrefetch ? t : pj_old[size_t(s) * stride + i];
    }
    // Past the last entry the registers are reloaded from the last slot again
    // (just read, so cached): a conditional reload would leave two values for
    // the compiler to merge, and the merge waits for the load at once.
    __device__ void next() {
        p++;
        const uint32_t pc = p < cnt_old ? p : p - 1;
        ej = pj_old[size_t(pc) * stride + i];
        if constexpr (EAGER) rec = ld_rec(ps_old + (size_t(pc) * stride + i) * kSpringRec);
    }
    __device__ spring9 load() const {
        if constexpr (EAGER) return rec;
        else return ld_rec(ps_old + (size_t(p) * stride + i) * kSpringRec);
    }
    __device__ void store(uint32_t tag, const spring9 &s) {
        if (q < cap) {
            if (q >= kNewTagBuf) pj_new[size_t(q) * stride + i] = tag;
#pragma unroll
            for (uint32_t k = 0; k < kNewTagBuf; k++) nt[k] = q == k ? tag : nt[k];
            st_rec(ps_new + (size_t(q) * stride + i) * kSpringRec, s);
        }
        q++;
    }
    // Copies the entries below j (pairs no longer listed; rare). The loop is
    // entered through a separate test and its path ends with a full wait for
    // its loads, so that the common path past it keeps exact wait counts
    // (the compiler's count after a loop that loads is otherwise zero).
    __device__ void advance_to(uint32_t j) {
        if (p < cnt_old && (ej & kIdxMask) < j) {
            do {
                store(ej, load());
                next();
            } while (p < cnt_old && (ej & kIdxMask) < j);
            __builtin_amdgcn_s_waitcnt(0);
        }
    }
    __device__ bool take(uint32_t j, spring9 &s, bool &bonded) {
        if (p < cnt_old && (ej & kIdxMask) == j) {
            s = load();
            bonded = (ej & kBondBit) != 0;
            next();
            return true;
        }
        return false;
    }
    // Neck strength of the bond between i and j (infinite when none is set).
    __device__ double neck(uint32_t j) const {
        if (!bcnt) return __builtin_huge_val();
        const uint32_t c = bcnt[i];
        for (uint32_t s = 0; s < c; s++)
            if (bj[size_t(s) * stride + i] == j) return bstr[size_t(s) * stride + i];
        return __builtin_huge_val();
    }
    __device__ void finish() {
        while (p < cnt_old) {
            store(ej, load());
            next();
        }
        const uint32_t m = q < cap ? q : cap;
#pragma unroll
        for (uint32_t s = 0; s < kNewTagBuf; s++)
            if (s < m) pj_new[size_t(s) * stride + i] = nt[s];
    }
};

// A particle's listed partners; the first kIdPrefetch slots are loaded
// together with the count (slots below the list capacity only).
struct id_list {
    const uint32_t *nb;
    size_t stride;
    uint32_t i, n;
    uint32_t pre[kIdPrefetch];

    __device__ void init(const uint32_t *nb_, const uint32_t *ncnt, size_t stride_, uint32_t i_, uint32_t cap) {
        nb = nb_;
        stride = stride_;
        i = i_;
        n = ncnt[i];
#pragma unroll
        for (uint32_t k = 0; k < kIdPrefetch; k++) pre[k] = k < cap ? nb[size_t(k) * stride + i] : 0u;
    }
    __device__ uint32_t at(uint32_t k) const {
        uint32_t t = pre[0];
#pragma unroll
        for (uint32_t s = 1; s < kIdPrefetch; s++) t = k == s ? pre[s] : t;
        return k < kIdPrefetch ? t : nb[size_t(k) * stride + i];
    }
    // Slot k, or the last slot for k >= n (n >= 1), always from memory.
    __device__ uint32_t at_clamped(uint32_t k) const { return nb[size_t(k < n ? k : n - 1) * stride + i]; }
};

// The first n listed partners of local particle t of a system, copied to LDS
// as offsets from the system's first particle `base`: slot k at
// ids[k*nsys + t] (consecutive lanes read consecutive halfwords). LDS loads
// are counted apart from global loads and stores, so waiting for an id does
// not wait for the pair-state stores issued after it.
struct lds_id_list {
    const uint16_t *ids;
    uint32_t nsys, t, base, n;
    __device__ uint32_t at(uint32_t k) const { return base + ids[k * nsys + t]; }
    __device__ uint32_t at_clamped(uint32_t k) const { return at(k < n ? k : n - 1); }
};

// A particle's listed partners in list order: fold(j, xj, vj, wj) for each,
// with load(j, x, v, w) fetching a partner's state. The id of partner k+2 and
// the state of partner k+1 are loaded before pair k is evaluated, so the state
// load waits on an id that arrived an iteration earlier and both loads overlap
// the pair's arithmetic. The loads are unconditional (past the end of the list
// they fetch the last partner again, from cache): a conditional load leaves
// two values for the compiler to merge, and the merge waits for the load and
// every memory operation issued before it.
// IDS is id_list (global memory) or lds_id_list. load(j) returns what is
// gathered ahead for partner j, the fold gets it: positions alone where the
// state sits in LDS (the fold then loads a partner's velocities itself, for
// the bonded and overlapping pairs only, a few of a 3.2r list), or the whole
// pstate from global memory, whose latency only the prefetch hides. PREFETCH
// = false loads each partner when its pair comes.
struct pstate {
    d3 x, v, w;
};
template <bool PREFETCH = true, class IDS, class LOAD, class FOLD>
__device__ __forceinline__ void nl_partners(const IDS &ids, LOAD &&load, FOLD &&fold) {
    if (ids.n == 0) return;
    if constexpr (!PREFETCH) {
        uint32_t jn = ids.at(0);
        for (uint32_t k = 0; k < ids.n; k++) {
            const uint32_t j = jn;
            const auto sj = load(j);
            jn = ids.at_clamped(k + 1);
            fold(j, sj);
        }
        return;
    }
    uint32_t jn = ids.at(0), jn2 = ids.at_clamped(1);
    auto sn = load(jn);
    for (uint32_t k = 0; k < ids.n; k++) {
        const uint32_t j = jn;
        const auto sj = sn;
        jn = jn2;
        sn = load(jn);
        jn2 = ids.at_clamped(k + 2);
        fold(j, sj);
    }
}

// The binary functor of one ordered pair: alt_sinter's bonded model for a
// bonded pair, the contact functor otherwise; keep says whether the pair's
// springs stay in the table; neck() gives the bond's neck strength (read only
// when bonds break); vw(vj, wj) the partner's velocities (called for bonded
// and overlapping pairs only).
// TRACK adds the pair's dissipated work to d4 (pair_dissipation).
template <int BIN, bool TRACK = false, class VW, class NECK>
__device__ __forceinline__ acc6 pair_binary(const hg_params &P, d3 xi, d3 xj, d3 vi, d3 wi, VW &&vw, spring9 &sp,
                                            bool bonded, bool &keep, NECK &&neck, double *d4 = nullptr) {
    acc6 b;
    if (BIN == HG_BINARY_SINTER && bonded) {
        d3 ftot, vj, wj;
        vw(vj, wj);
        b = sinter_eval<TRACK>(P.sinter, xi, xj, vi, vj, wi, wj, sp, ftot, d4);
        // alt_breaking_sinter_bridge.h:117-119: (i, j) stops being bonded
        // once its force exceeds the neck strength; dropping the entry
        // leaves the contact functor a pair with zero springs, as the
        // reference's untouched contact table holds.
        keep = !(P.sinter_breaking && norm(ftot) > neck());
    } else {
        b = contact_eval_vw<TRACK>(P.contact, xi, xj, vi, wi, vw, sp, keep, d4);
    }
    return b;
}

template <int HAM>
__device__ __forceinline__ acc6 pair_hamaker(const hg_params &P, d3 xi, d3 xj) {
    acc6 h{zero3(), zero3()};
    if constexpr (HAM == HG_HAMAKER_FULL) h = hamaker_eval(P.hamaker, xi, xj);
    if constexpr (HAM == HG_HAMAKER_LIMITED) h = lr_hamaker_eval(P.hamaker, xi, xj);
    if constexpr (HAM == HG_HAMAKER_SIMPLIFIED) h = simplified_hamaker_eval(P.hamaker, xi, xj);
    return h;
}

// The pair's term (the binary container's fold of its functors,
// operator_overloads.h:10-20) added to the particle's sum.
template <int BIN, int HAM>
__device__ __forceinline__ acc6 pair_term(const acc6 &b, const acc6 &h) {
    if constexpr (BIN != HG_BINARY_NONE && HAM != HG_HAMAKER_NONE) return b + h;
    else if constexpr (BIN != HG_BINARY_NONE) return b;
    else return h;
}
template <int BIN, int HAM>
__device__ __forceinline__ void fold_pair(acc6 &acc, const acc6 &b, const acc6 &h) {
    const acc6 e = pair_term<BIN, HAM>(b, h);
    acc.lin = acc.lin + e.lin;
    acc.ang = acc.ang + e.ang;
}

// One ordered pair through the binary container fold
// (binary_force_container.h:20-32): {contact | alt_sinter} + {hamaker}.
// vw(vj, wj) loads the partner's velocities when the pair needs them.
template <int BIN, int HAM, bool TRACK = false, bool EAGER, class VW>
__device__ __forceinline__ void pair_fold_vw(const hg_params &P, uint32_t j, d3 xi, d3 vi, d3 wi, d3 xj, VW &&vw,
                                             walker_t<EAGER> &W, acc6 &acc, double *d4 = nullptr) {
    acc6 b{zero3(), zero3()}, h{zero3(), zero3()};
    uint32_t lslot_ij = 0;
    if (BIN == HG_BINARY_SINTER_LEGACY && W.legacy_bonded(j, lslot_ij)) {
        d3 vj, wj;
        vw(vj, wj);
        // sinter_bridge.h:222-223: the bond's connectors, i's at its slot for
        // j, j's at its slot for i
        const double *ri = W.lconn + (size_t(lslot_ij) * W.stride + W.i) * kConnRec;
        const double *rj = W.lconn + (size_t(W.lslot[size_t(lslot_ij) * W.stride + W.i]) * W.stride + j) * kConnRec;
        d3 ci[5], cj[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            ci[q] = d3{ri[3 * q], ri[3 * q + 1], ri[3 * q + 2]};
            cj[q] = d3{rj[3 * q], rj[3 * q + 1], rj[3 * q + 2]};
        }
        b = legacy_bonded_eval(P.sinter.k, P.sinter.gamma_n, P.contact.r_part, P.contact.mass, P.contact.inertia, ci,
                               cj, xi, xj, vi, vj, wi, wj);
    } else if constexpr (BIN != HG_BINARY_NONE) {
        W.advance_to(j);
        spring9 sp = zero_spring();
        bool bonded = false;
        W.take(j, sp, bonded);
        bool keep;
        b = pair_binary<BIN, TRACK>(P, xi, xj, vi, wi, vw, sp, bonded, keep, [&] { return W.neck(j); }, d4);
        if (keep) W.store(j | (bonded ? kBondBit : 0u), sp);
    }
    h = pair_hamaker<HAM>(P, xi, xj);
    fold_pair<BIN, HAM>(acc, b, h);
}
template <int BIN, int HAM, bool EAGER>
__device__ __forceinline__ void pair_fold(const hg_params &P, uint32_t j, d3 xi, d3 vi, d3 wi, d3 xj, d3 vj, d3 wj,
                                          walker_t<EAGER> &W, acc6 &acc) {
    pair_fold_vw<BIN, HAM>(P, j, xi, vi, wi, xj, [&](d3 &v, d3 &w) { v = vj, w = wj; }, W, acc);
}

// Unary container over facets (unary_force_container.h:20-32, a right fold),
// each facet folding its surface functors (triangular_facet.h:32-51). Surface
// springs of facet f for particle i at ss + f*9*stride + i.
__device__ __forceinline__ acc6 facet_fold(const hg_params &P, const facet_dev *facets, int nf, double *ss,
                                           size_t stride, uint32_t i, d3 xi, d3 vi, d3 wi) {
    acc6 s{zero3(), zero3()};
    for (int f = nf - 1; f >= 0; f--) {
        const facet_dev F = facets[f];
        d3 xf = closest_point_on_triangle(F, xi);
        acc6 c{zero3(), zero3()}, h{zero3(), zero3()};
        if (P.facet_contact) {
            double *b = ss + size_t(f) * 9 * stride + i;
            spring9 sp = ld_spring(b, stride);
            c = surface_contact_eval(P.surface_contact, xf, F.v_facet, xi, vi, wi, sp);
            st_spring(b, stride, sp);
        }
        if (P.facet_hamaker) h = surface_hamaker_eval(P.surface_hamaker, xf, xi);
        acc6 e = (P.facet_contact && P.facet_hamaker) ? c + h : (P.facet_contact ? c : h);
        s = (f == nf - 1) ? e : e + s;
    }
    return s;
}

// Mixed-precision fold of one ordered pair (dem_mixed.h); the spring walk is
// the FP64 one. vw as in pair_fold_vw.
// A pair that is not bonded and lies beyond far2 is out of contact (its
// springs, if any, are dropped as the contact functor drops them) and takes
// the FP32-distance Hamaker form; the others take the FP64 kinematics.
template <int BIN, int HAM, bool EAGER, class VW>
__device__ __forceinline__ void pair_fold_mx(const mixed_consts &m, uint32_t j, d3 xi, d3 vi, d3 wi, d3 xj, VW &&vw,
                                             walker_t<EAGER> &W, accf &acc) {
    const d3 d = xi - xj;
    const double d2 = dotm(d, d);
    bool near = !(d2 > m.far2);
    accf b{zerof(), zerof()};
    f3 h = zerof();
    if constexpr (BIN != HG_BINARY_NONE) {
        W.advance_to(j);
        spring9 sp = zero_spring();
        bool bonded = false;
        W.take(j, sp, bonded);
        near = near || (BIN == HG_BINARY_SINTER && bonded);
        if (near) {
            double dist;
            d3 n64;
            f3 n;
            pair_kinematics_mx(xi, xj, dist, n64, n);
            bool keep;
            if (BIN == HG_BINARY_SINTER && bonded) {
                b = bonded_or_contact_mx<true>(m, dist, n64, n, vi, wi, vw, sp, keep);
                if (m.s_breaking && sqrtf(dotf(b.lin, b.lin)) * m.s_m > float(W.neck(j))) keep = false;
            } else
                b = bonded_or_contact_mx<false>(m, dist, n64, n, vi, wi, vw, sp, keep);
            if (keep) W.store(j | (bonded ? kBondBit : 0u), sp);
            if constexpr (HAM != HG_HAMAKER_NONE) h = hamaker_mx<HAM>(m, dist, n);
        }
    } else if (near) {
        if constexpr (HAM != HG_HAMAKER_NONE) {
            double dist;
            d3 n64;
            f3 n;
            pair_kinematics_mx(xi, xj, dist, n64, n);
            h = hamaker_mx<HAM>(m, dist, n);
        }
    }
    if constexpr (HAM != HG_HAMAKER_NONE)
        if (!near) h = hamaker_far_mx<HAM>(m, d, d2);
    acc.lin = acc.lin + (b.lin + h);
    acc.ang = acc.ang + b.ang;
}

__device__ __forceinline__ accf facet_fold_mx(const hg_params &P, const mixed_consts &m, const facet_dev *facets,
                                              int nf, double *ss, size_t stride, uint32_t i, d3 xi, d3 vi, d3 wi) {
    accf s{zerof(), zerof()};
    for (int f = nf - 1; f >= 0; f--) {
        const facet_dev F = facets[f];
        const d3 xf = closest_point_on_triangle(F, xi);
        if (P.facet_contact) {
            double *b = ss + size_t(f) * 9 * stride + i;
            spring9 sp = ld_spring(b, stride);
            s = s + surface_contact_mx(m, xf, F.v_facet, xi, vi, wi, sp);
            st_spring(b, stride, sp);
        }
        if (P.facet_hamaker) s.lin = s.lin + surface_hamaker_mx(m, xf, xi);
    }
    return s;
}

// Closes particle i's new pair-state row. An overflowing row raises the
// flags (and, in a pool, its system's fault word `sys_fault`, may be null).
template <int BIN, bool EAGER>
__device__ __forceinline__ void close_table(walker_t<EAGER> &W, uint32_t *pcnt_new, uint32_t i, int *flags,
                                            int *sys_fault = nullptr) {
    if constexpr (BIN != HG_BINARY_NONE) {
        W.finish();
        pcnt_new[i] = W.q < W.cap ? W.q : W.cap;
        if (W.q > W.cap) {
            atomicOr(&flags[FLAG_PAIR_OVERFLOW], 1);
            atomicMax(&flags[FLAG_PAIR_MAXCOUNT], int(W.q));
            if (sys_fault) atomicOr(sys_fault, HG_FAULT_PAIR_CAPACITY);
        }
    }
}

// True when any component of the closing kick's v or omega is NaN or infinite.
__device__ __forceinline__ bool nonfinite(d3 v, d3 w) {
    return !(__builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z) &&
             __builtin_isfinite(w.x) && __builtin_isfinite(w.y) && __builtin_isfinite(w.z));
}

// Facet vertices advance by v_facet*dt (triangular_facet.h:62-66).
__device__ __forceinline__ void advance_facets(facet_dev *facets, int nf, double dt) {
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) facets[f].vert[k] = facets[f].vert[k] + facets[f].v_facet * dt;
}

}  // namespace hg
