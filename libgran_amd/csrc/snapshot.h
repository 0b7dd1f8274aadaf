// Snapshot of one system (hg_save / hg_load, hg_pool_save / hg_pool_load):
// everything a step reads, so that a system reloaded from it continues
// exactly as the original would have. Host code only.
//
// The reference has no such format: its only dump is the VTK file of
// positions and orientation axes (writer.cpp:51-118), which cannot restart a
// run. Rows (neighbour list, pair-state table, neck strengths, legacy
// connectors) are stored as CSR over local particle indices, so a snapshot
// taken from a pool loads into a single system and the other way round.
//
// Byte layout (little-endian, as written by the host):
//   "HGSNAP01"  u32 version  u32 sizeof(hg_params)  hg_params
//   u64 n  f64 t  u64 step_index  u32 acc_valid  u32 needs_rebuild  u32 nf
//   u32 has_nl  u32 has_table  u32 has_neck  u32 has_legacy  u32 reserved
//   u64 nbonds  u64 skipped  u64 nl_entries  u64 pt_entries  u64 neck_entries
//   u64 legacy_entries
//   x, v, theta, omega, a, alpha         6 x 3n f64
//   [has_nl]    nl_count[n] u32, nl_ids[nl_entries] u32
//   [has_table] pt_count[n] u32, pt_tags[pt_entries] u32 (local j | bond bit),
//               pt_springs[9 pt_entries] f64
//   bonds[2 nbonds] u32
//   facets[nf] (the device records: vertices, velocity, area, normal),
//   facet springs[nf * 9 * n] f64 (facet f, component c, particle i)
//   [has_neck]   neck_count[n] u32, neck_j[neck_entries] u32,
//                neck_strength[neck_entries] f64
//   [has_legacy] leg_count[n] u32, leg_j[legacy_entries] u32,
//                leg_slot[legacy_entries] u32, leg_conn[15 legacy_entries] f64
//   u64 FNV-1a hash of every byte before it
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "hg_internal.h"

namespace hg {

struct snapshot {
    hg_params p{};
    uint64_t n = 0, step_index = 0;
    double t = 0.0;
    uint32_t acc_valid = 0, needs_rebuild = 0, nf = 0;
    bool has_nl = false, has_table = false, has_neck = false, has_legacy = false;
    std::vector<double> x, v, th, w, a, al;  // 3n each
    std::vector<uint32_t> nl_count, nl_ids;
    std::vector<uint32_t> pt_count, pt_tags;
    std::vector<double> pt_springs;  // 9 per entry
    std::vector<uint32_t> bonds;     // 2 per bond
    uint64_t skipped = 0;
    std::vector<facet_dev> facets;
    std::vector<double> ss;  // nf * 9 * n
    std::vector<uint32_t> neck_count, neck_j;
    std::vector<double> neck_str;
    std::vector<uint32_t> leg_count, leg_j, leg_slot;
    std::vector<double> leg_conn;  // 15 per entry
};

constexpr char kSnapMagic[8] = {'H', 'G', 'S', 'N', 'A', 'P', '0', '1'};
constexpr uint32_t kSnapVersion = 1;

inline uint64_t fnv1a(const uint8_t *b, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t k = 0; k < n; k++) {
        h ^= b[k];
        h *= 1099511628211ull;
    }
    return h;
}

inline std::vector<uint8_t> encode_snapshot(const snapshot &s) {
    std::vector<uint8_t> out;
    auto put = [&](const void *p, size_t bytes) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        out.insert(out.end(), b, b + bytes);
    };
    auto u32 = [&](uint32_t v) { put(&v, 4); };
    auto u64 = [&](uint64_t v) { put(&v, 8); };
    auto f64 = [&](double v) { put(&v, 8); };
    auto vec = [&](const auto &v) { put(v.data(), v.size() * sizeof(v[0])); };
    put(kSnapMagic, 8);
    u32(kSnapVersion);
    u32(uint32_t(sizeof(hg_params)));
    put(&s.p, sizeof(hg_params));
    u64(s.n);
    f64(s.t);
    u64(s.step_index);
    u32(s.acc_valid);
    u32(s.needs_rebuild);
    u32(s.nf);
    u32(s.has_nl);
    u32(s.has_table);
    u32(s.has_neck);
    u32(s.has_legacy);
    u32(0);
    u64(s.bonds.size() / 2);
    u64(s.skipped);
    u64(s.nl_ids.size());
    u64(s.pt_tags.size());
    u64(s.neck_j.size());
    u64(s.leg_j.size());
    for (const auto *f : {&s.x, &s.v, &s.th, &s.w, &s.a, &s.al}) vec(*f);
    if (s.has_nl) {
        vec(s.nl_count);
        vec(s.nl_ids);
    }
    if (s.has_table) {
        vec(s.pt_count);
        vec(s.pt_tags);
        vec(s.pt_springs);
    }
    vec(s.bonds);
    vec(s.facets);
    vec(s.ss);
    if (s.has_neck) {
        vec(s.neck_count);
        vec(s.neck_j);
        vec(s.neck_str);
    }
    if (s.has_legacy) {
        vec(s.leg_count);
        vec(s.leg_j);
        vec(s.leg_slot);
        vec(s.leg_conn);
    }
    u64(fnv1a(out.data(), out.size()));
    return out;
}

// Checks the hash, the header and every count against the buffer size and
// the row counts before anything is read into `s`; HG_ERR_INVALID on any
// mismatch.
inline hg_status decode_snapshot(const void *buf, size_t size, snapshot &s) {
    const uint8_t *b = static_cast<const uint8_t *>(buf);
    if (!buf || size < 8 + 8 + 8) return fail(HG_ERR_INVALID, "snapshot: buffer too small");
    if (std::memcmp(b, kSnapMagic, 8) != 0) return fail(HG_ERR_INVALID, "snapshot: not a hipgran snapshot");
    uint64_t stored;
    std::memcpy(&stored, b + size - 8, 8);
    if (fnv1a(b, size - 8) != stored) return fail(HG_ERR_INVALID, "snapshot: checksum mismatch (damaged data)");
    size_t at = 8;
    const size_t end = size - 8;
    bool ok = true;
    auto get = [&](void *p, size_t bytes) {
        if (!ok || bytes > end - at) {
            ok = false;
            return;
        }
        std::memcpy(p, b + at, bytes);
        at += bytes;
    };
    auto u32 = [&] {
        uint32_t v = 0;
        get(&v, 4);
        return v;
    };
    auto u64 = [&] {
        uint64_t v = 0;
        get(&v, 8);
        return v;
    };
    auto vec = [&](auto &v, uint64_t count) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (!ok || count > (end - at) / sizeof(T)) {
            ok = false;
            return;
        }
        v.resize(count);
        get(v.data(), count * sizeof(T));
    };
    if (u32() != kSnapVersion) return fail(HG_ERR_INVALID, "snapshot: unknown version");
    if (u32() != sizeof(hg_params)) return fail(HG_ERR_INVALID, "snapshot: parameter block of another size");
    get(&s.p, sizeof(hg_params));
    s.n = u64();
    get(&s.t, 8);
    s.step_index = u64();
    s.acc_valid = u32();
    s.needs_rebuild = u32();
    s.nf = u32();
    s.has_nl = u32() != 0;
    s.has_table = u32() != 0;
    s.has_neck = u32() != 0;
    s.has_legacy = u32() != 0;
    (void)u32();
    const uint64_t nbonds = u64();
    s.skipped = u64();
    const uint64_t nl_entries = u64(), pt_entries = u64(), neck_entries = u64(), leg_entries = u64();
    if (!ok || s.n == 0 || s.n >= (uint64_t(1) << 31) || s.nf > uint32_t(kMaxFacets))
        return fail(HG_ERR_INVALID, "snapshot: bad header");
    const uint64_t n = s.n;
    for (auto *f : {&s.x, &s.v, &s.th, &s.w, &s.a, &s.al}) vec(*f, 3 * n);
    auto rows_ok = [&](const std::vector<uint32_t> &count, const std::vector<uint32_t> &ids, bool tags) {
        uint64_t tot = 0;
        for (uint32_t c : count) tot += c;
        if (tot != ids.size()) return false;
        for (uint32_t j : ids)
            if ((tags ? (j & kIdxMask) : j) >= n) return false;
        return true;
    };
    if (s.has_nl) {
        vec(s.nl_count, n);
        vec(s.nl_ids, nl_entries);
        if (ok && !rows_ok(s.nl_count, s.nl_ids, false)) return fail(HG_ERR_INVALID, "snapshot: bad neighbour rows");
    }
    if (s.has_table) {
        vec(s.pt_count, n);
        vec(s.pt_tags, pt_entries);
        vec(s.pt_springs, 9 * pt_entries);
        if (ok && !rows_ok(s.pt_count, s.pt_tags, true)) return fail(HG_ERR_INVALID, "snapshot: bad pair-state rows");
    }
    vec(s.bonds, 2 * nbonds);
    for (uint32_t j : s.bonds)
        if (ok && j >= n) return fail(HG_ERR_INVALID, "snapshot: bad bond");
    vec(s.facets, s.nf);
    vec(s.ss, uint64_t(s.nf) * 9 * n);
    if (s.has_neck) {
        vec(s.neck_count, n);
        vec(s.neck_j, neck_entries);
        vec(s.neck_str, neck_entries);
        if (ok && !rows_ok(s.neck_count, s.neck_j, false)) return fail(HG_ERR_INVALID, "snapshot: bad neck rows");
    }
    if (s.has_legacy) {
        vec(s.leg_count, n);
        vec(s.leg_j, leg_entries);
        vec(s.leg_slot, leg_entries);
        vec(s.leg_conn, 15 * leg_entries);
        if (ok && !rows_ok(s.leg_count, s.leg_j, false)) return fail(HG_ERR_INVALID, "snapshot: bad legacy rows");
    }
    if (!ok || at != end) return fail(HG_ERR_INVALID, "snapshot: sizes disagree with the buffer");
    return HG_OK;
}

// Host copies of transposed device rows (entry s of particle i at
// (s*n + i)*width, n the system's particle count) to CSR and back. `used`
// of the `width` values of an entry are kept. For index rows, `shift` is
// subtracted from the index bits on the way out (arena to local indices)
// and added on the way in; a set bond bit is kept.
template <typename T>
inline void rows_to_csr(const T *rows, const uint32_t *count, uint32_t n, size_t width, size_t used,
                        std::vector<T> &out) {
    out.clear();
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t s = 0; s < count[i]; s++)
            for (size_t c = 0; c < used; c++) out.push_back(rows[(size_t(s) * n + i) * width + c]);
}
inline void shift_ids(std::vector<uint32_t> &ids, int64_t shift) {
    for (auto &t : ids) t = (t & kBondBit) | uint32_t(int64_t(t & kIdxMask) - shift);
}
template <typename T>
inline void csr_to_rows(const std::vector<uint32_t> &count, const std::vector<T> &vals, uint32_t n, uint32_t cap,
                        size_t width, size_t used, T fill, std::vector<T> &rows) {
    rows.assign(size_t(cap) * n * width, fill);
    size_t k = 0;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t s = 0; s < count[i]; s++, k++)
            for (size_t c = 0; c < used; c++) rows[(size_t(s) * n + i) * width + c] = vals[k * used + c];
}
inline uint32_t max_count(const std::vector<uint32_t> &count) {
    uint32_t m = 0;
    for (uint32_t c : count) m = c > m ? c : m;
    return m;
}
// 4 doubles per particle (device layout) <-> 3.
inline std::vector<double> pack3(const std::vector<double> &q4, uint32_t n) {
    std::vector<double> out(3 * size_t(n));
    for (uint32_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) out[3 * size_t(i) + c] = q4[4 * size_t(i) + c];
    return out;
}

}  // namespace hg
