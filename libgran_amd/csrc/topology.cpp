// Bond topology of alt_sinter_functor, computed on the host at hg_create.
//
// Same decisions as the reference constructor (alt_sinter_bridge.h:56-90):
// candidate pairs i<j are taken in lexicographic order; a pair is a bond
// candidate when fabs(norm(x_i - x_j) - 2 r) < d_crit; a candidate whose two
// ends already share a component is skipped (the reference prints a warning
// and keeps the bond graph a forest); otherwise it is bonded and the two
// components merge. The reference scans all n^2/2 pairs and relabels a whole
// component per merge; here candidates come from a cell list and components
// from a union-find, which yields the identical bond list.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <utility>
#include <vector>

#include "hg_internal.h"

namespace hg {

namespace {
struct uf {
    std::vector<uint32_t> parent;
    explicit uf(size_t n) : parent(n) { std::iota(parent.begin(), parent.end(), 0u); }
    uint32_t find(uint32_t a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    }
};
}  // namespace

void sinter_topology_host(size_t n, const double *x0, double r_part, double d_crit,
                          std::vector<std::pair<uint32_t, uint32_t>> &bonds, size_t &skipped, bool prevent_cycles,
                          bool inclusive) {
    bonds.clear();
    skipped = 0;
    if (n < 2 || !(d_crit > 0.0 || (inclusive && d_crit == 0.0))) return;
    auto X = [&](size_t i) { return d3{x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}; };
    auto pred = [&](size_t i, size_t j) {
        const double gap = std::fabs(norm(X(i) - X(j)) - 2.0 * r_part);
        return inclusive ? gap <= d_crit : gap < d_crit;
    };

    std::vector<std::pair<uint32_t, uint32_t>> cand;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], x0[3 * i + c]);
            hi[c] = std::max(hi[c], x0[3 * i + c]);
        }
    // Any candidate has centre distance <= 2r + d_crit; cells of that edge
    // (plus a relative margin for rounding) contain every candidate in the
    // 27-cell neighbourhood.
    const double cs = (2.0 * r_part + d_crit) * (1.0 + 1e-9);
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    if (!(cs > 0.0) || ext / cs > 1.0e6 || n < 64) {
        for (size_t i = 0; i + 1 < n; i++)
            for (size_t j = i + 1; j < n; j++)
                if (pred(i, j)) cand.emplace_back(uint32_t(i), uint32_t(j));
    } else {
        auto key = [](long long a, long long b, long long c) {
            return (uint64_t(a) & 0x1fffff) | ((uint64_t(b) & 0x1fffff) << 21) | ((uint64_t(c) & 0x1fffff) << 42);
        };
        std::unordered_map<uint64_t, std::vector<uint32_t>> cells;
        cells.reserve(n);
        std::vector<long long> cc(3 * n);
        for (size_t i = 0; i < n; i++) {
            for (int c = 0; c < 3; c++) cc[3 * i + c] = (long long)((x0[3 * i + c] - lo[c]) / cs);
            cells[key(cc[3 * i], cc[3 * i + 1], cc[3 * i + 2])].push_back(uint32_t(i));
        }
        for (size_t i = 0; i < n; i++) {
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        auto it = cells.find(key(cc[3 * i] + dx, cc[3 * i + 1] + dy, cc[3 * i + 2] + dz));
                        if (it == cells.end()) continue;
                        for (uint32_t j : it->second)
                            if (j > i && pred(i, j)) cand.emplace_back(uint32_t(i), j);
                    }
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    }
    if (!prevent_cycles) {  // alt_breaking_sinter_bridge.h:63-75: every candidate
        bonds = std::move(cand);
        return;
    }
    uf comps(n);
    for (auto &c : cand) {
        uint32_t ri = comps.find(c.first), rj = comps.find(c.second);
        if (ri == rj) {
            skipped++;
            continue;
        }
        bonds.push_back(c);
        comps.parent[rj] = ri;
    }
}

}  // namespace hg

// The legacy sinter_functor constructor's connector frames
// (sinter_bridge.h:146-211): for each bond (i, j) in row order, with
// n_i = (x_j - x_i)/|x_j - x_i|, unit vectors v are drawn until v x (-n_i) is
// not zero; ds is v without its n_i component and db = v x (-n_i), both
// normalised; i holds x_i +- r ds, x_i +- r db, x_i + r n_i and j holds
// x_j +- r ds, x_j +- r db, x_j - r n_i.
extern "C" hg_status hg_legacy_sinter_frames(size_t n, const double *x0, double r_part, double critical_separation,
                                             hg_unit_vector_fn next, void *user, uint32_t *pairs,
                                             double *connectors, size_t cap, size_t *nbonds) {
    using namespace hg;
    if (!nbonds) return fail(HG_ERR_INVALID, "hg_legacy_sinter_frames: nbonds is required");
    if (n && !x0) return fail(HG_ERR_INVALID, "hg_legacy_sinter_frames: x0 is required");
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped = 0;
    sinter_topology_host(n, x0, r_part, critical_separation, bonds, skipped, false, true);
    *nbonds = bonds.size();
    if (!pairs || !connectors || cap < bonds.size()) return HG_OK;
    if (!next && !bonds.empty()) return fail(HG_ERR_INVALID, "hg_legacy_sinter_frames: a unit-vector source is required");
    constexpr long kMaxDraws = 100000;  // per bond: a source that only returns collinear vectors ends here
    for (size_t b = 0; b < bonds.size(); b++) {
        const uint32_t i = bonds[b].first, j = bonds[b].second;
        const d3 pi{x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]}, pj{x0[3 * j], x0[3 * j + 1], x0[3 * j + 2]};
        const d3 n_i = normalized(pj - pi);
        const d3 n_j = -n_i;
        d3 v, db;
        long draws = 0;
        do {
            if (++draws > kMaxDraws)
                return fail(HG_ERR_INVALID, "hg_legacy_sinter_frames: no unit vector off the normal of bond " +
                                                std::to_string(b) + " in 100,000 draws");
            double o[3] = {0.0, 0.0, 0.0};
            next(user, o);
            v = d3{o[0], o[1], o[2]};
            db = cross(v, n_j);
        } while (norm(db) == 0.0);
        d3 ds = v - dot(v, n_i) * n_i;
        db = normalized(db);
        ds = normalized(ds);
        const d3 pts[10] = {pi + ds * r_part, pi - ds * r_part, pi + db * r_part, pi - db * r_part, pi + n_i * r_part,
                            pj + ds * r_part, pj - ds * r_part, pj + db * r_part, pj - db * r_part, pj + n_j * r_part};
        pairs[2 * b] = i;
        pairs[2 * b + 1] = j;
        for (int q = 0; q < 10; q++) {
            connectors[30 * b + 3 * q] = pts[q].x;
            connectors[30 * b + 3 * q + 1] = pts[q].y;
            connectors[30 * b + 3 * q + 2] = pts[q].z;
        }
    }
    return HG_OK;
}
