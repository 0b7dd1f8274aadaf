// Seeded tunable-dimension cluster-cluster aggregation (TD-CCA) of
// monodisperse spheres: the input generator of the aggregate workloads. It has
// no counterpart in the reference, which ships one fixed 150-particle
// aggregate (resources/aggregate.h).
//
// Method (after Filippov, Zurita & Rosner, J. Colloid Interface Sci. 229,
// 261-273, 2000): clusters merge pairwise, level by level. For clusters of N1
// and N2 spheres with radii of gyration Rg1, Rg2, the centre-of-mass distance
// that keeps N = kf (Rg/r)^Df for the union is
//     G^2 = N^2 r^2 / (N1 N2) (N/kf)^(2/Df) - N Rg1^2 / N2 - N Rg2^2 / N1.
// Each trial draws a uniform random rotation of cluster 2 and a uniform
// direction u, and slides cluster 2 in from infinity along u until its first
// contact with cluster 1 (so the union touches and does not overlap). The
// trial whose contact distance is closest to G is kept. Random numbers come
// from mt19937_64 with explicit transforms, so a seed gives the same aggregate
// on every platform.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hg_internal.h"

namespace hg {
namespace {

struct rng {
    std::mt19937_64 g;
    explicit rng(uint64_t s) : g(s) {}
    double uni() { return double(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint64_t below(uint64_t m) { return g() % m; }
};

struct cluster {
    std::vector<d3> p;  // centred on its centre of mass
    double rg2 = 0.0;
};

void recentre(cluster &c) {
    d3 cm = zero3();
    for (auto &q : c.p) cm = cm + q;
    cm = cm / double(c.p.size());
    double s = 0.0;
    for (auto &q : c.p) {
        q = q - cm;
        s += dot(q, q);
    }
    c.rg2 = s / double(c.p.size());
}

// Largest s at which some pair (a in A, b + s*u in B) is at distance 2r, or
// -inf when no pair can touch along u. Uses a 2-D grid on the plane normal to u.
double first_contact(const std::vector<d3> &A, const std::vector<d3> &B, d3 u, double r) {
    d3 e1 = std::fabs(u.x) < 0.9 ? cross(u, d3{1, 0, 0}) : cross(u, d3{0, 1, 0});
    e1 = normalized(e1);
    d3 e2 = cross(u, e1);
    const double cs = 2.0 * r, four_r2 = 4.0 * r * r;
    auto key = [](long long a, long long b) { return (uint64_t(a) << 32) ^ (uint64_t(b) & 0xffffffffull); };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    grid.reserve(A.size() * 2);
    for (uint32_t k = 0; k < A.size(); k++)
        grid[key((long long)std::floor(dot(A[k], e1) / cs), (long long)std::floor(dot(A[k], e2) / cs))].push_back(k);
    double best = -INFINITY;
    for (const d3 &b : B) {
        long long gx = (long long)std::floor(dot(b, e1) / cs), gy = (long long)std::floor(dot(b, e2) / cs);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++) {
                auto it = grid.find(key(gx + dx, gy + dy));
                if (it == grid.end()) continue;
                for (uint32_t k : it->second) {
                    d3 q = A[k] - b;
                    double qu = dot(q, u);
                    double perp2 = dot(q, q) - qu * qu;
                    if (perp2 > four_r2) continue;
                    double s = qu + std::sqrt(four_r2 - perp2);
                    best = std::max(best, s);
                }
            }
    }
    return best;
}

d3 rotate(const double m[9], d3 a) {
    return {m[0] * a.x + m[1] * a.y + m[2] * a.z, m[3] * a.x + m[4] * a.y + m[5] * a.z,
            m[6] * a.x + m[7] * a.y + m[8] * a.z};
}

void random_rotation(rng &R, double m[9]) {
    double q0 = R.normal(), q1 = R.normal(), q2 = R.normal(), q3 = R.normal();
    double s = std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    q0 /= s; q1 /= s; q2 /= s; q3 /= s;
    m[0] = 1 - 2 * (q2 * q2 + q3 * q3); m[1] = 2 * (q1 * q2 - q0 * q3); m[2] = 2 * (q1 * q3 + q0 * q2);
    m[3] = 2 * (q1 * q2 + q0 * q3); m[4] = 1 - 2 * (q1 * q1 + q3 * q3); m[5] = 2 * (q2 * q3 - q0 * q1);
    m[6] = 2 * (q1 * q3 - q0 * q2); m[7] = 2 * (q2 * q3 + q0 * q1); m[8] = 1 - 2 * (q1 * q1 + q2 * q2);
}

d3 random_direction(rng &R) {
    d3 u{R.normal(), R.normal(), R.normal()};
    return normalized(u);
}

cluster merge(const cluster &A, const cluster &B, double r, double df, double kf, rng &R, int trials) {
    const double N1 = double(A.p.size()), N2 = double(B.p.size()), N = N1 + N2;
    double g2 = N * N * r * r / (N1 * N2) * std::pow(N / kf, 2.0 / df) - N * A.rg2 / N2 - N * B.rg2 / N1;
    const double G = g2 > 0.0 ? std::sqrt(g2) : 0.0;
    double best_err = INFINITY, best_s = 0.0;
    double best_m[9];
    d3 best_u{1, 0, 0};
    std::vector<d3> Bt(B.p.size());
    for (int t = 0; t < trials; t++) {
        double m[9];
        random_rotation(R, m);
        d3 u = random_direction(R);
        for (size_t k = 0; k < B.p.size(); k++) Bt[k] = rotate(m, B.p[k]);
        double s = first_contact(A.p, Bt, u, r);
        if (!std::isfinite(s)) continue;
        double err = std::fabs(s - G);
        if (err < best_err) {
            best_err = err;
            best_s = s;
            best_u = u;
            std::copy(m, m + 9, best_m);
        }
        if (err < 0.02 * G + 1e-3 * r) break;
    }
    cluster C;
    C.p = A.p;
    C.p.reserve(A.p.size() + B.p.size());
    for (const d3 &b : B.p) C.p.push_back(rotate(best_m, b) + best_s * best_u);
    recentre(C);
    return C;
}

}  // namespace
}  // namespace hg

namespace hg {
namespace {
void generate_one(size_t n, double r_part, double df, double kf, uint64_t seed, double *x_out);
}
}  // namespace hg

extern "C" hg_status hg_generate_aggregate(size_t n, double r_part, double df, double kf, uint64_t seed,
                                           double *x_out) {
    using namespace hg;
    if (n == 0 || x_out == nullptr || !(r_part > 0.0) || !(df > 0.0) || !(kf > 0.0))
        return fail(HG_ERR_INVALID, "hg_generate_aggregate: need n > 0, r_part > 0, df > 0, kf > 0 and x_out");
    generate_one(n, r_part, df, kf, seed, x_out);
    return HG_OK;
}

extern "C" hg_status hg_generate_aggregates(size_t count, const uint32_t *n, const double *df, double kf,
                                            const uint64_t *seeds, double r_part, double *x_out, int threads) {
    using namespace hg;
    if (count == 0) return HG_OK;
    if (!n || !df || !seeds || !x_out || !(r_part > 0.0) || !(kf > 0.0))
        return fail(HG_ERR_INVALID, "hg_generate_aggregates: invalid arguments");
    std::vector<size_t> off(count + 1, 0);
    for (size_t k = 0; k < count; k++) {
        if (n[k] == 0 || !(df[k] > 0.0)) return fail(HG_ERR_INVALID, "hg_generate_aggregates: empty aggregate or df <= 0");
        off[k + 1] = off[k] + n[k];
    }
    unsigned nt = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&] {
            for (size_t k; (k = next.fetch_add(1)) < count;)
                generate_one(n[k], r_part, df[k], kf, seeds[k], x_out + 3 * off[k]);
        });
    for (auto &th : pool) th.join();
    return HG_OK;
}

namespace hg {
namespace {
void generate_one(size_t n, double r_part, double df, double kf, uint64_t seed, double *x_out) {
    rng R(seed);
    std::vector<cluster> cl(n);
    for (auto &c : cl) c.p.assign(1, zero3());
    while (cl.size() > 1) {
        for (size_t k = cl.size() - 1; k > 0; k--) std::swap(cl[k], cl[R.below(k + 1)]);
        std::vector<cluster> next;
        next.reserve(cl.size() / 2 + 1);
        for (size_t k = 0; k + 1 < cl.size(); k += 2) next.push_back(merge(cl[k], cl[k + 1], r_part, df, kf, R, 48));
        if (cl.size() % 2) next.push_back(std::move(cl.back()));
        cl.swap(next);
    }
    for (size_t i = 0; i < n; i++) {
        x_out[3 * i] = cl[0].p[i].x;
        x_out[3 * i + 1] = cl[0].p[i].y;
        x_out[3 * i + 2] = cl[0].p[i].z;
    }
}
}  // namespace
}  // namespace hg
