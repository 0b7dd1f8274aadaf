// CPU oracle for the libgran DEM per-timestep path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
// check and bench.py's cpu_baseline leg may compile against it, link it or run
// it. The product (libgran_amd/) never includes or calls anything here.
//
// It restates, in FP64 and in the reference's arithmetic order, the functors of
// egor-demidov/libgran and the documented contract of its time-stepping
// dependency (libtimestep, absent from the reference tree):
//
//   vec3 ops                 Eigen::Vector3d semantics as used by the functors
//                            (dot/norm/normalized/cross; normalized() of a zero
//                            vector returns it unchanged, as Eigen 3.4 does)
//   contact_eval             include/libgran/contact_force/contact_force.h:57-96
//   shear_contact            include/libgran/contact_force/contact_force.h:105-147
//   sinter_bonded_eval       include/libgran/sinter_bridge/alt_sinter_bridge.h:92-129
//   shear_sinter             include/libgran/sinter_bridge/alt_sinter_bridge.h:137-156
//   sinter_topology          include/libgran/sinter_bridge/alt_sinter_bridge.h:19-90
//   legacy_sinter_frames     include/libgran/sinter_bridge/sinter_bridge.h:138-211
//   legacy_bonded_eval       include/libgran/sinter_bridge/sinter_bridge.h:235-259
//   legacy connector moves   include/libgran/sinter_bridge/sinter_bridge.h:24-75
//   hamaker_eval             include/libgran/hamaker_force/hamaker_force.h:28-43
//   lr_hamaker_eval          include/libgran/hamaker_force/limited_range_hamaker_force.h:29-47
//   simplified_hamaker_eval  include/libgran/hamaker_force/simplified_hamaker_force.h:18-30
//   pair fold                include/libgran/granular_system/binary_force_container.h:20-32
//                            + include/libgran/utility/operator_overloads.h:10-20
//   unary fold               include/libgran/granular_system/unary_force_container.h:20-32
//   facet geometry           include/libgran/surface_force/triangular_facet.h:18-137
//   surface_contact_eval     include/libgran/contact_force/surface_contact_force.h:57-147
//   surface_hamaker_eval     include/libgran/hamaker_force/surface_hamaker_force.h:18-30
//   system / integrator      granular_system.h:53-81, granular_system_neighbor_list.h:54-82
//                            and the libtimestep contract in
//                            Writerside/topics/Custom-step-handler.md:82-145,
//                            Writerside/topics/Granular-system.md:30-244
//   diagnostics              test/compute_energy.cpp:7-58, test/mass_distribution.cpp:7-27
//
// Decisions where the reference source is absent (libtimestep), fixed here and
// pinned by closed-form tests rather than by the reference:
//   * acceleration pass: for each i, acc = 0; for j ascending (j != i, all j in
//     all-pairs mode, listed j in neighbour mode): acc += fold(i, j);
//     then acc = acc + unary(i) when facets are present.
//   * integrator (rotational velocity Verlet, kick-drift-kick):
//       v += a*(dt/2); w += alpha*(dt/2); x += v*dt; theta += w*dt; t += dt;
//       recompute (a, alpha) at the new state; v += a*(dt/2); w += alpha*(dt/2).
//     a(t0) is computed lazily by the first step, with whatever neighbour list
//     exists at that moment.
//   * neighbour list: j != i with norm(x_i - x_j) < r_verlet, ascending j.
//   * sinter_step_handler: within one particle's drift, increment_x (the
//     connectors follow dx) comes before increment_theta (they turn about the
//     new centre); the two orders are the same rotation and differ only in
//     rounding.
//   * the alt_sinter bond predicate uses fabs (the author's intent); the
//     reference's unqualified abs() can bind to abs(int) on some toolchains.
//   * spring tables are sparse rows keyed by j; an absent entry is the zero
//     spring, which reproduces the dense n*n tables exactly.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- vec3 ------
struct v3 {
    double x, y, z;
};

inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
inline v3 operator*(double s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline v3 operator*(v3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator/(v3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double norm(v3 a) { return std::sqrt(dot(a, a)); }
inline v3 normalized(v3 a) {
    double z = dot(a, a);
    if (z > 0.0) return a / std::sqrt(z);
    return a;
}
inline v3 cross(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
constexpr v3 vzero{0.0, 0.0, 0.0};

struct acc2 {
    v3 lin, ang;
};
inline acc2 operator+(acc2 a, acc2 b) { return {a.lin + b.lin, a.ang + b.ang}; }

template <typename T> inline T square(T x) { return x * x; }
template <typename T> inline T cube(T x) { return x * x * x; }

// ------------------------------------------------------------- params -------
struct contact_params {
    double k, gamma_n, k_t, gamma_t, mu_s, phi_d, k_r, gamma_r, mu_r, phi_r,
        k_o, gamma_o, mu_o, phi_o, r_part, mass, inertia, dt;
};
struct sinter_params {
    double k, gamma_n, k_t, gamma_t, k_r, gamma_r, k_o, gamma_o,
        r_part, mass, inertia, dt, d_crit;
};
struct hamaker_params {
    double A, h0, r_cutoff, r_part, mass;
};

enum { BIN_NONE = 0, BIN_CONTACT = 1, BIN_SINTER = 2, BIN_SINTER_LEGACY = 3 };
enum { HAM_NONE = 0, HAM_FULL = 1, HAM_LIMITED = 2, HAM_SIMPLIFIED = 3 };

struct spring3 {
    v3 s[3];
};

// Sparse spring rows: row i holds (j, spring) sorted by j.
struct spring_rows {
    std::vector<std::vector<std::pair<uint32_t, spring3>>> rows;
    void resize(size_t n) { rows.assign(n, {}); }
    spring3 &at(size_t i, uint32_t j) {
        auto &r = rows[i];
        auto it = std::lower_bound(r.begin(), r.end(), j,
                                   [](const std::pair<uint32_t, spring3> &e, uint32_t key) { return e.first < key; });
        if (it == r.end() || it->first != j) it = r.insert(it, {j, spring3{{vzero, vzero, vzero}}});
        return it->second;
    }
    void reset(size_t i, uint32_t j) {
        auto &r = rows[i];
        auto it = std::lower_bound(r.begin(), r.end(), j,
                                   [](const std::pair<uint32_t, spring3> &e, uint32_t key) { return e.first < key; });
        if (it != r.end() && it->first == j) r.erase(it);
    }
};

// ------------------------------------------------------- contact model ------
// contact_force.h:105-147 (Luding slip model on a persistent spring).
inline v3 shear_contact(v3 &xi, v3 n, double stiffness, double damping, double normal_force,
                        double mu_static, double phi_dynamic, v3 rel_v, double dt) {
    double mu_dynamic = mu_static * phi_dynamic;
    v3 xi_new = xi - dot(xi, n) * n;
    if (norm(xi_new) > 0.0) {
        xi_new = xi_new / std::sqrt(dot(xi_new, xi_new));
        xi_new = xi_new * norm(xi);
    }
    xi = xi_new;
    v3 f0 = (-stiffness) * xi - damping * rel_v;
    double static_friction = mu_static * normal_force;
    double dynamic_friction = mu_dynamic * normal_force;
    v3 t = norm(f0) > 0.0 ? normalized(f0) : vzero;
    double f_selected;
    if (norm(f0) <= static_friction) {
        xi = xi + rel_v * dt;
        f_selected = static_friction;
    } else {
        xi = (-1.0 / stiffness) * (dynamic_friction * t + damping * rel_v);
        f_selected = dynamic_friction;
    }
    return std::min(f_selected, norm(f0)) * t;
}

// Dissipated work of one ordered pair over one evaluation (the builder's
// ledger term, not in the reference): the pair's forces take work
// -f_n v_n + f_t.v_t + (r/rpp) f_r.v_r + f_o.v_o per unit time from the two
// particles' motion (derived from the torques the functors apply: the rolling
// torque uses r, the rolling velocity rpp). gamma_n v_n^2 of it is lost, and
// of each shear term the work minus the change of its spring energy
// k/2 |xi|^2. An ordered pair holds half of its pair.
inline void pair_dissipation(double *d4, double gamma_n, double v_n, double dt, double w_r, const double k[3],
                             const v3 f[3], const v3 v[3], const double u_old[3], const spring3 &sp) {
    d4[0] = d4[0] + 0.5 * gamma_n * v_n * v_n * dt;
    for (int c = 0; c < 3; c++) {
        const double w = c == 1 ? w_r : 1.0;
        d4[1 + c] = d4[1 + c] + (-0.5 * w * dot(f[c], v[c]) * dt - 0.25 * k[c] * (dot(sp.s[c], sp.s[c]) - u_old[c]));
    }
}

// contact_force.h:57-96. `sp` is the (i,j) spring triple; `reset` is set when the
// reference would reset it (overlap <= 0). With d4, the pair's dissipated work
// (pair_dissipation; a reset spring's energy is lost whole) is added to it.
inline acc2 contact_eval(const contact_params &p, v3 xi_, v3 xj, v3 vi, v3 vj, v3 wi, v3 wj,
                         spring3 &sp, bool &reset, double *d4 = nullptr) {
    v3 n = normalized(xi_ - xj);
    double overlap = 2.0 * p.r_part - dot(xi_ - xj, n);
    if (overlap <= 0) {
        if (d4) {
            const double k[3] = {p.k_t, p.k_r, p.k_o};
            for (int c = 0; c < 3; c++) d4[1 + c] = d4[1 + c] + 0.25 * k[c] * dot(sp.s[c], sp.s[c]);
        }
        reset = true;
        sp = spring3{{vzero, vzero, vzero}};
        return {vzero, vzero};
    }
    reset = false;
    double rpp = p.r_part - overlap / 2.0;
    double v_n = -dot(vi - vj, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    v3 v_ij = ((vi - vj) + rpp * cross(n, wi)) + rpp * cross(n, wj);
    v3 v_t = v_ij - dot(v_ij, n) * n;
    v3 v_r = (-rpp) * (cross(n, wi) - cross(n, wj));
    v3 v_o = (p.r_part * (dot(n, wi) - dot(n, wj))) * n;
    const double u_old[3] = {dot(sp.s[0], sp.s[0]), dot(sp.s[1], sp.s[1]), dot(sp.s[2], sp.s[2])};
    v3 f_t = shear_contact(sp.s[0], n, p.k_t, p.gamma_t, f_n, p.mu_s, p.phi_d, v_t, p.dt);
    v3 f_r = shear_contact(sp.s[1], n, p.k_r, p.gamma_r, f_n, p.mu_r, p.phi_r, v_r, p.dt);
    v3 f_o = shear_contact(sp.s[2], n, p.k_o, p.gamma_o, f_n, p.mu_o, p.phi_o, v_o, p.dt);
    if (d4) {
        const double k[3] = {p.k_t, p.k_r, p.k_o};
        const v3 f[3] = {f_t, f_r, f_o}, vv[3] = {v_t, v_r, v_o};
        pair_dissipation(d4, p.gamma_n, v_n, p.dt, p.r_part / rpp, k, f, vv, u_old, sp);
    }
    v3 tau_t = rpp * cross(n, f_t);
    v3 tau_r = p.r_part * cross(n, f_r);
    v3 tau_o = p.r_part * f_o;
    return {(f_n * n + f_t) / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}

// ----------------------------------------------------- bonded (sinter) ------
// alt_sinter_bridge.h:137-156
inline v3 shear_sinter(v3 &xi, v3 n, double stiffness, double damping, v3 rel_v, double dt) {
    v3 xi_new = xi - dot(xi, n) * n;
    if (norm(xi_new) > 0.0) {
        xi_new = xi_new / std::sqrt(dot(xi_new, xi_new));
        xi_new = xi_new * norm(xi);
    }
    xi = xi_new;
    v3 f0 = (-stiffness) * xi - damping * rel_v;
    xi = xi + rel_v * dt;
    return f0;
}

// alt_sinter_bridge.h:92-129, bonded branch (no overlap gate).
// alt_breaking_sinter_bridge.h:85-125 evaluates the same expressions and also
// needs the force before the division by the mass: *ftot receives f_n n + f_t.
inline acc2 sinter_bonded_eval(const sinter_params &p, v3 xi_, v3 xj, v3 vi, v3 vj, v3 wi, v3 wj,
                               spring3 &sp, v3 *ftot = nullptr, double *d4 = nullptr) {
    v3 n = normalized(xi_ - xj);
    double overlap = 2.0 * p.r_part - dot(xi_ - xj, n);
    double rpp = p.r_part - overlap / 2.0;
    double v_n = -dot(vi - vj, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    v3 v_ij = ((vi - vj) + rpp * cross(n, wi)) + rpp * cross(n, wj);
    v3 v_t = v_ij - dot(v_ij, n) * n;
    v3 v_r = (-rpp) * (cross(n, wi) - cross(n, wj));
    v3 v_o = (p.r_part * (dot(n, wi) - dot(n, wj))) * n;
    const double u_old[3] = {dot(sp.s[0], sp.s[0]), dot(sp.s[1], sp.s[1]), dot(sp.s[2], sp.s[2])};
    v3 f_t = shear_sinter(sp.s[0], n, p.k_t, p.gamma_t, v_t, p.dt);
    v3 f_r = shear_sinter(sp.s[1], n, p.k_r, p.gamma_r, v_r, p.dt);
    v3 f_o = shear_sinter(sp.s[2], n, p.k_o, p.gamma_o, v_o, p.dt);
    if (d4) {
        const double k[3] = {p.k_t, p.k_r, p.k_o};
        const v3 f[3] = {f_t, f_r, f_o}, vv[3] = {v_t, v_r, v_o};
        pair_dissipation(d4, p.gamma_n, v_n, p.dt, p.r_part / rpp, k, f, vv, u_old, sp);
    }
    v3 tau_t = rpp * cross(n, f_t);
    v3 tau_r = p.r_part * cross(n, f_r);
    v3 tau_o = p.r_part * f_o;
    const v3 f_tot = f_n * n + f_t;
    if (ftot) *ftot = f_tot;
    return {f_tot / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}

// alt_sinter_bridge.h:56-90: lexicographic scan over i<j; a pair whose gap is
// within d_crit is bonded unless both ends already share a component (cycle),
// which is skipped (the reference prints a warning). Components merge by
// relabelling every member of j's subset with i's label, as the reference does.
// With prevent_cycles = false it is the bond search of
// alt_breaking_sinter_functor (alt_breaking_sinter_bridge.h:63-75): every pair
// within d_crit is bonded.
inline void sinter_topology(size_t n, const std::vector<v3> &x0, double r_part, double d_crit,
                            std::vector<std::pair<uint32_t, uint32_t>> &bonds, size_t &skipped,
                            bool prevent_cycles = true) {
    std::vector<size_t> subset(n);
    for (size_t i = 0; i < n; i++) subset[i] = i;
    bonds.clear();
    skipped = 0;
    if (n == 0) return;
    for (size_t i = 0; i + 1 < n; i++) {
        for (size_t j = i + 1; j < n; j++) {
            if (std::fabs(norm(x0[i] - x0[j]) - 2.0 * r_part) < d_crit) {
                if (prevent_cycles && subset[i] == subset[j]) {
                    skipped++;
                    continue;
                }
                bonds.emplace_back(uint32_t(i), uint32_t(j));
                size_t sj = subset[j];
                for (auto &s : subset)
                    if (s == sj) s = subset[i];
            }
        }
    }
}

// --------------------------------------------------------------- hamaker ----
// ------------------------------------------------- legacy sinter_functor --
// Spring connectors of one bond: the five points attached to each of its two
// particles, moved with that particle by the step handler.
struct legacy_bond {
    uint32_t i, j;
    v3 ci[5], cj[5];
};

// The constructor's bond search and connector frames
// (sinter_bridge.h:146-211): pairs i < j in row order whose gap
// |x_j - x_i| - 2r has magnitude at most critical_separation (fabs, as for
// alt_sinter), each given a frame (ds, db, n) from the caller's random unit
// vectors, drawn until one is not collinear with the normal.
template <class NextVec>
inline void legacy_sinter_frames(size_t n, const std::vector<v3> &x0, double r_part, double d_crit, NextVec next,
                                 std::vector<legacy_bond> &out) {
    out.clear();
    for (size_t i = 0; i + 1 < n; i++) {
        for (size_t j = i + 1; j < n; j++) {
            const v3 pi = x0[i], pj = x0[j];
            double separation = norm(pj - pi) - 2.0 * r_part;
            if (std::fabs(separation) > d_crit) continue;
            v3 n_i = normalized(pj - pi);
            v3 n_j = -n_i;
            v3 v, db;
            do {
                v = next();
                db = cross(v, n_j);
            } while (norm(db) == 0.0);
            v3 ds = v - dot(v, n_i) * n_i;
            db = normalized(db);
            ds = normalized(ds);
            legacy_bond b;
            b.i = uint32_t(i);
            b.j = uint32_t(j);
            b.ci[0] = pi + ds * r_part;
            b.ci[1] = pi - ds * r_part;
            b.ci[2] = pi + db * r_part;
            b.ci[3] = pi - db * r_part;
            b.ci[4] = pi + n_i * r_part;
            b.cj[0] = pj + ds * r_part;
            b.cj[1] = pj - ds * r_part;
            b.cj[2] = pj + db * r_part;
            b.cj[3] = pj - db * r_part;
            b.cj[4] = pj + n_j * r_part;
            out.push_back(b);
        }
    }
}

// Bonded pair (i, j) of sinter_functor (sinter_bridge.h:235-259): five
// springs between matching connectors, rest length 2r for the four side
// springs and 0 for the one at the contact point; ci are i's connectors of
// the bond, cj j's.
inline acc2 legacy_bonded_eval(double k_d, double gamma_d, double r_part, double mass, double inertia,
                               const v3 *ci, const v3 *cj, v3 xi_, v3 xj, v3 vi, v3 vj, v3 wi, v3 wj) {
    v3 force = vzero, torque = vzero;
    for (int q = 0; q < 5; q++) {
        v3 spring = cj[q] - ci[q];
        double spring_length = norm(spring);
        double equilibrium_length = q < 4 ? 2.0 * r_part : 0.0;
        v3 conn_i_v = vi + cross(wi, ci[q] - xi_);
        v3 conn_j_v = vj + cross(wj, cj[q] - xj);
        v3 rel_v = conn_j_v - conn_i_v;
        v3 dforce = -(equilibrium_length - spring_length) * k_d * normalized(spring) + gamma_d * rel_v;
        force = force + dforce;
        v3 arm = ci[q] - xi_;
        torque = torque + cross(arm, dforce);
    }
    return {force / mass, torque / inertia};
}

// sinter_step_handler (sinter_bridge.h:24-42, 55-75): a particle's connectors
// follow its position increment dx, then turn with its orientation increment
// dtheta about its (new) centre.
inline void legacy_move_x(v3 *c, v3 dx) {
    for (int q = 0; q < 5; q++) c[q] = c[q] + dx;
}
inline void legacy_move_theta(v3 *c, v3 x, v3 dtheta) {
    for (int q = 0; q < 5; q++) {
        v3 spring_vector = c[q] - x;
        c[q] = c[q] + cross(dtheta, spring_vector);
    }
}

inline double hamaker_scalar(double A, double rp, double h) {
    return (-A) / 6.0 *
           ((((4.0 * rp + 2.0 * h) / (4.0 * rp + h) / h - 2.0 / (2.0 * rp + h)) -
             4.0 * square(rp) / cube(2.0 * rp + h)) -
            2.0 * square(rp) * (4.0 * rp + 2.0 * h) / square(4.0 * rp + h) / square(h));
}

// hamaker_force.h:28-43
inline acc2 hamaker_eval(const hamaker_params &p, v3 xi_, v3 xj) {
    v3 r = xj - xi_;
    double h = norm(r) - 2.0 * p.r_part;
    if (h < p.h0) h = p.h0;
    v3 f = hamaker_scalar(p.A, p.r_part, h) * normalized(r);
    return {f / p.mass, vzero};
}

// limited_range_hamaker_force.h:29-47 (strict '>' on the centre distance)
inline acc2 lr_hamaker_eval(const hamaker_params &p, v3 xi_, v3 xj) {
    if (norm(xi_ - xj) > p.r_cutoff) return {vzero, vzero};
    return hamaker_eval(p, xi_, xj);
}

// simplified_hamaker_force.h:18-30: 2 A r / h^2 along x_j - x_i, h clamped at h0
inline acc2 simplified_hamaker_eval(const hamaker_params &p, v3 xi_, v3 xj) {
    v3 r = xj - xi_;
    double h = norm(r) - 2.0 * p.r_part;
    if (h < p.h0) h = p.h0;
    v3 f = 2.0 * p.A * p.r_part / square(h) * normalized(r);
    return {f / p.mass, vzero};
}

// Energy ledger, Hamaker part (the reference defines no potential energy).
// README.md:154-157 gives U(d) for two spheres at surface gap d; its
// derivative dU/dd is hamaker_scalar. Below h0 the functors apply the force at
// h0, so U continues linearly; the limited-range potential is shifted to zero
// at r_cutoff. The simplified functor's potential is -2 A r / d.
inline double hamaker_u(int model, const hamaker_params &p, double d) {
    if (model == HAM_SIMPLIFIED) return -2.0 * p.A * p.r_part / d;
    const double r = p.r_part;
    return (-p.A) / 6.0 *
           (2.0 * square(r) / ((4.0 * r + d) * d) + 2.0 * square(r) / square(2.0 * r + d) +
            std::log((4.0 * r + d) * d / square(2.0 * r + d)));
}
inline double hamaker_dudd(int model, const hamaker_params &p, double d) {
    if (model == HAM_SIMPLIFIED) return 2.0 * p.A * p.r_part / square(d);
    return hamaker_scalar(p.A, p.r_part, d);
}
inline double hamaker_u_sat(int model, const hamaker_params &p, double d) {
    return d < p.h0 ? hamaker_u(model, p, p.h0) + hamaker_dudd(model, p, p.h0) * (d - p.h0) : hamaker_u(model, p, d);
}
inline double hamaker_potential(int model, const hamaker_params &p, double dist) {
    const double d = dist - 2.0 * p.r_part;
    if (model == HAM_LIMITED) {
        if (dist > p.r_cutoff) return 0.0;
        return hamaker_u_sat(model, p, d) - hamaker_u_sat(model, p, p.r_cutoff - 2.0 * p.r_part);
    }
    return hamaker_u_sat(model, p, d);
}

// the second functor of the binary container, by model
inline acc2 binary_hamaker_eval(int model, const hamaker_params &p, v3 xi_, v3 xj) {
    if (model == HAM_FULL) return hamaker_eval(p, xi_, xj);
    if (model == HAM_SIMPLIFIED) return simplified_hamaker_eval(p, xi_, xj);
    return lr_hamaker_eval(p, xi_, xj);
}

// ------------------------------------------------------------- facets -------
struct facet {
    v3 v_facet;
    v3 vert[3];
    double area;
    v3 normal;
};

// triangular_facet.h:18-30
inline facet make_facet(v3 v_facet, v3 a, v3 b, v3 c) {
    facet f;
    f.v_facet = v_facet;
    f.vert[0] = a;
    f.vert[1] = b;
    f.vert[2] = c;
    v3 u = b - a, w = c - a;
    f.area = 1.0 / 2.0 * norm(cross(u, w));
    f.normal = normalized(cross(u, w));
    return f;
}

// triangular_facet.h:73-137
inline v3 closest_point_on_triangle(const facet &f, v3 p) {
    v3 offset = dot(f.vert[0], f.normal) * f.normal;
    v3 proj = p - dot(p, f.normal) * f.normal;
    v3 cp = proj + offset;
    {
        double area = 0.0;
        v3 pa = f.vert[0] - cp, pb = f.vert[1] - cp, pc = f.vert[2] - cp;
        area += 1.0 / 2.0 * norm(cross(pa, pb));
        area += 1.0 / 2.0 * norm(cross(pb, pc));
        area += 1.0 / 2.0 * norm(cross(pc, pa));
        if (!(area > f.area)) return cp;
    }
    auto edge = [&](int a, int b) -> v3 {
        v3 lv = f.vert[b] - f.vert[a];
        v3 ulv = normalized(lv);
        v3 tpv = p - f.vert[a];
        double len = norm(lv);
        double pr = dot(tpv, ulv);
        if (pr < 0.0) return f.vert[a];
        if (pr > len) return f.vert[b];
        return f.vert[a] + pr * ulv;
    };
    v3 e1 = edge(0, 1), e2 = edge(1, 2), e3 = edge(2, 0);
    v3 d1 = p - e1, d2 = p - e2, d3 = p - e3;
    double s1 = dot(d1, d1), s2 = dot(d2, d2), s3 = dot(d3, d3);
    if (s1 < s2) {
        if (s1 < s3) return e1;
    } else {
        if (s2 < s3) return e2;
    }
    return e3;
}

// surface_contact_force.h:57-96 (shear part identical to contact_force.h:105-147)
inline acc2 surface_contact_eval(const contact_params &p, v3 xf, v3 vf, v3 xi_, v3 vi, v3 wi,
                                 spring3 &sp) {
    v3 n = normalized(xi_ - xf);
    double overlap = p.r_part - dot(xi_ - xf, n);
    if (overlap <= 0) {
        sp = spring3{{vzero, vzero, vzero}};
        return {vzero, vzero};
    }
    double rpp = p.r_part - overlap;
    double v_n = -dot(vi - vf, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    v3 v_ij = (vi - vf) + rpp * cross(n, wi);
    v3 v_t = v_ij - dot(v_ij, n) * n;
    v3 v_r = (-rpp) * cross(n, wi);
    v3 v_o = (p.r_part * dot(n, wi)) * n;
    v3 f_t = shear_contact(sp.s[0], n, p.k_t, p.gamma_t, f_n, p.mu_s, p.phi_d, v_t, p.dt);
    v3 f_r = shear_contact(sp.s[1], n, p.k_r, p.gamma_r, f_n, p.mu_r, p.phi_r, v_r, p.dt);
    v3 f_o = shear_contact(sp.s[2], n, p.k_o, p.gamma_o, f_n, p.mu_o, p.phi_o, v_o, p.dt);
    v3 tau_t = rpp * cross(n, f_t);
    v3 tau_r = p.r_part * cross(n, f_r);
    v3 tau_o = p.r_part * f_o;
    return {(f_n * n + f_t) / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}

// surface_hamaker_force.h:18-30
inline acc2 surface_hamaker_eval(const hamaker_params &p, v3 xf, v3 xi_) {
    v3 r = xf - xi_;
    double h = norm(r) - p.r_part;
    if (h < p.h0) h = p.h0;
    v3 f = 2.0 * p.A / 3.0 * (cube(p.r_part) / square(h) / square(h + 2.0 * p.r_part)) * normalized(r);
    return {f / p.mass, vzero};
}

// ------------------------------------------------------------- system -------
struct system_params {
    int binary_model;   // BIN_*
    int hamaker_model;  // HAM_*
    int neighbor_mode;  // 0: all pairs (granular_system), 1: list (granular_system_neighbor_list)
    int facet_contact;  // facets carry a surface contact functor
    int facet_hamaker;  // facets carry a surface hamaker functor
    int sinter_breaking;  // BIN_SINTER as alt_breaking_sinter_functor
    double r_verlet;
    contact_params contact;
    sinter_params sinter;
    hamaker_params hamaker;
    contact_params surface_contact;
    hamaker_params surface_hamaker;
};

class dem_system {
public:
    system_params p;
    size_t n;
    std::vector<v3> x, v, th, w, a, al;
    double t = 0.0;
    bool acc_valid = false;
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t bonds_skipped = 0;
    std::vector<std::vector<uint32_t>> bonded_to;  // sorted per i
    spring_rows contact_springs, sinter_springs;
    std::vector<std::vector<uint32_t>> nbr;
    std::vector<facet> facets;
    std::vector<spring_rows> surface_springs;  // one row set per facet, key 0
    // cumulative dissipated work per particle (normal, tangential, rolling,
    // torsion), accumulated while track_diss is set
    bool track_diss = false;
    std::vector<double> diss;

    dem_system(const system_params &prm, size_t n_, const double *x0, const double *v0,
               const double *th0, const double *w0)
        : p(prm), n(n_) {
        x.resize(n); v.resize(n); th.resize(n); w.resize(n);
        a.assign(n, vzero); al.assign(n, vzero);
        for (size_t i = 0; i < n; i++) {
            x[i] = {x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]};
            v[i] = {v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]};
            th[i] = {th0[3 * i], th0[3 * i + 1], th0[3 * i + 2]};
            w[i] = {w0[3 * i], w0[3 * i + 1], w0[3 * i + 2]};
        }
        contact_springs.resize(n);
        sinter_springs.resize(n);
        bonded_to.assign(n, {});
        if (p.binary_model == BIN_SINTER) {
            sinter_topology(n, x, p.sinter.r_part, p.sinter.d_crit, bonds, bonds_skipped, !p.sinter_breaking);
            for (auto &b : bonds) {
                bonded_to[b.first].push_back(b.second);
                bonded_to[b.second].push_back(b.first);
            }
            for (auto &r : bonded_to) std::sort(r.begin(), r.end());
            neck.assign(n, {});
            for (auto &b : bonds) {
                neck[b.first].push_back({b.second, INFINITY});
                neck[b.second].push_back({b.first, INFINITY});
            }
            for (auto &r : neck) std::sort(r.begin(), r.end());
        }
        if (p.neighbor_mode == 1) rebuild_neighbors(true);
    }

    // sinter_functor (BIN_SINTER_LEGACY): bonds with their connectors, and per
    // particle its (partner, bond index) pairs sorted by partner. Set by
    // set_legacy before the first step.
    std::vector<legacy_bond> legacy;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> legacy_of;
    void set_legacy(const std::vector<legacy_bond> &b) {
        legacy = b;
        legacy_of.assign(n, {});
        bonds.clear();
        for (size_t k = 0; k < b.size(); k++) {
            legacy_of[b[k].i].push_back({b[k].j, uint32_t(k)});
            legacy_of[b[k].j].push_back({b[k].i, uint32_t(k)});
            bonds.push_back({b[k].i, b[k].j});
        }
        for (auto &r : legacy_of) std::sort(r.begin(), r.end());
    }
    // i's connectors of bond k
    v3 *legacy_side(size_t i, uint32_t k) { return legacy[k].i == i ? legacy[k].ci : legacy[k].cj; }
    const v3 *legacy_side(size_t i, uint32_t k) const { return legacy[k].i == i ? legacy[k].ci : legacy[k].cj; }
    int64_t legacy_bond_of(size_t i, uint32_t j) const {
        if (legacy_of.empty()) return -1;
        const auto &r = legacy_of[i];
        auto it = std::lower_bound(r.begin(), r.end(), std::make_pair(j, uint32_t(0)));
        return (it != r.end() && it->first == j) ? int64_t(it->second) : -1;
    }

    bool is_bonded(size_t i, uint32_t j) const {
        const auto &r = bonded_to[i];
        return std::binary_search(r.begin(), r.end(), j);
    }

    // Neck strength of bond {i, j} (both ordered pairs share it). The
    // reference stores one draw per bond at neck_strengths[i] and
    // neck_strengths[j] but reads neck_strengths[i*n + j]
    // (alt_breaking_sinter_bridge.h:71-72, 118); here each bond keeps its own
    // draw, which is what the read intends.
    std::vector<std::vector<std::pair<uint32_t, double>>> neck;
    double neck_strength(size_t i, uint32_t j) const {
        const auto &r = neck[i];
        auto it = std::lower_bound(r.begin(), r.end(), j,
                                   [](const std::pair<uint32_t, double> &e, uint32_t key) { return e.first < key; });
        return (it != r.end() && it->first == j) ? it->second : INFINITY;
    }
    void set_neck_strengths(const double *s) {
        for (auto &r : neck) r.clear();
        for (size_t b = 0; b < bonds.size(); b++) {
            neck[bonds[b].first].push_back({bonds[b].second, s[b]});
            neck[bonds[b].second].push_back({bonds[b].first, s[b]});
        }
        for (auto &r : neck) std::sort(r.begin(), r.end());
    }

    // One ordered pair through the binary container fold.
    acc2 pair(size_t i, uint32_t j, double *d4 = nullptr) {
        bool have_b = p.binary_model != BIN_NONE, have_h = p.hamaker_model != HAM_NONE;
        acc2 b{vzero, vzero}, h{vzero, vzero};
        int64_t lb = p.binary_model == BIN_SINTER_LEGACY ? legacy_bond_of(i, j) : -1;
        if (have_b && lb >= 0) {
            // sinter_bridge.h:222-224
            b = legacy_bonded_eval(p.sinter.k, p.sinter.gamma_n, p.contact.r_part, p.contact.mass, p.contact.inertia,
                                   legacy_side(i, uint32_t(lb)), legacy_side(j, uint32_t(lb)), x[i], x[j], v[i], v[j],
                                   w[i], w[j]);
        } else if (have_b) {
            if (p.binary_model == BIN_SINTER && is_bonded(i, j)) {
                v3 ftot;
                b = sinter_bonded_eval(p.sinter, x[i], x[j], v[i], v[j], w[i], w[j], sinter_springs.at(i, j), &ftot,
                                       d4);
                // alt_breaking_sinter_bridge.h:117-119: the ordered pair (i, j)
                // stops being bonded once its force exceeds the neck strength;
                // from then on the embedded contact functor takes it, with
                // contact springs that were never touched (zero).
                if (p.sinter_breaking && norm(ftot) > neck_strength(i, j)) {
                    auto &r = bonded_to[i];
                    r.erase(std::lower_bound(r.begin(), r.end(), j));
                }
            } else {
                spring3 tmp{{vzero, vzero, vzero}};
                auto &row = contact_springs.rows[i];
                auto it = std::lower_bound(row.begin(), row.end(), j,
                                           [](const std::pair<uint32_t, spring3> &e, uint32_t key) { return e.first < key; });
                bool present = it != row.end() && it->first == j;
                if (present) tmp = it->second;
                bool reset = false;
                b = contact_eval(p.contact, x[i], x[j], v[i], v[j], w[i], w[j], tmp, reset, d4);
                if (reset) {
                    if (present) row.erase(it);
                } else if (present) {
                    it->second = tmp;
                } else {
                    row.insert(it, {j, tmp});
                }
            }
        }
        if (have_h) {
            h = binary_hamaker_eval(p.hamaker_model, p.hamaker, x[i], x[j]);
        }
        if (have_b && have_h) return b + h;
        return have_b ? b : h;
    }

    // Energy ledger terms of every particle, out[8*i + c]: kinetic translational
    // and rotational; the particle's half of its evaluated pairs' normal
    // elastic energy (k/2 overlap^2 for overlapping contacts and for bonds of
    // either sign); its half of its pairs' tangential, rolling and torsion
    // spring energy (each ordered pair's springs k/2 |xi|^2, halved); its half
    // of the pairs' Hamaker potential; 0. The builder's own ledger: the
    // reference reports kinetic energy only (test/compute_energy.cpp:7-58).
    void energy_terms(double mass, double inertia, double *out) const {
        for (size_t i = 0; i < n; i++) {
            double e[8] = {0.5 * mass * dot(v[i], v[i]), 0.5 * inertia * dot(w[i], w[i]), 0, 0, 0, 0, 0, 0};
            auto springs = [&](const spring_rows &rows, double kt, double kr, double ko) {
                for (const auto &en : rows.rows[i]) {
                    e[3] += 0.25 * kt * dot(en.second.s[0], en.second.s[0]);
                    e[4] += 0.25 * kr * dot(en.second.s[1], en.second.s[1]);
                    e[5] += 0.25 * ko * dot(en.second.s[2], en.second.s[2]);
                }
            };
            if (p.binary_model != BIN_NONE) {
                springs(contact_springs, p.contact.k_t, p.contact.k_r, p.contact.k_o);
                springs(sinter_springs, p.sinter.k_t, p.sinter.k_r, p.sinter.k_o);
            }
            auto pair = [&](uint32_t j) {
                const double dist = norm(x[i] - x[j]);
                const int64_t lb = p.binary_model == BIN_SINTER_LEGACY ? legacy_bond_of(i, j) : -1;
                if (lb >= 0) {
                    // the bond's five connector springs, k_d/2 (|s| - L0)^2
                    // each, half to each ordered pair
                    const v3 *ci = legacy_side(i, uint32_t(lb)), *cj = legacy_side(j, uint32_t(lb));
                    for (int q = 0; q < 5; q++) {
                        const double d = (q < 4 ? 2.0 * p.contact.r_part : 0.0) - norm(cj[q] - ci[q]);
                        e[2] += 0.25 * p.sinter.k * d * d;
                    }
                } else if (p.binary_model != BIN_NONE) {
                    if (p.binary_model == BIN_SINTER && is_bonded(i, j)) {
                        const double ov = 2.0 * p.sinter.r_part - dist;
                        e[2] += 0.25 * p.sinter.k * ov * ov;
                    } else {
                        const double ov = 2.0 * p.contact.r_part - dist;
                        if (ov > 0.0) e[2] += 0.25 * p.contact.k * ov * ov;
                    }
                }
                if (p.hamaker_model != HAM_NONE) e[6] += 0.5 * hamaker_potential(p.hamaker_model, p.hamaker, dist);
            };
            if (p.neighbor_mode == 1) {
                for (uint32_t j : nbr[i]) pair(j);
            } else {
                for (uint32_t j = 0; j < uint32_t(n); j++)
                    if (j != i) pair(j);
            }
            for (int c = 0; c < 8; c++) out[8 * i + c] = e[c];
        }
    }

    // The fold of one ordered pair on copies of its springs (no state change):
    // the denominator of per-particle force errors, sum_j |a_ij|.
    acc2 pair_peek(size_t i, uint32_t j) const {
        acc2 b, h;
        pair_peek_parts(i, j, b, h);
        bool have_b = p.binary_model != BIN_NONE, have_h = p.hamaker_model != HAM_NONE;
        if (have_b && have_h) return b + h;
        return have_b ? b : h;
    }

    // The two functors of one ordered pair separately (zero when absent).
    void pair_peek_parts(size_t i, uint32_t j, acc2 &b, acc2 &h) const {
        bool have_b = p.binary_model != BIN_NONE, have_h = p.hamaker_model != HAM_NONE;
        b = acc2{vzero, vzero};
        h = acc2{vzero, vzero};
        auto find = [&](const spring_rows &rows) {
            spring3 s{{vzero, vzero, vzero}};
            const auto &r = rows.rows[i];
            auto it = std::lower_bound(r.begin(), r.end(), j,
                                       [](const std::pair<uint32_t, spring3> &e, uint32_t key) { return e.first < key; });
            if (it != r.end() && it->first == j) s = it->second;
            return s;
        };
        if (have_b) {
            if (p.binary_model == BIN_SINTER && is_bonded(i, j)) {
                spring3 s = find(sinter_springs);
                b = sinter_bonded_eval(p.sinter, x[i], x[j], v[i], v[j], w[i], w[j], s);
            } else {
                spring3 s = find(contact_springs);
                bool reset;
                b = contact_eval(p.contact, x[i], x[j], v[i], v[j], w[i], w[j], s, reset);
            }
        }
        if (have_h)
            h = binary_hamaker_eval(p.hamaker_model, p.hamaker, x[i], x[j]);
    }

    // triangular_facet.h:32-51 then the unary container's right fold over facets.
    acc2 unary(size_t i) {
        size_t nf = facets.size();
        std::vector<acc2> per(nf);
        for (size_t f = 0; f < nf; f++) {
            v3 xf = closest_point_on_triangle(facets[f], x[i]);
            acc2 c{vzero, vzero}, h{vzero, vzero};
            if (p.facet_contact)
                c = surface_contact_eval(p.surface_contact, xf, facets[f].v_facet, x[i], v[i], w[i],
                                         surface_springs[f].at(i, 0));
            if (p.facet_hamaker) h = surface_hamaker_eval(p.surface_hamaker, xf, x[i]);
            if (p.facet_contact && p.facet_hamaker) per[f] = c + h;
            else per[f] = p.facet_contact ? c : h;
        }
        acc2 s = per[nf - 1];
        for (size_t f = nf - 1; f-- > 0;) s = per[f] + s;
        return s;
    }

    void compute_accelerations() {
        const long nn = long(n);
#pragma omp parallel for schedule(dynamic, 64)
        for (long ii = 0; ii < nn; ii++) {
            size_t i = size_t(ii);
            acc2 acc{vzero, vzero};
            double d4[4] = {0.0, 0.0, 0.0, 0.0};
            double *dp = track_diss ? d4 : nullptr;
            if (p.neighbor_mode == 1) {
                for (uint32_t j : nbr[i]) {
                    acc2 e = pair(i, j, dp);
                    acc.lin = acc.lin + e.lin;
                    acc.ang = acc.ang + e.ang;
                }
            } else {
                for (uint32_t j = 0; j < uint32_t(n); j++) {
                    if (j == i) continue;
                    acc2 e = pair(i, j, dp);
                    acc.lin = acc.lin + e.lin;
                    acc.ang = acc.ang + e.ang;
                }
            }
            if (track_diss)
                for (int c = 0; c < 4; c++) diss[4 * i + c] = diss[4 * i + c] + d4[c];
            if (!facets.empty()) acc = acc + unary(i);
            a[i] = acc.lin;
            al[i] = acc.ang;
        }
        acc_valid = true;
    }

    // Neighbour list: brute force (the checker) or a cell list that must agree.
    void rebuild_neighbors(bool brute) {
        nbr.assign(n, {});
        const double rv = p.r_verlet;
        if (brute || n < 64) {
#pragma omp parallel for schedule(dynamic, 64)
            for (long ii = 0; ii < long(n); ii++) {
                size_t i = size_t(ii);
                for (size_t j = 0; j < n; j++) {
                    if (j == i) continue;
                    if (norm(x[i] - x[j]) < rv) nbr[i].push_back(uint32_t(j));
                }
            }
            return;
        }
        // Cell list over the bounding box; cells of edge >= r_verlet.
        v3 lo = x[0], hi = x[0];
        for (auto &q : x) {
            lo = {std::min(lo.x, q.x), std::min(lo.y, q.y), std::min(lo.z, q.z)};
            hi = {std::max(hi.x, q.x), std::max(hi.y, q.y), std::max(hi.z, q.z)};
        }
        const double cs = rv;
        long nx = long((hi.x - lo.x) / cs) + 1, ny = long((hi.y - lo.y) / cs) + 1, nz = long((hi.z - lo.z) / cs) + 1;
        // sparse cells: sort particles by linear cell id
        std::vector<std::pair<long, uint32_t>> key(n);
        auto cell_of = [&](v3 q, long &cx, long &cy, long &cz) {
            cx = std::min(nx - 1, long((q.x - lo.x) / cs));
            cy = std::min(ny - 1, long((q.y - lo.y) / cs));
            cz = std::min(nz - 1, long((q.z - lo.z) / cs));
        };
        for (size_t i = 0; i < n; i++) {
            long cx, cy, cz;
            cell_of(x[i], cx, cy, cz);
            key[i] = {(cz * ny + cy) * nx + cx, uint32_t(i)};
        }
        std::sort(key.begin(), key.end());
#pragma omp parallel for schedule(dynamic, 64)
        for (long ii = 0; ii < long(n); ii++) {
            size_t i = size_t(ii);
            long cx, cy, cz;
            cell_of(x[i], cx, cy, cz);
            auto &out = nbr[i];
            for (long dz = -1; dz <= 1; dz++)
                for (long dy = -1; dy <= 1; dy++)
                    for (long dx = -1; dx <= 1; dx++) {
                        long qx = cx + dx, qy = cy + dy, qz = cz + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx >= nx || qy >= ny || qz >= nz) continue;
                        long c = (qz * ny + qy) * nx + qx;
                        auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(c, uint32_t(0)));
                        for (; it != key.end() && it->first == c; ++it) {
                            uint32_t j = it->second;
                            if (j == i) continue;
                            if (norm(x[i] - x[j]) < rv) out.push_back(j);
                        }
                    }
            std::sort(out.begin(), out.end());
        }
    }

    void step(double dt) {
        if (!acc_valid) compute_accelerations();
        const double half = dt / 2.0;
        for (size_t i = 0; i < n; i++) {
            v[i] = v[i] + a[i] * half;
            w[i] = w[i] + al[i] * half;
        }
        for (size_t i = 0; i < n; i++) {
            const v3 dx = v[i] * dt, dth = w[i] * dt;
            x[i] = x[i] + dx;
            if (!legacy_of.empty())
                for (auto &e : legacy_of[i]) legacy_move_x(legacy_side(i, e.second), dx);
            th[i] = th[i] + dth;
            if (!legacy_of.empty())
                for (auto &e : legacy_of[i]) legacy_move_theta(legacy_side(i, e.second), x[i], dth);
        }
        t = t + dt;
        compute_accelerations();
        for (size_t i = 0; i < n; i++) {
            v[i] = v[i] + a[i] * half;
            w[i] = w[i] + al[i] * half;
        }
    }

    // triangular_facet.h:62-66
    void advance_facets(double dt) {
        for (auto &f : facets)
            for (int k = 0; k < 3; k++) f.vert[k] = f.vert[k] + f.v_facet * dt;
    }
};

}  // namespace oracle
