// Vector algebra and pair/facet force models of the DEM step, for gfx950
// device code and the host-side set-up code.
//
// The FP64 functions follow the reference functors' expression order
// term by term (cited per function) so that, compiled with
// -ffp-contract=off, they reproduce the reference's double-precision results
// bit for bit. The vector type carries Eigen::Vector3d's semantics as the
// functors use them: dot = (x*x' + y*y') + z*z', norm = sqrt(dot),
// normalized() of a zero vector returns it unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipgran.h"

#define HGD __host__ __device__ __forceinline__

namespace hg {

struct d3 {
    double x, y, z;
};

HGD d3 mk(double x, double y, double z) { return d3{x, y, z}; }
HGD d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HGD d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HGD d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
HGD d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
HGD d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
HGD d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
HGD double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
HGD double norm(d3 a) { return sqrt(dot(a, a)); }
HGD d3 normalized(d3 a) {
    double z = dot(a, a);
    if (z > 0.0) return a / sqrt(z);
    return a;
}
HGD d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HGD d3 zero3() { return {0.0, 0.0, 0.0}; }
// std::min(a, b) semantics: returns a unless b < a.
HGD double min_ref(double a, double b) { return (b < a) ? b : a; }

struct acc6 {
    d3 lin, ang;
};
HGD acc6 operator+(acc6 a, acc6 b) { return {a.lin + b.lin, a.ang + b.ang}; }

// Tangential, rolling and torsion spring of one ordered pair.
struct spring9 {
    d3 s[3];
};
HGD spring9 zero_spring() { return spring9{{zero3(), zero3(), zero3()}}; }

template <typename T> HGD T sq(T x) { return x * x; }
template <typename T> HGD T cb(T x) { return x * x * x; }

// ---- contact_force.h:105-147 (compute_shear_contribution) ------------------
HGD d3 shear_contact(d3 &xi, d3 n, double stiffness, double damping, double normal_force, double mu_static,
                     double phi_dynamic, d3 rel_v, double dt) {
    double mu_dynamic = mu_static * phi_dynamic;
    d3 xi_new = xi - dot(xi, n) * n;
    if (norm(xi_new) > 0.0) {
        xi_new = xi_new / sqrt(dot(xi_new, xi_new));
        xi_new = xi_new * norm(xi);
    }
    xi = xi_new;
    d3 f0 = (-stiffness) * xi - damping * rel_v;
    double static_friction = mu_static * normal_force;
    double dynamic_friction = mu_dynamic * normal_force;
    double f0n = norm(f0);
    d3 t = f0n > 0.0 ? normalized(f0) : zero3();
    double f_selected;
    if (f0n <= static_friction) {
        xi = xi + rel_v * dt;
        f_selected = static_friction;
    } else {
        xi = (-1.0 / stiffness) * (dynamic_friction * t + damping * rel_v);
        f_selected = dynamic_friction;
    }
    return min_ref(f_selected, f0n) * t;
}

// Dissipated work of one ordered pair over one step of length dt (the
// builder's ledger term; the reference has none). The pair's forces on both
// particles take work -f_n v_n + f_t.v_t + (r/rpp) f_r.v_r + f_o.v_o per unit
// time from their motion (the rolling torque uses r, its velocity rpp); the
// elastic part of f_n is the gradient of U_normal, the rest of the normal
// work, gamma_n v_n^2, is lost, and of each shear term the work minus the
// change of its spring's energy k/2 |xi|^2 is lost (damping, Coulomb slip,
// the rescaled rotation). Each ordered pair holds half of the pair: d4 +=
// {gamma_n v_n^2 dt / 2, then per shear term c: -w_c (f_c.v_c) dt / 2 -
// k_c/4 (|xi_c|^2 - |xi_c,old|^2)}, w = 1, r/rpp, 1.
HGD void pair_dissipation(double *d4, double gamma_n, double v_n, double dt, double w_r, const double k[3],
                          const d3 f[3], const d3 v[3], const double u_old[3], const spring9 &sp) {
    d4[0] = d4[0] + 0.5 * gamma_n * v_n * v_n * dt;
    for (int c = 0; c < 3; c++) {
        const double w = c == 1 ? w_r : 1.0;
        d4[1 + c] = d4[1 + c] + (-0.5 * w * dot(f[c], v[c]) * dt - 0.25 * k[c] * (dot(sp.s[c], sp.s[c]) - u_old[c]));
    }
}

// ---- contact_force.h:57-96 --------------------------------------------------
// `active` is false where the reference resets the pair's springs
// (overlap <= 0, contact_force.h:68-71); the spring is then zero.
// vw(vj, wj) gives the partner's velocity and angular velocity; it is called
// only for an overlapping pair (the others need the positions alone). TRACK
// adds the pair's dissipated work to d4 (pair_dissipation); the springs a
// separating pair drops are lost whole.
template <bool TRACK = false, class VW>
HGD acc6 contact_eval_vw(const hg_contact_params &p, d3 xi, d3 xj, d3 vi, d3 wi, VW &&vw, spring9 &sp,
                         bool &active, double *d4 = nullptr) {
    d3 d = xi - xj;
    d3 n = normalized(d);
    double overlap = 2.0 * p.r_part - dot(d, n);
    if (overlap <= 0) {
        if constexpr (TRACK) {
            const double k[3] = {p.k_t, p.k_r, p.k_o};
            for (int c = 0; c < 3; c++) d4[1 + c] = d4[1 + c] + 0.25 * k[c] * dot(sp.s[c], sp.s[c]);
        }
        active = false;
        sp = zero_spring();
        return {zero3(), zero3()};
    }
    active = true;
    d3 vj, wj;
    vw(vj, wj);
    double rpp = p.r_part - overlap / 2.0;
    double v_n = -dot(vi - vj, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    d3 cwi = cross(n, wi), cwj = cross(n, wj);
    d3 v_ij = ((vi - vj) + rpp * cwi) + rpp * cwj;
    d3 v_t = v_ij - dot(v_ij, n) * n;
    d3 v_r = (-rpp) * (cwi - cwj);
    d3 v_o = (p.r_part * (dot(n, wi) - dot(n, wj))) * n;
    double u_old[3] = {0.0, 0.0, 0.0};
    if constexpr (TRACK)
        for (int c = 0; c < 3; c++) u_old[c] = dot(sp.s[c], sp.s[c]);
    d3 f_t = shear_contact(sp.s[0], n, p.k_t, p.gamma_t, f_n, p.mu_s, p.phi_d, v_t, p.dt);
    d3 f_r = shear_contact(sp.s[1], n, p.k_r, p.gamma_r, f_n, p.mu_r, p.phi_r, v_r, p.dt);
    d3 f_o = shear_contact(sp.s[2], n, p.k_o, p.gamma_o, f_n, p.mu_o, p.phi_o, v_o, p.dt);
    if constexpr (TRACK) {
        const double k[3] = {p.k_t, p.k_r, p.k_o};
        const d3 f[3] = {f_t, f_r, f_o}, v[3] = {v_t, v_r, v_o};
        pair_dissipation(d4, p.gamma_n, v_n, p.dt, p.r_part / rpp, k, f, v, u_old, sp);
    }
    d3 tau_t = rpp * cross(n, f_t);
    d3 tau_r = p.r_part * cross(n, f_r);
    d3 tau_o = p.r_part * f_o;
    return {(f_n * n + f_t) / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}
HGD acc6 contact_eval(const hg_contact_params &p, d3 xi, d3 xj, d3 vi, d3 vj, d3 wi, d3 wj, spring9 &sp,
                      bool &active) {
    return contact_eval_vw(p, xi, xj, vi, wi, [&](d3 &v, d3 &w) { v = vj; w = wj; }, sp, active);
}

// ---- alt_sinter_bridge.h:137-156 -------------------------------------------
HGD d3 shear_sinter(d3 &xi, d3 n, double stiffness, double damping, d3 rel_v, double dt) {
    d3 xi_new = xi - dot(xi, n) * n;
    if (norm(xi_new) > 0.0) {
        xi_new = xi_new / sqrt(dot(xi_new, xi_new));
        xi_new = xi_new * norm(xi);
    }
    xi = xi_new;
    d3 f0 = (-stiffness) * xi - damping * rel_v;
    xi = xi + rel_v * dt;
    return f0;
}

// ---- alt_sinter_bridge.h:92-129 (bonded branch; no overlap gate) ----------
// ftot receives f_n n + f_t, the force alt_breaking_sinter_bridge.h:115-118
// compares with the neck strength. TRACK: as contact_eval_vw.
template <bool TRACK = false>
HGD acc6 sinter_eval(const hg_sinter_params &p, d3 xi, d3 xj, d3 vi, d3 vj, d3 wi, d3 wj, spring9 &sp, d3 &ftot,
                     double *d4 = nullptr) {
    d3 d = xi - xj;
    d3 n = normalized(d);
    double overlap = 2.0 * p.r_part - dot(d, n);
    double rpp = p.r_part - overlap / 2.0;
    double v_n = -dot(vi - vj, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    d3 cwi = cross(n, wi), cwj = cross(n, wj);
    d3 v_ij = ((vi - vj) + rpp * cwi) + rpp * cwj;
    d3 v_t = v_ij - dot(v_ij, n) * n;
    d3 v_r = (-rpp) * (cwi - cwj);
    d3 v_o = (p.r_part * (dot(n, wi) - dot(n, wj))) * n;
    double u_old[3] = {0.0, 0.0, 0.0};
    if constexpr (TRACK)
        for (int c = 0; c < 3; c++) u_old[c] = dot(sp.s[c], sp.s[c]);
    d3 f_t = shear_sinter(sp.s[0], n, p.k_t, p.gamma_t, v_t, p.dt);
    d3 f_r = shear_sinter(sp.s[1], n, p.k_r, p.gamma_r, v_r, p.dt);
    d3 f_o = shear_sinter(sp.s[2], n, p.k_o, p.gamma_o, v_o, p.dt);
    if constexpr (TRACK) {
        const double k[3] = {p.k_t, p.k_r, p.k_o};
        const d3 f[3] = {f_t, f_r, f_o}, v[3] = {v_t, v_r, v_o};
        pair_dissipation(d4, p.gamma_n, v_n, p.dt, p.r_part / rpp, k, f, v, u_old, sp);
    }
    d3 tau_t = rpp * cross(n, f_t);
    d3 tau_r = p.r_part * cross(n, f_r);
    d3 tau_o = p.r_part * f_o;
    ftot = f_n * n + f_t;
    return {ftot / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}

// ---- sinter_bridge.h:235-259 (legacy sinter_functor, bonded pair) --------
// Five springs between matching connectors, ci attached to i and cj to j:
// rest length 2r for the four side springs, 0 for the one at the contact
// point; force -(L0 - |s|) k_d s/|s| + gamma_d (relative connector velocity),
// torque about x_i.
HGD acc6 legacy_bonded_eval(double k_d, double gamma_d, double r_part, double mass, double inertia, const d3 *ci,
                            const d3 *cj, d3 xi, d3 xj, d3 vi, d3 vj, d3 wi, d3 wj) {
    d3 force = zero3(), torque = zero3();
    for (int q = 0; q < 5; q++) {
        const d3 spring = cj[q] - ci[q];
        const double spring_length = norm(spring);
        const double equilibrium_length = q < 4 ? 2.0 * r_part : 0.0;
        const d3 conn_i_v = vi + cross(wi, ci[q] - xi);
        const d3 conn_j_v = vj + cross(wj, cj[q] - xj);
        const d3 rel_v = conn_j_v - conn_i_v;
        const d3 dforce = -(equilibrium_length - spring_length) * k_d * normalized(spring) + gamma_d * rel_v;
        force = force + dforce;
        const d3 arm = ci[q] - xi;
        torque = torque + cross(arm, dforce);
    }
    return {force / mass, torque / inertia};
}

// ---- hamaker_force.h:36-39: the scalar factor of -dU/dh --------------------
HGD double hamaker_scalar(double A, double rp, double h) {
    return (-A) / 6.0 *
           ((((4.0 * rp + 2.0 * h) / (4.0 * rp + h) / h - 2.0 / (2.0 * rp + h)) - 4.0 * sq(rp) / cb(2.0 * rp + h)) -
            2.0 * sq(rp) * (4.0 * rp + 2.0 * h) / sq(4.0 * rp + h) / sq(h));
}

// ---- hamaker_force.h:28-43 -------------------------------------------------
HGD acc6 hamaker_eval(const hg_hamaker_params &p, d3 xi, d3 xj) {
    d3 r = xj - xi;
    double h = norm(r) - 2.0 * p.r_part;
    if (h < p.h0) h = p.h0;
    d3 f = hamaker_scalar(p.A, p.r_part, h) * normalized(r);
    return {f / p.mass, zero3()};
}

// ---- limited_range_hamaker_force.h:29-47 (strict '>' at the cutoff) -------
HGD acc6 lr_hamaker_eval(const hg_hamaker_params &p, d3 xi, d3 xj) {
    if (norm(xi - xj) > p.r_cutoff) return {zero3(), zero3()};
    return hamaker_eval(p, xi, xj);
}

// ---- energy ledger: Hamaker potentials (the reference defines no potential
// energy; README.md:154-157 gives the one below, whose derivative dU/dd is
// hamaker_scalar). Below the saturation gap h0 the functors apply a constant
// force, so U continues linearly there, and the limited-range functor's
// potential is shifted to zero at its cutoff: dU/dd is the applied force.
HGD double hamaker_potential_raw(double A, double r, double d) {
    return (-A) / 6.0 *
           (2.0 * sq(r) / ((4.0 * r + d) * d) + 2.0 * sq(r) / sq(2.0 * r + d) + log((4.0 * r + d) * d / sq(2.0 * r + d)));
}
HGD double hamaker_u(int model, const hg_hamaker_params &p, double d) {
    return model == HG_HAMAKER_SIMPLIFIED ? -2.0 * p.A * p.r_part / d : hamaker_potential_raw(p.A, p.r_part, d);
}
HGD double hamaker_dudd(int model, const hg_hamaker_params &p, double d) {
    return model == HG_HAMAKER_SIMPLIFIED ? 2.0 * p.A * p.r_part / sq(d) : hamaker_scalar(p.A, p.r_part, d);
}
HGD double hamaker_u_sat(int model, const hg_hamaker_params &p, double d) {
    return d < p.h0 ? hamaker_u(model, p, p.h0) + hamaker_dudd(model, p, p.h0) * (d - p.h0) : hamaker_u(model, p, d);
}
// Potential of one pair at centre distance `dist`.
HGD double hamaker_potential(int model, const hg_hamaker_params &p, double dist) {
    const double d = dist - 2.0 * p.r_part;
    if (model == HG_HAMAKER_LIMITED) {
        if (dist > p.r_cutoff) return 0.0;
        return hamaker_u_sat(model, p, d) - hamaker_u_sat(model, p, p.r_cutoff - 2.0 * p.r_part);
    }
    return hamaker_u_sat(model, p, d);
}

// ---- simplified_hamaker_force.h:18-30: 2 A r / h^2 along x_j - x_i ----------
HGD acc6 simplified_hamaker_eval(const hg_hamaker_params &p, d3 xi, d3 xj) {
    d3 r = xj - xi;
    double h = norm(r) - 2.0 * p.r_part;
    if (h < p.h0) h = p.h0;
    d3 f = 2.0 * p.A * p.r_part / sq(h) * normalized(r);
    return {f / p.mass, zero3()};
}

// ---- triangular facet (triangular_facet.h:18-137) --------------------------
struct facet_dev {
    d3 vert[3];
    d3 v_facet;
    d3 normal;
    double area;
    double pad[2];
};

HGD d3 facet_edge_point(const facet_dev &f, int a, int b, d3 p) {
    d3 lv = f.vert[b] - f.vert[a];
    d3 ulv = normalized(lv);
    d3 tpv = p - f.vert[a];
    double len = norm(lv);
    double pr = dot(tpv, ulv);
    if (pr < 0.0) return f.vert[a];
    if (pr > len) return f.vert[b];
    return f.vert[a] + pr * ulv;
}

HGD d3 closest_point_on_triangle(const facet_dev &f, d3 p) {
    d3 offset = dot(f.vert[0], f.normal) * f.normal;
    d3 proj = p - dot(p, f.normal) * f.normal;
    d3 cp = proj + offset;
    double area = 0.0;
    d3 pa = f.vert[0] - cp, pb = f.vert[1] - cp, pc = f.vert[2] - cp;
    area += 1.0 / 2.0 * norm(cross(pa, pb));
    area += 1.0 / 2.0 * norm(cross(pb, pc));
    area += 1.0 / 2.0 * norm(cross(pc, pa));
    if (!(area > f.area)) return cp;
    d3 e1 = facet_edge_point(f, 0, 1, p), e2 = facet_edge_point(f, 1, 2, p), e3 = facet_edge_point(f, 2, 0, p);
    d3 q1 = p - e1, q2 = p - e2, q3 = p - e3;
    double s1 = dot(q1, q1), s2 = dot(q2, q2), s3 = dot(q3, q3);
    if (s1 < s2) {
        if (s1 < s3) return e1;
    } else {
        if (s2 < s3) return e2;
    }
    return e3;
}

// ---- surface_contact_force.h:57-96 ------------------------------------------
HGD acc6 surface_contact_eval(const hg_contact_params &p, d3 xf, d3 vf, d3 xi, d3 vi, d3 wi, spring9 &sp) {
    d3 d = xi - xf;
    d3 n = normalized(d);
    double overlap = p.r_part - dot(d, n);
    if (overlap <= 0) {
        sp = zero_spring();
        return {zero3(), zero3()};
    }
    double rpp = p.r_part - overlap;
    double v_n = -dot(vi - vf, n);
    double f_n = p.k * overlap + p.gamma_n * v_n;
    d3 cw = cross(n, wi);
    d3 v_ij = (vi - vf) + rpp * cw;
    d3 v_t = v_ij - dot(v_ij, n) * n;
    d3 v_r = (-rpp) * cw;
    d3 v_o = (p.r_part * dot(n, wi)) * n;
    d3 f_t = shear_contact(sp.s[0], n, p.k_t, p.gamma_t, f_n, p.mu_s, p.phi_d, v_t, p.dt);
    d3 f_r = shear_contact(sp.s[1], n, p.k_r, p.gamma_r, f_n, p.mu_r, p.phi_r, v_r, p.dt);
    d3 f_o = shear_contact(sp.s[2], n, p.k_o, p.gamma_o, f_n, p.mu_o, p.phi_o, v_o, p.dt);
    d3 tau_t = rpp * cross(n, f_t);
    d3 tau_r = p.r_part * cross(n, f_r);
    d3 tau_o = p.r_part * f_o;
    return {(f_n * n + f_t) / p.mass, ((-tau_t + tau_r) + tau_o) / p.inertia};
}

// ---- surface_hamaker_force.h:18-30 ------------------------------------------
HGD acc6 surface_hamaker_eval(const hg_hamaker_params &p, d3 xf, d3 xi) {
    d3 r = xf - xi;
    double h = norm(r) - p.r_part;
    if (h < p.h0) h = p.h0;
    d3 f = 2.0 * p.A / 3.0 * (cb(p.r_part) / sq(h) / sq(h + 2.0 * p.r_part)) * normalized(r);
    return {f / p.mass, zero3()};
}

}  // namespace hg
