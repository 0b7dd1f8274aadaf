// C entry points of the CPU oracle (test infrastructure only; see oracle.hpp).
//
// Two groups:
//   * or_*_seq: drive one functor through a sequence of pair states, the same
//     way oracle/ref_harness.cpp drives the reference functor headers, so the
//     two can be compared output for output.
//   * or_sys_*: the whole system (functor fold + acceleration pass +
//     integrator + neighbour list), used as the parity reference for the HIP
//     path on identical inputs.
#include "oracle.hpp"

#include <cstring>

using namespace oracle;

namespace {
v3 ld(const double *p) { return {p[0], p[1], p[2]}; }
void st(double *p, v3 a) {
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}
void st_spring(double *p, const spring3 &s) {
    for (int k = 0; k < 3; k++) st(p + 3 * k, s.s[k]);
}
contact_params cp_from(const double *q) {
    contact_params c;
    std::memcpy(&c, q, sizeof(c));
    return c;
}
sinter_params sp_from(const double *q) {
    sinter_params s;
    std::memcpy(&s, q, sizeof(s));
    return s;
}
hamaker_params hp_from(const double *q) {
    hamaker_params h;
    std::memcpy(&h, q, sizeof(h));
    return h;
}
}  // namespace

extern "C" {

// Pair state per step: xi, xj, vi, vj, wi, wj (18 doubles). Output per step:
// (i=0,j=1) lin, ang then (i=1,j=0) lin, ang (12 doubles), and the two spring
// triples after the step (18 doubles) when springs_out is not null.
void or_contact_seq(const double *params18, size_t nsteps, const double *states, double *out,
                    double *springs_out) {
    contact_params p = cp_from(params18);
    spring3 s01{{vzero, vzero, vzero}}, s10{{vzero, vzero, vzero}};
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 18 * k;
        v3 x0 = ld(s), x1 = ld(s + 3), v0 = ld(s + 6), v1 = ld(s + 9), w0 = ld(s + 12), w1 = ld(s + 15);
        bool r;
        acc2 e01 = contact_eval(p, x0, x1, v0, v1, w0, w1, s01, r);
        acc2 e10 = contact_eval(p, x1, x0, v1, v0, w1, w0, s10, r);
        st(out + 12 * k, e01.lin);
        st(out + 12 * k + 3, e01.ang);
        st(out + 12 * k + 6, e10.lin);
        st(out + 12 * k + 9, e10.ang);
        if (springs_out) {
            st_spring(springs_out + 18 * k, s01);
            st_spring(springs_out + 18 * k + 9, s10);
        }
    }
}

// alt_sinter with its embedded contact functor; bonding decided from the first
// state's positions with the reference predicate. bonded_out gets 1/0.
void or_sinter_seq(const double *sinter13, const double *contact18, size_t nsteps, const double *states,
                   double *out, double *springs_out, int *bonded_out) {
    sinter_params sp = sp_from(sinter13);
    contact_params cp = cp_from(contact18);
    std::vector<v3> x0 = {ld(states), ld(states + 3)};
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped;
    sinter_topology(2, x0, sp.r_part, sp.d_crit, bonds, skipped);
    bool bonded = !bonds.empty();
    *bonded_out = bonded ? 1 : 0;
    spring3 s01{{vzero, vzero, vzero}}, s10{{vzero, vzero, vzero}};  // sinter springs
    spring3 c01{{vzero, vzero, vzero}}, c10{{vzero, vzero, vzero}};  // contact springs
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 18 * k;
        v3 x0_ = ld(s), x1 = ld(s + 3), v0 = ld(s + 6), v1 = ld(s + 9), w0 = ld(s + 12), w1 = ld(s + 15);
        acc2 e01, e10;
        bool r;
        if (bonded) {
            e01 = sinter_bonded_eval(sp, x0_, x1, v0, v1, w0, w1, s01);
            e10 = sinter_bonded_eval(sp, x1, x0_, v1, v0, w1, w0, s10);
        } else {
            e01 = contact_eval(cp, x0_, x1, v0, v1, w0, w1, c01, r);
            e10 = contact_eval(cp, x1, x0_, v1, v0, w1, w0, c10, r);
        }
        st(out + 12 * k, e01.lin);
        st(out + 12 * k + 3, e01.ang);
        st(out + 12 * k + 6, e10.lin);
        st(out + 12 * k + 9, e10.ang);
        if (springs_out) {
            st_spring(springs_out + 18 * k, s01);
            st_spring(springs_out + 18 * k + 9, s10);
        }
    }
}

// alt_breaking_sinter on the ordered pair (0, 1) only, with neck strength
// `strength`: out = 6 doubles per step (a, alpha of particle 0); broken_out[k]
// = 1 once the pair has broken (after step k's evaluation). Bonding as the
// reference decides it on the first state (every pair within d_crit).
void or_breaking_seq(const double *sinter13, const double *contact18, double strength, size_t nsteps,
                     const double *states, double *out, int *broken_out) {
    sinter_params sp = sp_from(sinter13);
    contact_params cp = cp_from(contact18);
    std::vector<v3> x0 = {ld(states), ld(states + 3)};
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped;
    sinter_topology(2, x0, sp.r_part, sp.d_crit, bonds, skipped, false);
    bool bonded = !bonds.empty();
    spring3 s01{{vzero, vzero, vzero}}, c01{{vzero, vzero, vzero}};
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 18 * k;
        v3 x0_ = ld(s), x1 = ld(s + 3), v0 = ld(s + 6), v1 = ld(s + 9), w0 = ld(s + 12), w1 = ld(s + 15);
        acc2 e01;
        if (bonded) {
            v3 ftot;
            e01 = sinter_bonded_eval(sp, x0_, x1, v0, v1, w0, w1, s01, &ftot);
            if (norm(ftot) > strength) bonded = false;
        } else {
            bool r;
            e01 = contact_eval(cp, x0_, x1, v0, v1, w0, w1, c01, r);
        }
        st(out + 6 * k, e01.lin);
        st(out + 6 * k + 3, e01.ang);
        broken_out[k] = bonded ? 0 : 1;
    }
}

void or_sys_set_neck_strengths(void *s, const double *strengths) {
    static_cast<dem_system *>(s)->set_neck_strengths(strengths);
}

// mode 1: hamaker_functor, 2: limited_range_hamaker_functor,
// 3: simplified_hamaker_functor.
void or_hamaker_seq(int mode, const double *ham5, size_t nsteps, const double *states, double *out) {
    hamaker_params hp = hp_from(ham5);
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 18 * k;
        v3 x0 = ld(s), x1 = ld(s + 3);
        acc2 e01 = binary_hamaker_eval(mode, hp, x0, x1);
        acc2 e10 = binary_hamaker_eval(mode, hp, x1, x0);
        st(out + 12 * k, e01.lin);
        st(out + 12 * k + 3, e01.ang);
        st(out + 12 * k + 6, e10.lin);
        st(out + 12 * k + 9, e10.ang);
    }
}

// Container fold of {alt_sinter (with contact), limited_range_hamaker}.
void or_fold_seq(const double *sinter13, const double *contact18, const double *ham5, size_t nsteps,
                 const double *states, double *out) {
    sinter_params sp = sp_from(sinter13);
    contact_params cp = cp_from(contact18);
    hamaker_params hp = hp_from(ham5);
    std::vector<v3> x0 = {ld(states), ld(states + 3)};
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    size_t skipped;
    sinter_topology(2, x0, sp.r_part, sp.d_crit, bonds, skipped);
    bool bonded = !bonds.empty();
    spring3 s01{{vzero, vzero, vzero}}, s10{{vzero, vzero, vzero}};
    spring3 c01{{vzero, vzero, vzero}}, c10{{vzero, vzero, vzero}};
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 18 * k;
        v3 x0_ = ld(s), x1 = ld(s + 3), v0 = ld(s + 6), v1 = ld(s + 9), w0 = ld(s + 12), w1 = ld(s + 15);
        bool r;
        acc2 b01 = bonded ? sinter_bonded_eval(sp, x0_, x1, v0, v1, w0, w1, s01)
                          : contact_eval(cp, x0_, x1, v0, v1, w0, w1, c01, r);
        acc2 b10 = bonded ? sinter_bonded_eval(sp, x1, x0_, v1, v0, w1, w0, s10)
                          : contact_eval(cp, x1, x0_, v1, v0, w1, w0, c10, r);
        acc2 e01 = b01 + lr_hamaker_eval(hp, x0_, x1);
        acc2 e10 = b10 + lr_hamaker_eval(hp, x1, x0_);
        st(out + 12 * k, e01.lin);
        st(out + 12 * k + 3, e01.ang);
        st(out + 12 * k + 6, e10.lin);
        st(out + 12 * k + 9, e10.ang);
    }
}

// One particle against one triangular facet carrying {surface contact,
// surface hamaker}. Facet: 3 vertices (9) + velocity (3). Per step state:
// x, v, w (9). Output per step: lin, ang (6). The facet is advanced by
// dt_advance after every step (triangular_facet.h:62-66).
void or_facet_seq(const double *verts9, const double *vfacet3, const double *contact18, const double *ham5,
                  double dt_advance, size_t nsteps, const double *states, double *out) {
    contact_params cp = cp_from(contact18);
    hamaker_params hp = hp_from(ham5);
    facet f = make_facet(ld(vfacet3), ld(verts9), ld(verts9 + 3), ld(verts9 + 6));
    spring3 sp{{vzero, vzero, vzero}};
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 9 * k;
        v3 x = ld(s), v = ld(s + 3), w = ld(s + 6);
        v3 xf = closest_point_on_triangle(f, x);
        acc2 e = surface_contact_eval(cp, xf, f.v_facet, x, v, w, sp) + surface_hamaker_eval(hp, xf, x);
        st(out + 6 * k, e.lin);
        st(out + 6 * k + 3, e.ang);
        for (int q = 0; q < 3; q++) f.vert[q] = f.vert[q] + f.v_facet * dt_advance;
    }
}

// Closest point on a facet for a batch of query points (3 per point).
void or_closest_point(const double *verts9, size_t npts, const double *pts, double *out) {
    facet f = make_facet(vzero, ld(verts9), ld(verts9 + 3), ld(verts9 + 6));
    for (size_t k = 0; k < npts; k++) st(out + 3 * k, closest_point_on_triangle(f, ld(pts + 3 * k)));
}

// alt_sinter topology; returns the number of bonds (pairs i<j in insertion
// order written to pairs_out when it has room), skipped cycles to *skipped.
size_t or_sinter_topology(size_t n, const double *x0, double r_part, double d_crit, uint32_t *pairs_out,
                          size_t cap, size_t *skipped) {
    std::vector<v3> x(n);
    for (size_t i = 0; i < n; i++) x[i] = ld(x0 + 3 * i);
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    sinter_topology(n, x, r_part, d_crit, bonds, *skipped);
    for (size_t b = 0; b < bonds.size() && b < cap; b++) {
        pairs_out[2 * b] = bonds[b].first;
        pairs_out[2 * b + 1] = bonds[b].second;
    }
    return bonds.size();
}

// ------------------------------------------------------------ system API ----
void *or_sys_create(const system_params *p, size_t n, const double *x0, const double *v0, const double *th0,
                    const double *w0) {
    return new dem_system(*p, n, x0, v0, th0, w0);
}
void or_sys_destroy(void *s) { delete static_cast<dem_system *>(s); }

void or_sys_set_facets(void *s, size_t nf, const double *verts, const double *vel) {
    auto *S = static_cast<dem_system *>(s);
    S->facets.clear();
    S->surface_springs.assign(nf, {});
    for (size_t f = 0; f < nf; f++) {
        S->facets.push_back(make_facet(ld(vel + 3 * f), ld(verts + 9 * f), ld(verts + 9 * f + 3), ld(verts + 9 * f + 6)));
        S->surface_springs[f].resize(S->n);
    }
}

void or_sys_rebuild_neighbors(void *s, int brute) { static_cast<dem_system *>(s)->rebuild_neighbors(brute != 0); }
void or_sys_compute_accelerations(void *s) { static_cast<dem_system *>(s)->compute_accelerations(); }

// nsteps of: [rebuild when step_index % rebuild_every == 0] -> step -> advance facets.
void or_sys_step(void *s, double dt, uint64_t nsteps, uint32_t rebuild_every, uint64_t *step_index) {
    auto *S = static_cast<dem_system *>(s);
    for (uint64_t k = 0; k < nsteps; k++) {
        if (rebuild_every && S->p.neighbor_mode == 1 && (*step_index % rebuild_every) == 0)
            S->rebuild_neighbors(false);
        S->step(dt);
        if (!S->facets.empty()) S->advance_facets(dt);
        (*step_index)++;
    }
}

void or_sys_get_state(void *s, double *x, double *v, double *th, double *w, double *a, double *al) {
    auto *S = static_cast<dem_system *>(s);
    for (size_t i = 0; i < S->n; i++) {
        if (x) st(x + 3 * i, S->x[i]);
        if (v) st(v + 3 * i, S->v[i]);
        if (th) st(th + 3 * i, S->th[i]);
        if (w) st(w + 3 * i, S->w[i]);
        if (a) st(a + 3 * i, S->a[i]);
        if (al) st(al + 3 * i, S->al[i]);
    }
}

void or_sys_set_state(void *s, const double *x, const double *v, const double *th, const double *w) {
    auto *S = static_cast<dem_system *>(s);
    for (size_t i = 0; i < S->n; i++) {
        if (x) S->x[i] = ld(x + 3 * i);
        if (v) S->v[i] = ld(v + 3 * i);
        if (th) S->th[i] = ld(th + 3 * i);
        if (w) S->w[i] = ld(w + 3 * i);
    }
}

double or_sys_time(void *s) { return static_cast<dem_system *>(s)->t; }

// CSR neighbour list; returns the total entry count (idx written when it fits).
size_t or_sys_get_neighbors(void *s, uint64_t *offsets, uint32_t *idx, size_t cap) {
    auto *S = static_cast<dem_system *>(s);
    size_t tot = 0;
    for (size_t i = 0; i < S->n; i++) {
        offsets[i] = tot;
        for (uint32_t j : S->nbr[i]) {
            if (tot < cap) idx[tot] = j;
            tot++;
        }
    }
    offsets[S->n] = tot;
    return tot;
}

size_t or_sys_get_bonds(void *s, uint32_t *pairs, size_t cap, size_t *skipped) {
    auto *S = static_cast<dem_system *>(s);
    *skipped = S->bonds_skipped;
    for (size_t b = 0; b < S->bonds.size() && b < cap; b++) {
        pairs[2 * b] = S->bonds[b].first;
        pairs[2 * b + 1] = S->bonds[b].second;
    }
    return S->bonds.size();
}

// which = 0: contact springs, 1: sinter (bond) springs. Entries row-major by
// (i, j ascending). Returns the entry count.
size_t or_sys_get_springs(void *s, int which, uint32_t *ij, double *springs, size_t cap) {
    auto *S = static_cast<dem_system *>(s);
    auto &rows = which == 0 ? S->contact_springs.rows : S->sinter_springs.rows;
    size_t tot = 0;
    for (size_t i = 0; i < S->n; i++)
        for (auto &e : rows[i]) {
            if (tot < cap) {
                ij[2 * tot] = uint32_t(i);
                ij[2 * tot + 1] = e.first;
                st_spring(springs + 9 * tot, e.second);
            }
            tot++;
        }
    return tot;
}

// sinter_functor's constructor (sinter_bridge.h:146-211) on positions x0:
// the random unit vectors come from `vecs` (nvecs draws of 3 doubles) in
// order. Returns the bond count; pairs (i < j) and connectors (30 doubles per
// bond: i's five points, then j's) are written when cap allows. *used gets
// the number of draws taken, or nvecs + 1 when they ran out (the frames are
// then not valid).
size_t or_legacy_frames(size_t n, const double *x0, double r_part, double d_crit, const double *vecs, size_t nvecs,
                        uint32_t *pairs, double *conn, size_t cap, size_t *used) {
    std::vector<v3> x(n);
    for (size_t i = 0; i < n; i++) x[i] = ld(x0 + 3 * i);
    size_t k = 0;
    bool out_of_draws = false;
    auto next = [&]() -> v3 {
        if (k >= nvecs) {
            out_of_draws = true;
            return v3{1.0, 0.0, 0.0};
        }
        v3 r = ld(vecs + 3 * k);
        k++;
        return r;
    };
    std::vector<legacy_bond> b;
    legacy_sinter_frames(n, x, r_part, d_crit, next, b);
    *used = out_of_draws ? nvecs + 1 : k;
    for (size_t q = 0; q < b.size() && q < cap; q++) {
        pairs[2 * q] = b[q].i;
        pairs[2 * q + 1] = b[q].j;
        for (int c = 0; c < 5; c++) {
            st(conn + 30 * q + 3 * c, b[q].ci[c]);
            st(conn + 30 * q + 15 + 3 * c, b[q].cj[c]);
        }
    }
    return b.size();
}

// Gives a BIN_SINTER_LEGACY system its bonds and connectors (the layout of
// or_legacy_frames).
void or_sys_legacy_set(void *s, size_t nb, const uint32_t *pairs, const double *conn) {
    auto *S = static_cast<dem_system *>(s);
    std::vector<legacy_bond> b(nb);
    for (size_t q = 0; q < nb; q++) {
        b[q].i = pairs[2 * q];
        b[q].j = pairs[2 * q + 1];
        for (int c = 0; c < 5; c++) {
            b[q].ci[c] = ld(conn + 30 * q + 3 * c);
            b[q].cj[c] = ld(conn + 30 * q + 15 + 3 * c);
        }
    }
    S->set_legacy(b);
}

// The current connectors of every bond, in the order they were set.
void or_sys_legacy_get(void *s, double *conn) {
    auto *S = static_cast<dem_system *>(s);
    for (size_t q = 0; q < S->legacy.size(); q++)
        for (int c = 0; c < 5; c++) {
            st(conn + 30 * q + 3 * c, S->legacy[q].ci[c]);
            st(conn + 30 * q + 15 + 3 * c, S->legacy[q].cj[c]);
        }
}

// sinter_functor at functor level, as oracle/ref_harness/ref_legacy.cpp
// drives the reference: frames from x0 and `vecs`; then per step, for each
// particle p in order, its increments (12 doubles per particle per step: dx,
// dtheta, and the v, omega the pairs see) through the step handler
// (increment_x, then increment_theta), and every ordered pair (i, j), i != j,
// evaluated: out[((k*n + i)*n + j)*6] = linear, angular acceleration (zero
// for i == j). *nbonds and *used as or_legacy_frames.
void or_legacy_seq(size_t n, const double *x0, const double *th0, const double *s13, const double *c18,
                   const double *vecs, size_t nvecs, size_t nsteps, const double *incs, double *out, size_t *nbonds,
                   size_t *used) {
    sinter_params sp = sp_from(s13);
    contact_params cp = cp_from(c18);
    std::vector<v3> x(n), th(n), v(n), w(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = ld(x0 + 3 * i);
        th[i] = ld(th0 + 3 * i);
    }
    size_t k = 0;
    bool out_of_draws = false;
    auto next = [&]() -> v3 {
        if (k >= nvecs) {
            out_of_draws = true;
            return v3{1.0, 0.0, 0.0};
        }
        return ld(vecs + 3 * (k++));
    };
    std::vector<legacy_bond> b;
    legacy_sinter_frames(n, x, cp.r_part, sp.d_crit, next, b);
    *nbonds = b.size();
    *used = out_of_draws ? nvecs + 1 : k;
    std::vector<int64_t> bond_of(n * n, -1);
    for (size_t q = 0; q < b.size(); q++) {
        bond_of[b[q].i * n + b[q].j] = int64_t(q);
        bond_of[b[q].j * n + b[q].i] = int64_t(q);
    }
    auto side = [&](size_t i, size_t q) { return b[q].i == i ? b[q].ci : b[q].cj; };
    std::vector<spring3> springs(n * n, spring3{{vzero, vzero, vzero}});
    for (size_t step = 0; step < nsteps; step++) {
        for (size_t p = 0; p < n; p++) {
            const double *in = incs + 12 * (step * n + p);
            v3 dx = ld(in), dth = ld(in + 3);
            x[p] = x[p] + dx;
            for (size_t q = 0; q < b.size(); q++)
                if (b[q].i == p || b[q].j == p) legacy_move_x(side(p, q), dx);
            th[p] = th[p] + dth;
            for (size_t q = 0; q < b.size(); q++)
                if (b[q].i == p || b[q].j == p) legacy_move_theta(side(p, q), x[p], dth);
            v[p] = ld(in + 6);
            w[p] = ld(in + 9);
        }
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                double *o = out + ((step * n + i) * n + j) * 6;
                acc2 e{vzero, vzero};
                if (i != j) {
                    const int64_t q = bond_of[i * n + j];
                    if (q >= 0) {
                        e = legacy_bonded_eval(sp.k, sp.gamma_n, cp.r_part, cp.mass, cp.inertia, side(i, size_t(q)),
                                               side(j, size_t(q)), x[i], x[j], v[i], v[j], w[i], w[j]);
                    } else {
                        bool r;
                        e = contact_eval(cp, x[i], x[j], v[i], v[j], w[i], w[j], springs[i * n + j], r);
                    }
                }
                st(o, e.lin);
                st(o + 3, e.ang);
            }
    }
}

// Energy ledger terms, 8 per particle (dem_system::energy_terms).
void or_sys_energy_terms(void *s, double mass, double inertia, double *out) {
    static_cast<dem_system *>(s)->energy_terms(mass, inertia, out);
}

// Dissipated-work ledger: enable zeroes and starts it; out[5] = normal,
// tangential, rolling, torsion and total, summed in particle order.
void or_sys_set_dissipation_tracking(void *s, int enable) {
    auto *S = static_cast<dem_system *>(s);
    if (enable) S->diss.assign(4 * S->n, 0.0);
    S->track_diss = enable != 0;
}
void or_sys_dissipated_work(void *s, double *out) {
    auto *S = static_cast<dem_system *>(s);
    for (int c = 0; c < 4; c++) {
        double sum = 0.0;
        for (size_t i = 0; i < S->diss.size() / 4; i++) sum = sum + S->diss[4 * i + c];
        out[c] = sum;
    }
    out[4] = ((out[0] + out[1]) + out[2]) + out[3];
}

// Per particle: sum over evaluated partners and over the functors of each
// pair of |functor acceleration| (linear, angular), without changing any
// spring: the scale of the terms a rounded fold adds, including terms that
// cancel inside one pair (a bond holding a clamped Hamaker pull). out = 2*n.
void or_sys_functor_magnitudes(void *s, double *out) {
    auto *S = static_cast<dem_system *>(s);
    const long nn = long(S->n);
#pragma omp parallel for schedule(dynamic, 64)
    for (long ii = 0; ii < nn; ii++) {
        size_t i = size_t(ii);
        double sl = 0.0, sa = 0.0;
        auto add = [&](uint32_t j) {
            acc2 b, h;
            S->pair_peek_parts(i, j, b, h);
            sl += norm(b.lin) + norm(h.lin);
            sa += norm(b.ang) + norm(h.ang);
        };
        if (S->p.neighbor_mode == 1) {
            for (uint32_t j : S->nbr[i]) add(j);
        } else {
            for (uint32_t j = 0; j < uint32_t(S->n); j++)
                if (j != i) add(j);
        }
        out[2 * i] = sl;
        out[2 * i + 1] = sa;
    }
}

// Per particle: sum over evaluated partners of |a_ij| (linear) and |alpha_ij|
// (angular) at the current state, without changing any spring. out = 2*n.
void or_sys_pair_magnitudes(void *s, double *out) {
    auto *S = static_cast<dem_system *>(s);
    const long nn = long(S->n);
#pragma omp parallel for schedule(dynamic, 64)
    for (long ii = 0; ii < nn; ii++) {
        size_t i = size_t(ii);
        double sl = 0.0, sa = 0.0;
        auto add = [&](uint32_t j) {
            acc2 e = S->pair_peek(i, j);
            sl += norm(e.lin);
            sa += norm(e.ang);
        };
        if (S->p.neighbor_mode == 1) {
            for (uint32_t j : S->nbr[i]) add(j);
        } else {
            for (uint32_t j = 0; j < uint32_t(S->n); j++)
                if (j != i) add(j);
        }
        out[2 * i] = sl;
        out[2 * i + 1] = sa;
    }
}

// test/compute_energy.cpp:7-58 and test/mass_distribution.cpp:7-27.
// out: ke_total, ke_trans, ke_rot, |P|, |L|, rg, cm[3]
void or_sys_diagnostics(void *s, double mass, double inertia, double *out) {
    auto *S = static_cast<dem_system *>(s);
    double ke = 0.0;
    for (size_t i = 0; i < S->n; i++) {
        ke += mass * dot(S->v[i], S->v[i]) / 2.0;
        ke += inertia * dot(S->w[i], S->w[i]) / 2.0;
    }
    double ket = 0.0, ker = 0.0;
    for (size_t i = 0; i < S->n; i++) ket += mass * dot(S->v[i], S->v[i]);
    for (size_t i = 0; i < S->n; i++) ker += inertia * dot(S->w[i], S->w[i]);
    v3 P = vzero;
    for (size_t i = 0; i < S->n; i++) P = P + S->v[i];
    v3 L = vzero;
    for (size_t i = 0; i < S->n; i++) L = L + inertia * S->w[i];
    for (size_t i = 0; i < S->n; i++) L = L + cross(S->x[i], mass * S->v[i]);
    v3 cm = vzero;
    for (size_t i = 0; i < S->n; i++) cm = cm + S->x[i];
    cm = cm / double(S->n);
    double rg = 0.0;
    for (size_t i = 0; i < S->n; i++) {
        v3 d = S->x[i] - cm;
        rg += dot(d, d);
    }
    out[0] = ke;
    out[1] = ket / 2.0;
    out[2] = ker / 2.0;
    out[3] = mass * norm(P);
    out[4] = norm(L);
    out[5] = std::sqrt(rg / double(S->n));
    st(out + 6, cm);
}

}  // extern "C"
