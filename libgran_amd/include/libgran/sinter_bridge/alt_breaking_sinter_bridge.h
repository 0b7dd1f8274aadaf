// alt_breaking_sinter_functor: bonded (sintered) contact for every touching
// pair found at construction, bonds that break under load, and the embedded
// contact functor for all other pairs. Same template and constructor as the
// reference (include/libgran/sinter_bridge/alt_breaking_sinter_bridge.h:15-38).
// The bond search runs on the host (hg_breaking_sinter_topology: every pair
// i < j within critical_separation, no cycle check); one neck strength is drawn
// per bond, in that order, as the reference's constructor draws them; the
// model runs on the GPU (params.sinter_breaking).
//
// Differences from the reference, both deliberate: each bond keeps its own
// draw (the reference stores it at neck_strengths[i] and [j] and reads
// [i*n + j], alt_breaking_sinter_bridge.h:71-72, 118; neck_strengths here is
// laid out the way that read expects), and the include guard is this file's
// own (the reference reuses alt_sinter_bridge.h's, so the two cannot be
// included together there).
#ifndef LIBGRAN_ALT_BREAKING_SINTER_BRIDGE_H
#define LIBGRAN_ALT_BREAKING_SINTER_BRIDGE_H

#include <algorithm>
#include <tuple>
#include <vector>

#include "../contact_force/contact_force.h"
#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t, typename neck_strength_distribution_t>
struct alt_breaking_sinter_functor {
    static constexpr int hg_kind = libgran_amd::kind_sinter;

    // Read access to the current bond state as the reference's dense
    // bonded_contacts[i * n_part + j]: the ordered pairs still bonded on the
    // device, refreshed whenever the system has stepped since the last read.
    struct bond_table {
        const alt_breaking_sinter_functor *owner = nullptr;
        bool operator[](size_t idx) const {
            owner->refresh();
            const size_t n = owner->n_part;
            return std::binary_search(owner->current.begin(), owner->current.end(),
                                      uint64_t(idx / n) << 32 | uint64_t(idx % n));
        }
        size_t size() const { return owner->n_part * owner->n_part; }
    };

    alt_breaking_sinter_functor(size_t n_part, std::vector<field_value_t> x0, real_t k, real_t gamma_n, real_t k_t,
                                real_t gamma_t, real_t k_r, real_t gamma_r, real_t k_o, real_t gamma_o,
                                real_t r_part, real_t mass, real_t inertia, real_t dt,
                                field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]],
                                real_t critical_separation, neck_strength_distribution_t neck_strength_distribution,
                                contact_force_functor<field_value_t, real_t> contact_force)
        : n_part(n_part),
          params{double(k), double(gamma_n), double(k_t), double(gamma_t), double(k_r), double(gamma_r), double(k_o),
                 double(gamma_o), double(r_part), double(mass), double(inertia), double(dt),
                 double(critical_separation)},
          contact(contact_force.params) {
        std::vector<double> xd;
        libgran_amd::to_doubles(x0, xd);
        size_t nb = 0;
        libgran_amd::check(hg_breaking_sinter_topology(x0.size(), xd.data(), params.r_part, params.d_crit, nullptr, 0,
                                                       &nb),
                           "alt_breaking_sinter_functor");
        bonds.resize(2 * nb);
        libgran_amd::check(hg_breaking_sinter_topology(x0.size(), xd.data(), params.r_part, params.d_crit,
                                                       bonds.data(), nb, &nb),
                           "alt_breaking_sinter_functor");
        neck_strengths.assign(n_part * n_part, real_t(0));
        strengths.resize(nb);
        for (size_t b = 0; b < nb; b++) {
            const real_t s = neck_strength_distribution();
            strengths[b] = double(s);
            neck_strengths[size_t(bonds[2 * b]) * n_part + bonds[2 * b + 1]] = s;
            neck_strengths[size_t(bonds[2 * b + 1]) * n_part + bonds[2 * b]] = s;
            current.push_back(uint64_t(bonds[2 * b]) << 32 | bonds[2 * b + 1]);
            current.push_back(uint64_t(bonds[2 * b + 1]) << 32 | bonds[2 * b]);
        }
        std::sort(current.begin(), current.end());
        bonded_contacts.owner = this;
    }

    alt_breaking_sinter_functor(alt_breaking_sinter_functor const &o)
        : bonded_contacts(o.bonded_contacts), neck_strengths(o.neck_strengths), n_part(o.n_part),
          params(o.params), contact(o.contact), bonds(o.bonds), strengths(o.strengths), current(o.current),
          attached(o.attached), seen_step(o.seen_step) {
        bonded_contacts.owner = this;
    }

    void configure(hg_params &p) const {
        p.binary_model = HG_BINARY_SINTER;
        p.sinter_breaking = 1;
        p.sinter = params;
        p.contact = contact;
    }

    void attach(hg_system *sys) {
        attached = sys;
        libgran_amd::check(hg_set_neck_strengths(sys, strengths.data()), "alt_breaking_sinter_functor");
    }

    bond_table bonded_contacts;
    std::vector<real_t> neck_strengths;  // [i * n_part + j] = strength of bond {i, j}

    size_t n_part;
    hg_sinter_params params;
    hg_contact_params contact;
    std::vector<uint32_t> bonds;     // i < j pairs, lexicographic
    std::vector<double> strengths;   // one per bond, same order

private:
    // Re-reads the still-bonded ordered pairs from the device after a step.
    void refresh() const {
        if (!attached) return;
        double t = 0.0;
        uint64_t k = 0;
        libgran_amd::check(hg_get_time(attached, &t, &k), "bonded_contacts");
        if (k == seen_step) return;
        size_t total = 0;
        libgran_amd::check(hg_get_springs(attached, 1, nullptr, nullptr, 0, &total), "bonded_contacts");
        std::vector<uint32_t> ij(2 * total);
        std::vector<double> s(9 * total);
        libgran_amd::check(hg_get_springs(attached, 1, ij.data(), s.data(), total, &total), "bonded_contacts");
        current.clear();
        for (size_t e = 0; e < total; e++) current.push_back(uint64_t(ij[2 * e]) << 32 | ij[2 * e + 1]);
        std::sort(current.begin(), current.end());
        seen_step = k;
    }

    mutable std::vector<uint64_t> current;  // sorted (i << 32 | j) of bonded ordered pairs
    hg_system *attached = nullptr;
    mutable uint64_t seen_step = 0;
};

#endif  // LIBGRAN_ALT_BREAKING_SINTER_BRIDGE_H
