// alt_sinter_functor: bonded (sintered) contact for touching pairs found at
// construction, the embedded contact functor for all other pairs. Same
// template and constructor as the reference
// (include/libgran/sinter_bridge/alt_sinter_bridge.h:17-90). The bond search
// runs here on the host with the reference's decision rule (hg_sinter_topology:
// lexicographic pairs, gap within critical_separation, cycle-closing bonds
// skipped with the same warning); the model runs on the GPU
// (dem_math.h: sinter_eval).
#ifndef LIBGRAN_ALT_SINTER_BRIDGE_H
#define LIBGRAN_ALT_SINTER_BRIDGE_H

#include <algorithm>
#include <iostream>
#include <tuple>
#include <vector>

#include "../contact_force/contact_force.h"
#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct alt_sinter_functor {
    static constexpr int hg_kind = libgran_amd::kind_sinter;

    // Read access to the bond table as the reference's dense
    // bonded_contacts[i * n_part + j] (alt_sinter_bridge.h:163), stored
    // sparsely as sorted (i, j) keys.
    struct bond_table {
        size_t n = 0;
        std::vector<uint64_t> keys;
        bool operator[](size_t idx) const {
            return std::binary_search(keys.begin(), keys.end(), uint64_t(idx / n) << 32 | uint64_t(idx % n));
        }
        size_t size() const { return n * n; }
    };

    alt_sinter_functor(size_t n_part, std::vector<field_value_t> x0, real_t k, real_t gamma_n, real_t k_t,
                       real_t gamma_t, real_t k_r, real_t gamma_r, real_t k_o, real_t gamma_o, real_t r_part,
                       real_t mass, real_t inertia, real_t dt, field_value_t field_zero [[maybe_unused]],
                       real_t real_zero [[maybe_unused]], real_t critical_separation,
                       contact_force_functor<field_value_t, real_t> contact_force)
        : n_part(n_part),
          params{double(k), double(gamma_n), double(k_t), double(gamma_t), double(k_r), double(gamma_r), double(k_o),
                 double(gamma_o), double(r_part), double(mass), double(inertia), double(dt),
                 double(critical_separation)},
          contact(contact_force.params) {
        std::vector<double> xd;
        libgran_amd::to_doubles(x0, xd);
        size_t nb = 0, skipped = 0;
        libgran_amd::check(hg_sinter_topology(x0.size(), xd.data(), params.r_part, params.d_crit, nullptr, 0, &nb,
                                              &skipped),
                           "alt_sinter_functor");
        bonds.resize(2 * nb);
        libgran_amd::check(hg_sinter_topology(x0.size(), xd.data(), params.r_part, params.d_crit, bonds.data(), nb, &nb,
                                              &skipped),
                           "alt_sinter_functor");
        for (size_t s = 0; s < skipped; s++) std::cout << "Warning: preventing neck insertion to avoid a cycle" << std::endl;
        bonded_contacts.n = n_part;
        for (size_t b = 0; b < nb; b++) {
            const uint64_t i = bonds[2 * b], j = bonds[2 * b + 1];
            bonded_contacts.keys.push_back(i << 32 | j);
            bonded_contacts.keys.push_back(j << 32 | i);
        }
        std::sort(bonded_contacts.keys.begin(), bonded_contacts.keys.end());
    }

    void configure(hg_params &p) const {
        p.binary_model = HG_BINARY_SINTER;
        p.sinter = params;
        p.contact = contact;
    }

    // The bonded springs as the reference's dense per-ordered-pair table
    // (alt_sinter_bridge.h:158-161): n_part * n_part triples, read from the
    // device of the system this functor drives.
    std::vector<std::tuple<field_value_t, field_value_t, field_value_t>> get_contact_springs() const {
        std::vector<std::tuple<field_value_t, field_value_t, field_value_t>> out(
            n_part * n_part, std::make_tuple(field_value_t(0, 0, 0), field_value_t(0, 0, 0), field_value_t(0, 0, 0)));
        if (!attached) return out;
        size_t total = 0;
        libgran_amd::check(hg_get_springs(attached, 1, nullptr, nullptr, 0, &total), "get_contact_springs");
        std::vector<uint32_t> ij(2 * total);
        std::vector<double> s(9 * total);
        libgran_amd::check(hg_get_springs(attached, 1, ij.data(), s.data(), total, &total), "get_contact_springs");
        for (size_t e = 0; e < total; e++) {
            const double *q = &s[9 * e];
            out[size_t(ij[2 * e]) * n_part + ij[2 * e + 1]] =
                std::make_tuple(field_value_t(q[0], q[1], q[2]), field_value_t(q[3], q[4], q[5]),
                                field_value_t(q[6], q[7], q[8]));
        }
        return out;
    }

    void attach(hg_system *sys) { attached = sys; }

    size_t n_part;
    hg_sinter_params params;
    hg_contact_params contact;
    std::vector<uint32_t> bonds;
    bond_table bonded_contacts;
    hg_system *attached = nullptr;
};

#endif  // LIBGRAN_ALT_SINTER_BRIDGE_H
