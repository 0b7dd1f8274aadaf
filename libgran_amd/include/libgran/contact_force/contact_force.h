// contact_force_functor: Luding frictional contact with persistent
// tangential, rolling and torsion springs. Same template and constructor as
// the reference (include/libgran/contact_force/contact_force.h:8-55); the
// model runs on the GPU (libgran_amd/csrc/dem_math.h: contact_eval), with the
// springs in the device pair-state table instead of an n*n host table.
#ifndef LIBGRAN_CONTACT_FORCE_H
#define LIBGRAN_CONTACT_FORCE_H

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct contact_force_functor {
    static constexpr int hg_kind = libgran_amd::kind_contact;

    contact_force_functor(size_t n_part, real_t k, real_t gamma_n, real_t k_t, real_t gamma_t, real_t mu_s,
                          real_t phi_d, real_t k_r, real_t gamma_r, real_t mu_r, real_t phi_r, real_t k_o,
                          real_t gamma_o, real_t mu_o, real_t phi_o, real_t r_part, real_t mass, real_t inertia,
                          real_t dt, field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]])
        : n_part(n_part),
          params{double(k), double(gamma_n), double(k_t), double(gamma_t), double(mu_s), double(phi_d),
                 double(k_r), double(gamma_r), double(mu_r), double(phi_r), double(k_o), double(gamma_o),
                 double(mu_o), double(phi_o), double(r_part), double(mass), double(inertia), double(dt)} {}

    void configure(hg_params &p) const {
        p.binary_model = HG_BINARY_CONTACT;
        p.contact = params;
    }

    size_t n_part;
    hg_contact_params params;
};

#endif  // LIBGRAN_CONTACT_FORCE_H
