// sinter_functor and sinter_step_handler: the legacy sinter bridge. Bonds
// between every pair whose gap is within critical_separation, each held by
// five springs between connector points that move with the two particles;
// the embedded contact functor for all other pairs. Same templates and
// constructor as the reference (include/libgran/sinter_bridge/sinter_bridge.h:
// 13-272). The constructor draws its random unit vectors here on the host, in
// the reference's order (hg_legacy_sinter_frames); the bonded force and the
// step handler's connector moves run on the GPU (HG_BINARY_SINTER_LEGACY).
// granular_system accepts this functor together with the step handler it
// hands out, and that handler only with this functor.
#ifndef LIBGRAN_SINTER_BRIDGE_H
#define LIBGRAN_SINTER_BRIDGE_H

#include <cstdint>
#include <functional>
#include <vector>

#include "../contact_force/contact_force.h"
#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct sinter_functor;

// The step handler of the reference (sinter_bridge.h:13-94). On the device the
// integrator moves the connectors itself after each drift (increment_x, then
// increment_theta); the increments below keep the documented interface for
// host containers. Made by sinter_functor::get_step_handler.
template <typename field_container_t, typename field_value_t, typename real_t>
struct sinter_step_handler {
    static constexpr bool hg_legacy_sinter_handler = true;
    using it = typename field_container_t::iterator;

    void increment_x(size_t n, field_value_t const &dx, it x, it, it, it, it, it) const { *(x + n) += dx; }
    void increment_v(size_t n, field_value_t const &dv, it, it v, it, it, it, it) const { *(v + n) += dv; }
    void increment_theta(size_t n, field_value_t const &dth, it, it, it, it th, it, it) const { *(th + n) += dth; }
    void increment_omega(size_t n, field_value_t const &dw, it, it, it, it, it w, it) const { *(w + n) += dw; }

private:
    friend struct sinter_functor<field_value_t, real_t>;
    sinter_step_handler() = default;
};

template <typename field_value_t, typename real_t>
struct sinter_functor {
    static constexpr int hg_kind = libgran_amd::kind_legacy_sinter;

    sinter_functor(size_t n_part, real_t k_d, real_t gamma_d, real_t k, real_t gamma_n, real_t k_t, real_t gamma_t,
                   real_t mu_s, real_t phi_d, real_t k_r, real_t gamma_r, real_t mu_r, real_t phi_r, real_t k_o,
                   real_t gamma_o, real_t mu_o, real_t phi_o, real_t r_part, real_t mass, real_t inertia, real_t dt,
                   field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]],
                   typename std::vector<field_value_t>::const_iterator x0_begin_itr,
                   std::function<field_value_t()> random_unit_vector_generator, real_t critical_separation)
        : n_part(n_part),
          contact{double(k), double(gamma_n), double(k_t), double(gamma_t), double(mu_s), double(phi_d),
                  double(k_r), double(gamma_r), double(mu_r), double(phi_r), double(k_o), double(gamma_o),
                  double(mu_o), double(phi_o), double(r_part), double(mass), double(inertia), double(dt)} {
        sinter = hg_sinter_params{};
        sinter.k = double(k_d);
        sinter.gamma_n = double(gamma_d);
        sinter.r_part = double(r_part);
        sinter.mass = double(mass);
        sinter.inertia = double(inertia);
        sinter.dt = double(dt);
        sinter.d_crit = double(critical_separation);
        std::vector<field_value_t> x0(x0_begin_itr, x0_begin_itr + n_part);
        std::vector<double> xd;
        libgran_amd::to_doubles(x0, xd);
        size_t nb = 0;
        libgran_amd::check(hg_legacy_sinter_frames(n_part, xd.data(), sinter.r_part, sinter.d_crit, nullptr, nullptr,
                                                   nullptr, nullptr, 0, &nb),
                           "sinter_functor");
        pairs.resize(2 * nb);
        connectors.resize(30 * nb);
        auto next = [](void *user, double out[3]) {
            const field_value_t v = (*static_cast<std::function<field_value_t()> *>(user))();
            for (int c = 0; c < 3; c++) out[c] = double(v[c]);
        };
        libgran_amd::check(hg_legacy_sinter_frames(n_part, xd.data(), sinter.r_part, sinter.d_crit, next,
                                                   &random_unit_vector_generator, pairs.data(), connectors.data(), nb,
                                                   &nb),
                           "sinter_functor");
    }

    // sinter_bridge.h:228-232
    template <typename field_container_t>
    sinter_step_handler<field_container_t, field_value_t, real_t> get_step_handler() {
        return sinter_step_handler<field_container_t, field_value_t, real_t>();
    }

    void configure(hg_params &p) const {
        p.binary_model = HG_BINARY_SINTER_LEGACY;
        p.sinter = sinter;
        p.contact = contact;
    }

    void attach(hg_system *sys) {
        libgran_amd::check(hg_set_legacy_connectors(sys, pairs.size() / 2, pairs.data(), connectors.data()),
                           "sinter_functor");
    }

    size_t n_part;
    hg_contact_params contact;
    hg_sinter_params sinter;
    std::vector<uint32_t> pairs;     // bonds (i < j), row order
    std::vector<double> connectors;  // 30 per bond: i's five points, then j's
};

#endif  // LIBGRAN_SINTER_BRIDGE_H
