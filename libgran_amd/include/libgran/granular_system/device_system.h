// The device side of granular_system / granular_system_neighbor_list: reads
// the functor set out of the binary and unary containers at compile time,
// creates the hipgran system (include/hipgran.h) and serves the libtimestep
// system interface (do_step, get_x/get_v/get_theta/get_omega,
// reset_acceleration_buffers, update_neighbor_list) from it.
//
// Supported functor sets (anything else is a compile-time error: there is no
// host evaluation path): in the binary container at most one of
// contact_force_functor / alt_sinter_functor / sinter_functor and at most one of
// hamaker_functor / limited_range_hamaker_functor / simplified_hamaker_functor;
// in the unary container
// triangular facets carrying surface_contact_force_functor and/or
// surface_hamaker_functor with the same parameters on every facet; the
// default integrator, and the default step handler (sinter_functor's own
// step handler when the binary container holds a sinter_functor).
#ifndef LIBGRAN_DEVICE_SYSTEM_H
#define LIBGRAN_DEVICE_SYSTEM_H

#include <cstring>
#include <tuple>
#include <type_traits>
#include <vector>

#include <libtimestep/rotational_integrator/rotational_integrator.h>
#include <libtimestep/rotational_step_handler/rotational_step_handler.h>

#include "../hipgran_bridge.h"

namespace libgran_amd {

template <typename Tuple>
struct tuple_kinds;
template <typename... F>
struct tuple_kinds<std::tuple<F &...>> {
    static constexpr int count(int k) {
        (void)k;  // unused for an empty container
        return ((kind_of<F>::value == k ? 1 : 0) + ... + 0);
    }
    static constexpr bool all_known() { return ((kind_of<F>::value != 0) && ... && true); }
    static constexpr size_t size = sizeof...(F);
};

template <typename field_value_t, typename real_t, typename binary_container_t, typename unary_container_t>
class device_system {
    using BK = tuple_kinds<decltype(binary_container_t::binary_force_functors)>;
    using UK = tuple_kinds<decltype(unary_container_t::unary_force_functors)>;
    static_assert(BK::size > 0, "at least one binary force functor must be provided as a template parameter");
    static_assert(BK::all_known(), "the device path implements contact_force_functor, alt_sinter_functor, "
                                   "sinter_functor, hamaker_functor, limited_range_hamaker_functor and "
                                   "simplified_hamaker_functor only");
    static_assert(BK::count(kind_contact) + BK::count(kind_sinter) + BK::count(kind_legacy_sinter) <= 1,
                  "at most one contact-type functor (contact_force_functor, alt_sinter_functor or sinter_functor)");
    static_assert(BK::count(kind_hamaker) + BK::count(kind_limited_hamaker) + BK::count(kind_simplified_hamaker) <= 1,
                  "at most one Hamaker functor");
    static_assert(BK::count(kind_contact) + BK::count(kind_sinter) + BK::count(kind_legacy_sinter) +
                          BK::count(kind_hamaker) + BK::count(kind_limited_hamaker) +
                          BK::count(kind_simplified_hamaker) == BK::size,
                  "unsupported binary functor");
    static_assert(UK::count(kind_facet) == UK::size, "the unary container may hold triangular_facet objects only");

public:
    typedef std::vector<field_value_t> field_container_t;
    // the binary container holds the legacy sinter_functor
    static constexpr bool uses_legacy_sinter = BK::count(kind_legacy_sinter) == 1;

    device_system(int pair_mode, double r_verlet, const field_container_t &x0, const field_container_t &v0,
                  const field_container_t &theta0, const field_container_t &omega0, real_t t0,
                  binary_container_t &binary, unary_container_t &unary)
        : t(t0) {
        hg_params p;
        std::memset(&p, 0, sizeof(p));
        p.pair_mode = pair_mode;
        p.r_verlet = r_verlet;
        p.precision = HG_PRECISION_FP64;
        p.facet_advance = 1;  // facets move when the driver calls update_positions
        std::apply([&](auto &...f) { (f.configure(p), ...); }, binary.binary_force_functors);
        // facets: one set of surface parameters shared by all
        bool first = true;
        hg_params facet_p;
        std::apply(
            [&](auto &...f) {
                auto one = [&](auto &facet) {
                    hg_params q;
                    std::memset(&q, 0, sizeof(q));
                    facet.configure(q);
                    if (first) {
                        facet_p = q;
                        first = false;
                    } else if (q.facet_contact != facet_p.facet_contact || q.facet_hamaker != facet_p.facet_hamaker ||
                               std::memcmp(&q.surface_contact, &facet_p.surface_contact, sizeof(q.surface_contact)) ||
                               std::memcmp(&q.surface_hamaker, &facet_p.surface_hamaker, sizeof(q.surface_hamaker))) {
                        throw std::runtime_error("granular_system: every facet must carry the same surface functors");
                    }
                };
                (void)one;
                (one(f), ...);
            },
            unary.unary_force_functors);
        if (!first) {
            p.facet_contact = facet_p.facet_contact;
            p.facet_hamaker = facet_p.facet_hamaker;
            p.surface_contact = facet_p.surface_contact;
            p.surface_hamaker = facet_p.surface_hamaker;
        }
        std::vector<double> xd, vd, td, wd;
        to_doubles(x0, xd);
        to_doubles(v0, vd);
        to_doubles(theta0, td);
        to_doubles(omega0, wd);
        n = x0.size();
        check(hg_create(&p, n, xd.data(), vd.data(), td.data(), wd.data(), &sys), "granular_system");
        std::apply([&](auto &...f) { (attach_functor(f), ...); }, binary.binary_force_functors);
        if (UK::size > 0) {
            std::vector<double> verts, vel;
            size_t idx = 0;
            std::apply(
                [&](auto &...f) {
                    auto one = [&](auto &facet) {
                        const auto &V = facet.get_vertices();
                        double q[3];
                        put3(q, std::get<0>(V)); verts.insert(verts.end(), q, q + 3);
                        put3(q, std::get<1>(V)); verts.insert(verts.end(), q, q + 3);
                        put3(q, std::get<2>(V)); verts.insert(verts.end(), q, q + 3);
                        put3(q, facet.v_facet); vel.insert(vel.end(), q, q + 3);
                        facet.attach(sys, idx++);
                    };
                    (void)one;
                    (one(f), ...);
                },
                unary.unary_force_functors);
            check(hg_set_facets(sys, UK::size, verts.data(), vel.data()), "granular_system");
        }
        stale = true;
    }

    device_system(device_system const &) = delete;
    ~device_system() {
        if (sys) hg_destroy(sys);
    }

    void do_step(real_t dt) {
        // queued: a capacity error surfaces at the next state read
        check(hg_step_async(sys, double(dt), 1, 0), "do_step");
        t += dt;
        stale = true;
    }

    field_container_t const &get_x() { return sync().x; }
    field_container_t const &get_v() { return sync().v; }
    field_container_t const &get_theta() { return sync().theta; }
    field_container_t const &get_omega() { return sync().omega; }
    real_t get_t() const { return t; }

    void reset_acceleration_buffers() { check(hg_reset_accelerations(sys), "reset_acceleration_buffers"); }

    hg_system *native_handle() const { return sys; }

protected:
    void rebuild() {
        check(hg_rebuild_neighbors(sys), "update_neighbor_list");
    }

private:
    template <typename F>
    void attach_functor(F &f) {
        if constexpr (kind_of<F>::value == kind_sinter || kind_of<F>::value == kind_legacy_sinter) f.attach(sys);
    }

    struct cache_t {
        field_container_t x, v, theta, omega;
    };
    cache_t &sync() {
        if (stale) {
            std::vector<double> xd(3 * n), vd(3 * n), td(3 * n), wd(3 * n);
            check(hg_get_state(sys, xd.data(), vd.data(), td.data(), wd.data()), "get_x/get_v/get_theta/get_omega");
            from_doubles(xd, cache.x);
            from_doubles(vd, cache.v);
            from_doubles(td, cache.theta);
            from_doubles(wd, cache.omega);
            stale = false;
        }
        return cache;
    }

    hg_system *sys = nullptr;
    size_t n = 0;
    real_t t;
    bool stale = true;
    cache_t cache;
};

template <template <typename, typename, typename, typename, template <typename, typename> typename> typename I>
struct is_default_integrator : std::false_type {};
template <>
struct is_default_integrator<rotational_velocity_verlet_half> : std::true_type {};

template <template <typename, typename> typename H>
struct is_default_step_handler : std::false_type {};
template <>
struct is_default_step_handler<rotational_step_handler> : std::true_type {};

// sinter_functor's step handler (sinter_bridge.h), recognised on the
// instantiated type: drivers name it through an alias template that fixes
// real_t (test/sintered.cpp:46-47).
template <typename T, typename = void>
struct is_legacy_sinter_handler : std::false_type {};
template <typename T>
struct is_legacy_sinter_handler<T, std::void_t<decltype(T::hg_legacy_sinter_handler)>> : std::true_type {};

// The step handler a granular_system accepts: the default one, or with a
// sinter_functor exactly that functor's handler (the device moves its
// connectors; without them the bonds would not follow the particles).
template <template <typename, typename> typename H, typename field_value_t, bool legacy>
constexpr bool step_handler_ok() {
    if constexpr (legacy)
        return is_legacy_sinter_handler<H<std::vector<field_value_t>, field_value_t>>::value;
    else
        return is_default_step_handler<H>::value;
}

}  // namespace libgran_amd

#endif
