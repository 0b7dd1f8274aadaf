// triangular_facet: a moving rigid triangle acting on every particle through
// its surface functors. Same template, constructor and members as the
// reference (include/libgran/surface_force/triangular_facet.h:10-70). The
// closest point and the surface functors run on the GPU
// (dem_math.h: closest_point_on_triangle, facet_fold); update_positions moves
// the host copy and, once a system drives this facet, the device copy with
// the current v_facet (hg_advance_facet), exactly as the reference driver does.
#ifndef TRIANGULAR_FACET_H
#define TRIANGULAR_FACET_H

#include <cmath>
#include <tuple>
#include <vector>

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t, typename... surface_force_functors_t>
struct triangular_facet {
    static constexpr int hg_kind = libgran_amd::kind_facet;
    typedef std::vector<field_value_t> field_container_t;

    explicit triangular_facet(field_value_t v_facet, std::tuple<field_value_t, field_value_t, field_value_t> vertices,
                              surface_force_functors_t &...force_functors)
        : v_facet(std::move(v_facet)), vertices(std::move(vertices)), surface_force_functors{force_functors...} {
        static_assert(sizeof...(surface_force_functors_t) > 0,
                      "at least one surface force functor must be provided as a template parameter");
        static_assert(((libgran_amd::kind_of<surface_force_functors_t>::value == libgran_amd::kind_surface_contact ||
                        libgran_amd::kind_of<surface_force_functors_t>::value == libgran_amd::kind_surface_hamaker) &&
                       ...),
                      "the device path implements surface_contact_force_functor and surface_hamaker_functor");
        double u[3], w[3], c[3];
        for (int k = 0; k < 3; k++) {
            u[k] = double(std::get<1>(this->vertices)[k]) - double(std::get<0>(this->vertices)[k]);
            w[k] = double(std::get<2>(this->vertices)[k]) - double(std::get<0>(this->vertices)[k]);
        }
        c[0] = u[1] * w[2] - u[2] * w[1];
        c[1] = u[2] * w[0] - u[0] * w[2];
        c[2] = u[0] * w[1] - u[1] * w[0];
        double l = std::sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
        plane_unit_normal = field_value_t(c[0] / l, c[1] / l, c[2] / l);
    }

    field_value_t const &get_unit_normal() const { return plane_unit_normal; }
    std::tuple<field_value_t, field_value_t, field_value_t> const &get_vertices() const { return vertices; }

    // Move the facet according to its velocity.
    void update_positions(real_t dt) {
        std::get<0>(vertices) += v_facet * dt;
        std::get<1>(vertices) += v_facet * dt;
        std::get<2>(vertices) += v_facet * dt;
        if (attached) {
            double v[3];
            libgran_amd::put3(v, v_facet);
            libgran_amd::check(hg_advance_facet(attached, index, v, double(dt)), "triangular_facet::update_positions");
        }
    }

    void configure(hg_params &p) const {
        std::apply([&](auto &...f) { (f.configure(p), ...); }, surface_force_functors);
    }
    void attach(hg_system *sys, size_t idx) {
        attached = sys;
        index = idx;
    }

    field_value_t v_facet;

private:
    std::tuple<field_value_t, field_value_t, field_value_t> vertices;
    std::tuple<surface_force_functors_t &...> surface_force_functors;
    field_value_t plane_unit_normal;
    hg_system *attached = nullptr;
    size_t index = 0;
};

#endif
