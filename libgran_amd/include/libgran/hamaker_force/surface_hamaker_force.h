// surface_hamaker_functor: sphere-plane van der Waals attraction. Same
// template and constructor as the reference
// (include/libgran/hamaker_force/surface_hamaker_force.h:8-18); evaluated on
// the GPU (dem_math.h: surface_hamaker_eval).
#ifndef SURFACE_HAMAKER_FORCE_H
#define SURFACE_HAMAKER_FORCE_H

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct surface_hamaker_functor {
    static constexpr int hg_kind = libgran_amd::kind_surface_hamaker;

    surface_hamaker_functor(real_t A, real_t h0, real_t r_part, real_t mass, field_value_t field_zero [[maybe_unused]],
                            real_t real_zero [[maybe_unused]])
        : params{double(A), double(h0), 0.0, double(r_part), double(mass)} {}

    void configure(hg_params &p) const {
        p.facet_hamaker = 1;
        p.surface_hamaker = params;
    }

    hg_hamaker_params params;
};

#endif
