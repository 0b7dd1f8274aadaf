// simplified_hamaker_functor: 2 A r / h^2 attraction along the centre line,
// h saturated at h0. Same template and constructor as the reference
// (include/libgran/hamaker_force/simplified_hamaker_force.h:8-16); evaluated on
// the GPU (dem_math.h: simplified_hamaker_eval).
#ifndef LIBGRAN_SIMPLIFIED_HAMAKER_FORCE_H
#define LIBGRAN_SIMPLIFIED_HAMAKER_FORCE_H

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct simplified_hamaker_functor {
    static constexpr int hg_kind = libgran_amd::kind_simplified_hamaker;

    simplified_hamaker_functor(real_t A, real_t h0, real_t r_part, real_t mass,
                               field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]])
        : params{double(A), double(h0), 0.0, double(r_part), double(mass)} {}

    void configure(hg_params &p) const {
        p.hamaker_model = HG_HAMAKER_SIMPLIFIED;
        p.hamaker = params;
    }

    hg_hamaker_params params;
};

#endif
