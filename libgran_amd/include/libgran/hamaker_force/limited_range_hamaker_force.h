// limited_range_hamaker_functor: Hamaker attraction cut off beyond a centre
// distance. Same template and constructor as the reference
// (include/libgran/hamaker_force/limited_range_hamaker_force.h:18-27);
// evaluated on the GPU (dem_math.h: lr_hamaker_eval).
#ifndef LIBGRAN_LIMITED_RANGE_HAMAKER_FORCE_H
#define LIBGRAN_LIMITED_RANGE_HAMAKER_FORCE_H

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct limited_range_hamaker_functor {
    static constexpr int hg_kind = libgran_amd::kind_limited_hamaker;

    limited_range_hamaker_functor(real_t A, real_t h0, real_t r_cutoff, real_t r_part, real_t mass,
                                  field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]])
        : params{double(A), double(h0), double(r_cutoff), double(r_part), double(mass)} {}

    void configure(hg_params &p) const {
        p.hamaker_model = HG_HAMAKER_LIMITED;
        p.hamaker = params;
    }

    hg_hamaker_params params;
};

#endif
