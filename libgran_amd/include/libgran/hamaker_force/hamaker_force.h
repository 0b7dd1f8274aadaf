// hamaker_functor: sphere-sphere van der Waals attraction saturated at h0.
// Same template and constructor as the reference
// (include/libgran/hamaker_force/hamaker_force.h:18-26); evaluated on the GPU
// (dem_math.h: hamaker_eval).
#ifndef LIBGRAN_HAMAKER_FORCE_H
#define LIBGRAN_HAMAKER_FORCE_H

#include "../hipgran_bridge.h"

template <typename field_value_t, typename real_t>
struct hamaker_functor {
    static constexpr int hg_kind = libgran_amd::kind_hamaker;

    hamaker_functor(real_t A, real_t h0, real_t r_part, real_t mass, field_value_t field_zero [[maybe_unused]],
                    real_t real_zero [[maybe_unused]])
        : params{double(A), double(h0), 0.0, double(r_part), double(mass)} {}

    void configure(hg_params &p) const {
        p.hamaker_model = HG_HAMAKER_FULL;
        p.hamaker = params;
    }

    hg_hamaker_params params;
};

#endif
