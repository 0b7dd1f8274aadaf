// Drives the reference simplified_hamaker_functor (simplified_hamaker_force.h),
// which takes square() from hamaker_force.h, included first as in a reference
// program that uses both.
#include "prelude.h"
#include <libgran/hamaker_force/hamaker_force.h>
#include <libgran/hamaker_force/simplified_hamaker_force.h>

extern "C" void ref_simplified_hamaker_seq(const double *h5, size_t nsteps, const double *states, double *out) {
    simplified_hamaker_functor<rvec, double> f(h5[0], h5[1], h5[3], h5[4], rvec(), 0.0);
    pair_state s;
    for (size_t k = 0; k < nsteps; k++) {
        s.load(states + 18 * k);
        auto e01 = f(0, 1, s.x, s.v, s.th, s.w, 0.0);
        auto e10 = f(1, 0, s.x, s.v, s.th, s.w, 0.0);
        rv_st(out + 12 * k, e01.first);
        rv_st(out + 12 * k + 3, e01.second);
        rv_st(out + 12 * k + 6, e10.first);
        rv_st(out + 12 * k + 9, e10.second);
    }
}
