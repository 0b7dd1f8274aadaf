// Drives the reference limited_range_hamaker_functor alone and inside a
// binary_force_functor_container together with alt_sinter_functor (the
// functor set of the aggregate configurations).
#include "prelude.h"
#include <libgran/utility/operator_overloads.h>
#include <libgran/granular_system/binary_force_container.h>
#include <libgran/hamaker_force/limited_range_hamaker_force.h>
#include <libgran/sinter_bridge/alt_sinter_bridge.h>

using contact_t = contact_force_functor<rvec, double>;
using sinter_t = alt_sinter_functor<rvec, double>;
using lrh_t = limited_range_hamaker_functor<rvec, double>;

extern "C" void ref_lr_hamaker_seq(const double *h5, size_t nsteps, const double *states, double *out) {
    lrh_t f(h5[0], h5[1], h5[2], h5[3], h5[4], rvec(), 0.0);
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

extern "C" void ref_fold_seq(const double *s, const double *c, const double *h5, size_t nsteps,
                             const double *states, double *out) {
    pair_state ps;
    ps.load(states);
    contact_t contact(2, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13],
                      c[14], c[15], c[16], c[17], rvec(), 0.0);
    sinter_t sinter(2, ps.x, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10], s[11], rvec(), 0.0,
                    s[12], contact);
    lrh_t ham(h5[0], h5[1], h5[2], h5[3], h5[4], rvec(), 0.0);
    binary_force_functor_container<rvec, double, sinter_t, lrh_t> box{sinter, ham};
    for (size_t k = 0; k < nsteps; k++) {
        ps.load(states + 18 * k);
        auto e01 = box(0, 1, ps.x, ps.v, ps.th, ps.w, 0.0);
        auto e10 = box(1, 0, ps.x, ps.v, ps.th, ps.w, 0.0);
        rv_st(out + 12 * k, e01.first);
        rv_st(out + 12 * k + 3, e01.second);
        rv_st(out + 12 * k + 6, e10.first);
        rv_st(out + 12 * k + 9, e10.second);
    }
}
