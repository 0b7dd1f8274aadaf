// Drives the reference contact_force_functor (contact_force.h) through a
// sequence of pair states; output layout matches or_contact_seq.
#include "prelude.h"
#include <libgran/contact_force/contact_force.h>

extern "C" void ref_contact_seq(const double *p, size_t nsteps, const double *states, double *out) {
    contact_force_functor<rvec, double> f(2, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                                          p[11], p[12], p[13], p[14], p[15], p[16], p[17], rvec(), 0.0);
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
