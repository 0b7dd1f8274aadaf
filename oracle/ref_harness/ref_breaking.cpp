// Drives the reference alt_breaking_sinter_functor (alt_breaking_sinter_bridge.h)
// on the ordered pair (0, 1) of a two-particle system; layout matches
// or_breaking_seq. Its own translation unit: the header's include guard is the
// one alt_sinter_bridge.h uses.
#include "prelude.h"
#include <libgran/sinter_bridge/alt_breaking_sinter_bridge.h>

namespace {
struct constant_strength {
    double value;
    double operator()() const { return value; }
};
using contact_t = contact_force_functor<rvec, double>;
using breaking_t = alt_breaking_sinter_functor<rvec, double, constant_strength>;
}  // namespace

// The functor reads bond (0, 1)'s strength at neck_strengths[0 * 2 + 1], which
// its constructor fills (it writes neck_strengths[i] and neck_strengths[j] for
// bond {i, j} = {0, 1}), so this direction sees the drawn strength.
extern "C" void ref_breaking_seq(const double *s, const double *c, double strength, size_t nsteps,
                                 const double *states, double *out, int *broken_out) {
    pair_state st;
    st.load(states);
    contact_t contact(2, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13],
                      c[14], c[15], c[16], c[17], rvec(), 0.0);
    breaking_t f(2, st.x, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10], s[11], rvec(), 0.0,
                 s[12], constant_strength{strength}, contact);
    for (size_t k = 0; k < nsteps; k++) {
        st.load(states + 18 * k);
        auto e01 = f(0, 1, st.x, st.v, st.th, st.w, 0.0);
        rv_st(out + 6 * k, e01.first);
        rv_st(out + 6 * k + 3, e01.second);
        broken_out[k] = f.bonded_contacts[0 * 2 + 1] ? 0 : 1;
    }
}
