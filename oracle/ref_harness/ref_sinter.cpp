// Drives the reference alt_sinter_functor (alt_sinter_bridge.h) and reads its
// public bond table and spring accessor; layout matches or_sinter_seq and
// or_sinter_topology.
#include "prelude.h"
#include <libgran/sinter_bridge/alt_sinter_bridge.h>

using contact_t = contact_force_functor<rvec, double>;
using sinter_t = alt_sinter_functor<rvec, double>;

static contact_t make_contact(size_t n, const double *c) {
    return contact_t(n, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13],
                     c[14], c[15], c[16], c[17], rvec(), 0.0);
}

static sinter_t make_sinter(size_t n, const std::vector<rvec> &x0, const double *s, const double *c) {
    return sinter_t(n, x0, s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10], s[11], rvec(), 0.0,
                    s[12], make_contact(n, c));
}

extern "C" void ref_sinter_seq(const double *s13, const double *c18, size_t nsteps, const double *states,
                               double *out, double *springs_out, int *bonded_out) {
    pair_state s;
    s.load(states);
    sinter_t f = make_sinter(2, s.x, s13, c18);
    *bonded_out = f.bonded_contacts[0 * 2 + 1] ? 1 : 0;
    for (size_t k = 0; k < nsteps; k++) {
        s.load(states + 18 * k);
        auto e01 = f(0, 1, s.x, s.v, s.th, s.w, 0.0);
        auto e10 = f(1, 0, s.x, s.v, s.th, s.w, 0.0);
        rv_st(out + 12 * k, e01.first);
        rv_st(out + 12 * k + 3, e01.second);
        rv_st(out + 12 * k + 6, e10.first);
        rv_st(out + 12 * k + 9, e10.second);
        const auto &sp = f.get_contact_springs();
        for (int q = 0; q < 2; q++) {
            const auto &t = sp[q == 0 ? 1 : 2];
            rv_st(springs_out + 18 * k + 9 * q, std::get<0>(t));
            rv_st(springs_out + 18 * k + 9 * q + 3, std::get<1>(t));
            rv_st(springs_out + 18 * k + 9 * q + 6, std::get<2>(t));
        }
    }
}

extern "C" size_t ref_sinter_topology(size_t n, const double *x0, const double *s13, const double *c18,
                                      uint32_t *pairs_out, size_t cap) {
    std::vector<rvec> x(n);
    for (size_t i = 0; i < n; i++) x[i] = rv_ld(x0 + 3 * i);
    sinter_t f = make_sinter(n, x, s13, c18);
    size_t nb = 0;
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++)
            if (f.bonded_contacts[i * n + j]) {
                if (nb < cap) {
                    pairs_out[2 * nb] = uint32_t(i);
                    pairs_out[2 * nb + 1] = uint32_t(j);
                }
                nb++;
            }
    return nb;
}
