// Drives the reference sinter_functor and its sinter_step_handler
// (sinter_bridge.h) through the sequence or_legacy_seq restates: the
// constructor draws its unit vectors from `vecs`; per step each particle's
// increments go through the step handler (increment_x, then increment_theta)
// and every ordered pair is evaluated. Layout as or_legacy_seq.
#include <array>
#include <functional>

#include "prelude.h"
#include <libgran/sinter_bridge/sinter_bridge.h>

using legacy_t = sinter_functor<rvec, double>;

extern "C" void ref_legacy_seq(size_t n, const double *x0, const double *th0, const double *s13, const double *c18,
                               const double *vecs, size_t nvecs, size_t nsteps, const double *incs, double *out,
                               size_t *used) {
    std::vector<rvec> x(n), th(n), v(n), w(n), a(n), al(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = rv_ld(x0 + 3 * i);
        th[i] = rv_ld(th0 + 3 * i);
    }
    size_t k = 0;
    bool out_of_draws = false;
    std::function<rvec()> gen = [&]() {
        if (k >= nvecs) {
            out_of_draws = true;
            return rvec(1.0, 0.0, 0.0);
        }
        rvec r = rv_ld(vecs + 3 * k);
        k++;
        return r;
    };
    const double *c = c18;
    legacy_t f(n, s13[0], s13[1], c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12],
               c[13], c[14], c[15], c[16], c[17], rvec(), 0.0, x.cbegin(), gen, s13[12]);
    *used = out_of_draws ? nvecs + 1 : k;
    auto h = f.get_step_handler<std::vector<rvec>>();
    for (size_t step = 0; step < nsteps; step++) {
        for (size_t p = 0; p < n; p++) {
            const double *in = incs + 12 * (step * n + p);
            h.increment_x(p, rv_ld(in), x.begin(), v.begin(), a.begin(), th.begin(), w.begin(), al.begin());
            h.increment_theta(p, rv_ld(in + 3), x.begin(), v.begin(), a.begin(), th.begin(), w.begin(), al.begin());
            v[p] = rv_ld(in + 6);
            w[p] = rv_ld(in + 9);
        }
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                double *o = out + ((step * n + i) * n + j) * 6;
                rvec lin, ang;
                if (i != j) {
                    auto e = f(i, j, x, v, th, w, 0.0);
                    lin = e.first;
                    ang = e.second;
                }
                rv_st(o, lin);
                rv_st(o + 3, ang);
            }
    }
}
