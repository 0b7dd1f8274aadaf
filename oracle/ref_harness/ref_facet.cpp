// Drives the reference triangular_facet (triangular_facet.h) carrying a
// surface_contact_force_functor and a surface_hamaker_functor, with
// update_positions() after every step; layout matches or_facet_seq.
#include "prelude.h"
#include <libgran/utility/operator_overloads.h>
#include <libgran/hamaker_force/hamaker_force.h>  // square/cube used by surface_hamaker_force.h
#include <libgran/contact_force/surface_contact_force.h>
#include <libgran/hamaker_force/surface_hamaker_force.h>
#include <libgran/surface_force/triangular_facet.h>

using sc_t = surface_contact_force_functor<rvec, double>;
using sh_t = surface_hamaker_functor<rvec, double>;
using facet_t = triangular_facet<rvec, double, sc_t, sh_t>;

extern "C" void ref_facet_seq(const double *verts9, const double *vf3, const double *c, const double *h5,
                              double dt_advance, size_t nsteps, const double *states, double *out) {
    sc_t sc(1, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13], c[14], c[15],
            c[16], c[17], rvec(), 0.0);
    sh_t sh(h5[0], h5[1], h5[3], h5[4], rvec(), 0.0);
    facet_t facet(rv_ld(vf3), std::make_tuple(rv_ld(verts9), rv_ld(verts9 + 3), rv_ld(verts9 + 6)), sc, sh);
    std::vector<rvec> x(1), v(1), th(1), w(1);
    for (size_t k = 0; k < nsteps; k++) {
        const double *s = states + 9 * k;
        x[0] = rv_ld(s);
        v[0] = rv_ld(s + 3);
        w[0] = rv_ld(s + 6);
        auto e = facet(0, x, v, th, w, 0.0);
        rv_st(out + 6 * k, e.first);
        rv_st(out + 6 * k + 3, e.second);
        facet.update_positions(dt_advance);
    }
}
