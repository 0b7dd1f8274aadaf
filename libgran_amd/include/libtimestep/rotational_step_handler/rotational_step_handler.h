// libtimestep's default step handler: increments positions, velocities,
// orientations and angular velocities (Custom-step-handler.md:82-145). The
// device integrator applies these increments itself; this class keeps the
// documented interface so handlers that forward to it compile.
#ifndef LIBTIMESTEP_ROTATIONAL_STEP_HANDLER_H
#define LIBTIMESTEP_ROTATIONAL_STEP_HANDLER_H

#include <cstddef>

template <typename field_container_t, typename field_value_t>
struct rotational_step_handler {
    using it = typename field_container_t::iterator;
    using cit = typename field_container_t::const_iterator;
    void increment_x(size_t n, field_value_t const &dx, it x, cit, cit, cit, cit, cit) { x[n] += dx; }
    void increment_v(size_t n, field_value_t const &dv, cit, it v, cit, cit, cit, cit) { v[n] += dv; }
    void increment_theta(size_t n, field_value_t const &dth, cit, cit, cit, it th, cit, cit) { th[n] += dth; }
    void increment_omega(size_t n, field_value_t const &dw, cit, cit, cit, cit, it w, cit) { w[n] += dw; }
};

#endif
