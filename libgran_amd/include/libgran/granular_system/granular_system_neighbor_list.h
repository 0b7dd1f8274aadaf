// granular_system_neighbor_list: as granular_system, but pairs come from a
// neighbour list of radius r_verlet that the driver refreshes with
// update_neighbor_list(). Same template parameters and constructor as the
// reference (include/libgran/granular_system/granular_system_neighbor_list.h:26-87);
// the list is built on the GPU by a cell list (kernels.hip: k_nl_fill).
#ifndef LIBGRAN_GRANULAR_SYSTEM_NEIGHBOR_LIST_H
#define LIBGRAN_GRANULAR_SYSTEM_NEIGHBOR_LIST_H

#include "../utility/operator_overloads.h"
#include "binary_force_container.h"
#include "device_system.h"
#include "unary_force_container.h"

template <typename field_value_t, typename real_t,
          template <typename _field_container_t, typename _field_value_t, typename _real_t, typename _functor_t,
                    template <typename __field_container_t, typename __field_value_t> typename _step_handler_t>
          typename integrator_t,
          template <typename _field_container_t, typename _field_value_t> typename step_handler_t,
          typename binary_force_functor_container_t, typename unary_force_functor_container_t>
class granular_system_neighbor_list
    : public libgran_amd::device_system<field_value_t, real_t, binary_force_functor_container_t,
                                        unary_force_functor_container_t> {
    using base = libgran_amd::device_system<field_value_t, real_t, binary_force_functor_container_t,
                                            unary_force_functor_container_t>;
    static_assert(libgran_amd::is_default_integrator<integrator_t>::value,
                  "the device path implements rotational_velocity_verlet_half");
    static_assert(libgran_amd::step_handler_ok<step_handler_t, field_value_t, base::uses_legacy_sinter>(),
                  "the device path implements the default rotational_step_handler, and sinter_functor's own "
                  "step handler together with sinter_functor");

public:
    typedef std::vector<field_value_t> field_container_t;

    granular_system_neighbor_list(granular_system_neighbor_list const &) = delete;

    granular_system_neighbor_list(long n_part [[maybe_unused]], real_t r_verlet, field_container_t x0,
                                  field_container_t v0, field_container_t theta0, field_container_t omega0, real_t t0,
                                  field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]],
                                  step_handler_t<field_container_t, field_value_t> &step_handler [[maybe_unused]],
                                  binary_force_functor_container_t binary_force_functors,
                                  unary_force_functor_container_t unary_force_functors)
        : base(HG_PAIRS_NEIGHBOR_LIST, double(r_verlet), x0, v0, theta0, omega0, t0, binary_force_functors,
               unary_force_functors) {}

    void update_neighbor_list() { this->rebuild(); }
};

#endif  // LIBGRAN_GRANULAR_SYSTEM_NEIGHBOR_LIST_H
