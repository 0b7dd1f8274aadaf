// granular_system: every ordered pair j != i through the binary functor fold,
// plus the unary (facet) functors, integrated by rotational velocity Verlet.
// Same template parameters and constructor as the reference
// (include/libgran/granular_system/granular_system.h:25-86); the whole step
// runs on the GPU through the hipgran C ABI (device_system.h).
#ifndef LIBGRAN_GRANULAR_SYSTEM_H
#define LIBGRAN_GRANULAR_SYSTEM_H

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
class granular_system : public libgran_amd::device_system<field_value_t, real_t, binary_force_functor_container_t,
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

    granular_system(granular_system const &) = delete;

    granular_system(field_container_t x0, field_container_t v0, field_container_t theta0, field_container_t omega0,
                    real_t t0, field_value_t field_zero [[maybe_unused]], real_t real_zero [[maybe_unused]],
                    step_handler_t<field_container_t, field_value_t> &step_handler [[maybe_unused]],
                    binary_force_functor_container_t binary_force_functors,
                    unary_force_functor_container_t unary_force_functors)
        : base(HG_PAIRS_ALL, 0.0, x0, v0, theta0, omega0, t0, binary_force_functors, unary_force_functors) {}
};

#endif  // LIBGRAN_GRANULAR_SYSTEM_H
