// unary_force_functor_container: holds references to the unary force functors
// (triangular facets) of a system (same shape as the reference,
// include/libgran/granular_system/unary_force_container.h:10-35).
#ifndef LIBGRAN_UNARY_FORCE_CONTAINER_H
#define LIBGRAN_UNARY_FORCE_CONTAINER_H

#include <tuple>
#include <vector>

template <typename field_value_t, typename real_t, typename... unary_force_functors_t>
struct unary_force_functor_container {
    typedef std::vector<field_value_t> field_container_t;

    unary_force_functor_container(unary_force_functors_t &...force_functors) : unary_force_functors{force_functors...} {}

    std::tuple<unary_force_functors_t &...> unary_force_functors;
};

#endif
