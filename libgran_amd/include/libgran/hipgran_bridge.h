// Shared plumbing of the libgran drop-in headers: the functor kinds the
// device path understands, vector conversions and error reporting over the
// hipgran C ABI (include/hipgran.h).
//
// field_value_t is any 3-vector with operator[] and a 3-argument constructor
// (Eigen::Vector3d, or the small vec3 in libgran/vec3.h).
#ifndef LIBGRAN_HIPGRAN_BRIDGE_H
#define LIBGRAN_HIPGRAN_BRIDGE_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../../include/hipgran.h"

namespace libgran_amd {

// Functor kinds the device fold implements.
enum functor_kind : int {
    kind_contact = 1,
    kind_sinter = 2,
    kind_hamaker = 3,
    kind_limited_hamaker = 4,
    kind_surface_contact = 5,
    kind_surface_hamaker = 6,
    kind_facet = 7,
    kind_simplified_hamaker = 8,
    kind_legacy_sinter = 9
};

template <typename T, typename = void>
struct kind_of : std::integral_constant<int, 0> {};
template <typename T>
struct kind_of<T, std::void_t<decltype(T::hg_kind)>> : std::integral_constant<int, T::hg_kind> {};

inline void check(hg_status st, const char *what) {
    if (st != HG_OK) throw std::runtime_error(std::string(what) + ": " + hg_last_error());
}

template <typename field_value_t>
inline void to_doubles(const std::vector<field_value_t> &in, std::vector<double> &out) {
    out.resize(3 * in.size());
    for (size_t i = 0; i < in.size(); i++)
        for (int c = 0; c < 3; c++) out[3 * i + c] = double(in[i][c]);
}

template <typename field_value_t>
inline void from_doubles(const std::vector<double> &in, std::vector<field_value_t> &out) {
    out.resize(in.size() / 3);
    for (size_t i = 0; i < out.size(); i++) out[i] = field_value_t(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}

template <typename field_value_t>
inline void put3(double *dst, const field_value_t &v) {
    for (int c = 0; c < 3; c++) dst[c] = double(v[c]);
}

}  // namespace libgran_amd

#endif
