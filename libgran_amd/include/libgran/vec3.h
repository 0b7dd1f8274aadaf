// A minimal 3-vector for drivers without Eigen: operator[], a 3-argument
// constructor and the arithmetic the drivers use. The drop-in headers only
// need operator[] and the constructor.
#ifndef LIBGRAN_VEC3_H
#define LIBGRAN_VEC3_H

#include <cmath>

namespace libgran_amd {
struct vec3 {
    double c[3];
    vec3() : c{0.0, 0.0, 0.0} {}
    vec3(double x, double y, double z) : c{x, y, z} {}
    static vec3 Zero() { return vec3(); }
    double &operator[](int k) { return c[k]; }
    double operator[](int k) const { return c[k]; }
    vec3 operator+(const vec3 &o) const { return {c[0] + o.c[0], c[1] + o.c[1], c[2] + o.c[2]}; }
    vec3 operator-(const vec3 &o) const { return {c[0] - o.c[0], c[1] - o.c[1], c[2] - o.c[2]}; }
    vec3 operator*(double s) const { return {c[0] * s, c[1] * s, c[2] * s}; }
    vec3 operator/(double s) const { return {c[0] / s, c[1] / s, c[2] / s}; }
    vec3 &operator+=(const vec3 &o) {
        for (int k = 0; k < 3; k++) c[k] += o.c[k];
        return *this;
    }
    double dot(const vec3 &o) const { return (c[0] * o.c[0] + c[1] * o.c[1]) + c[2] * o.c[2]; }
    double norm() const { return std::sqrt(dot(*this)); }
    double squaredNorm() const { return dot(*this); }
    // Eigen's semantics: a zero vector is returned unchanged
    vec3 normalized() const {
        const double z = dot(*this);
        return z > 0.0 ? *this / std::sqrt(z) : *this;
    }
    vec3 cross(const vec3 &o) const {
        return {c[1] * o.c[2] - c[2] * o.c[1], c[2] * o.c[0] - c[0] * o.c[2], c[0] * o.c[1] - c[1] * o.c[0]};
    }
};
inline vec3 operator*(double s, const vec3 &a) { return a * s; }
}  // namespace libgran_amd

#endif
