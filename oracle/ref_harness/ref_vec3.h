// Field type used to instantiate the reference functor templates in the
// reference harness. The functors are templates over `field_value_t` and only
// call the methods below (contact_force.h:65-95, hamaker_force.h:32-42,
// triangular_facet.h:23-137); the arithmetic order of each method matches the
// oracle's vec3 (oracle/oracle.hpp) so that any difference between the two
// comes from the functor code, not from the vector type.
#pragma once
#include <cmath>

struct rvec {
    double x_, y_, z_;
    rvec() : x_(0.0), y_(0.0), z_(0.0) {}
    rvec(double a, double b, double c) : x_(a), y_(b), z_(c) {}
    double dot(const rvec &o) const { return (x_ * o.x_ + y_ * o.y_) + z_ * o.z_; }
    double norm() const { return std::sqrt(dot(*this)); }
    rvec normalized() const {
        double z = dot(*this);
        if (z > 0.0) return *this / std::sqrt(z);
        return *this;
    }
    void normalize() {
        double z = dot(*this);
        if (z > 0.0) *this /= std::sqrt(z);
    }
    rvec cross(const rvec &o) const {
        return {y_ * o.z_ - z_ * o.y_, z_ * o.x_ - x_ * o.z_, x_ * o.y_ - y_ * o.x_};
    }
    rvec operator+(const rvec &o) const { return {x_ + o.x_, y_ + o.y_, z_ + o.z_}; }
    rvec operator-(const rvec &o) const { return {x_ - o.x_, y_ - o.y_, z_ - o.z_}; }
    rvec operator-() const { return {-x_, -y_, -z_}; }
    rvec operator*(double s) const { return {x_ * s, y_ * s, z_ * s}; }
    rvec operator/(double s) const { return {x_ / s, y_ / s, z_ / s}; }
    rvec &operator+=(const rvec &o) { x_ += o.x_; y_ += o.y_; z_ += o.z_; return *this; }
    rvec &operator-=(const rvec &o) { x_ -= o.x_; y_ -= o.y_; z_ -= o.z_; return *this; }
    rvec &operator*=(double s) { x_ *= s; y_ *= s; z_ *= s; return *this; }
    rvec &operator/=(double s) { x_ /= s; y_ /= s; z_ /= s; return *this; }
};
inline rvec operator*(double s, const rvec &a) { return {s * a.x_, s * a.y_, s * a.z_}; }

inline rvec rv_ld(const double *p) { return {p[0], p[1], p[2]}; }
inline void rv_st(double *p, const rvec &a) { p[0] = a.x_; p[1] = a.y_; p[2] = a.z_; }
