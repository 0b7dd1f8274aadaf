// Standard headers the reference functor headers expect their includer to
// provide (contact_force.h and hamaker_force.h include nothing themselves).
// `using std::abs` makes the unqualified abs() in alt_sinter_bridge.h:69 take
// a double, which is what the bond predicate means (a gap tolerance).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <forward_list>
#include <iostream>
#include <numeric>
#include <tuple>
#include <utility>
#include <vector>
using std::abs;
#include "ref_vec3.h"

// One pair state: xi, xj, vi, vj, wi, wj (18 doubles).
struct pair_state {
    std::vector<rvec> x, v, th, w;
    pair_state() : x(2), v(2), th(2), w(2) {}
    void load(const double *s) {
        x[0] = rv_ld(s); x[1] = rv_ld(s + 3);
        v[0] = rv_ld(s + 6); v[1] = rv_ld(s + 9);
        w[0] = rv_ld(s + 12); w[1] = rv_ld(s + 15);
    }
};
