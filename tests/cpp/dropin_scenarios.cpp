// Driver programs written against the libgran template interface, compiled
// against the drop-in headers (libgran_amd/include) and libhipgran.so. Each
// scenario sets up one of the reference's scenario geometries, runs the
// requested number of steps and writes the final x, v, theta, omega (binary
// doubles, preceded by n as uint64) for tests/test_gpu_dropin.py to compare
// with the CPU oracle.
//
//   dropin_scenarios <scenario> <steps> <output-file>
//   scenarios: contact hamaker hamaker_neighbor simplified alt_sintered sintered breaking surface plate
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <libgran/contact_force/contact_force.h>
#include <libgran/contact_force/surface_contact_force.h>
#include <libgran/granular_system/granular_system.h>
#include <libgran/granular_system/granular_system_neighbor_list.h>
#include <libgran/hamaker_force/hamaker_force.h>
#include <libgran/hamaker_force/simplified_hamaker_force.h>
#include <libgran/hamaker_force/surface_hamaker_force.h>
#include <libgran/sinter_bridge/alt_breaking_sinter_bridge.h>
#include <libgran/sinter_bridge/alt_sinter_bridge.h>
#include <libgran/sinter_bridge/sinter_bridge.h>
#include <libgran/surface_force/triangular_facet.h>
#include <libgran/vec3.h>
#include <writer.h>

using V3 = libgran_amd::vec3;
using VV = std::vector<V3>;

namespace {

const double r_part = 1.4e-8;
const double rho = 1700.0;
const double mass = 4.0 / 3.0 * M_PI * std::pow(r_part, 3.0) * rho;
const double inertia = 2.0 / 5.0 * mass * std::pow(r_part, 2.0);
const double k = 10000.0, gamma_n = 5.0e-9, mu = 1.0, phi = 1.0, mu_o = 0.1;
const double gamma_t = 0.2 * gamma_n, gamma_r = 0.05 * gamma_n, gamma_o = 0.05 * gamma_n;

void write_state(const char *path, const VV &x, const VV &v, const VV &th, const VV &w) {
    FILE *f = std::fopen(path, "wb");
    if (!f) std::exit(3);
    uint64_t n = x.size();
    std::fwrite(&n, sizeof(n), 1, f);
    for (const VV *b : {&x, &v, &th, &w})
        for (const V3 &q : *b) std::fwrite(q.c, sizeof(double), 3, f);
    std::fclose(f);
}

// Two 3x3 layers and one particle above them moving at (1, 0, -10) m/s.
void lattice_with_impactor(VV &x0, VV &v0) {
    for (int i = 0; i < 3; i++) {
        double y = -3.0 * r_part + double(i) * 2.0 * r_part;
        for (int j = 0; j < 3; j++) {
            double xx = -3.0 * r_part + double(j) * 2.0 * r_part;
            x0.emplace_back(xx, y, 0.0);
            x0.emplace_back(xx, y, -2.0 * r_part);
        }
    }
    x0.emplace_back(0.0, 0.0, 3.0 * r_part);
    v0.assign(x0.size(), V3());
    v0.back() = V3(1.0, 0.0, -10.0);
}

template <typename S>
void finish(S &system, const char *out) {
    write_state(out, system.get_x(), system.get_v(), system.get_theta(), system.get_omega());
}

int run_contact(size_t steps, const char *out, bool with_hamaker, bool neighbor) {
    VV x0, v0;
    lattice_with_impactor(x0, v0);
    VV th0(x0.size()), w0(x0.size());
    const double dt = 1e-13;
    rotational_step_handler<VV, V3> handler;
    contact_force_functor<V3, double> contact(x0.size(), k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o, phi, k,
                                              gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0);
    hamaker_functor<V3, double> hamaker(1.0e-20, 1.0e-9, r_part, mass, V3(), 0.0);
    unary_force_functor_container<V3, double> unary;
    if (!with_hamaker) {
        binary_force_functor_container<V3, double, contact_force_functor<V3, double>> binary{contact};
        granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                        decltype(unary)>
            system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
        for (size_t n = 0; n < steps; n++) system.do_step(dt);
        finish(system, out);
        return 0;
    }
    binary_force_functor_container<V3, double, contact_force_functor<V3, double>, hamaker_functor<V3, double>> binary{
        contact, hamaker};
    if (!neighbor) {
        granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                        decltype(unary)>
            system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
        for (size_t n = 0; n < steps; n++) system.do_step(dt);
        finish(system, out);
        return 0;
    }
    granular_system_neighbor_list<V3, double, rotational_velocity_verlet_half, rotational_step_handler,
                                  decltype(binary), decltype(unary)>
        system(x0.size(), 5.0 * r_part, x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
    for (size_t n = 0; n < steps; n++) {
        if (n % 20 == 0) system.update_neighbor_list();
        system.do_step(dt);
    }
    finish(system, out);
    return 0;
}

// test/hamaker.cpp's set-up with simplified_hamaker_functor in place of
// hamaker_functor (all pairs).
int run_simplified(size_t steps, const char *out) {
    VV x0, v0;
    lattice_with_impactor(x0, v0);
    VV th0(x0.size()), w0(x0.size());
    const double dt = 1e-13;
    rotational_step_handler<VV, V3> handler;
    contact_force_functor<V3, double> contact(x0.size(), k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o, phi, k,
                                              gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0);
    simplified_hamaker_functor<V3, double> hamaker(1.0e-20, 1.0e-9, r_part, mass, V3(), 0.0);
    unary_force_functor_container<V3, double> unary;
    binary_force_functor_container<V3, double, contact_force_functor<V3, double>,
                                   simplified_hamaker_functor<V3, double>>
        binary{contact, hamaker};
    granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                    decltype(unary)>
        system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
    for (size_t n = 0; n < steps; n++) system.do_step(dt);
    finish(system, out);
    return 0;
}

int run_alt_sintered(size_t steps, const char *out) {
    VV x0, v0;
    for (int i = 0; i < 3; i++) {
        double y = -1.0 * r_part + double(i) * 2.0 * r_part;
        for (int j = 0; j < 3; j++) {
            double xx = -1.0 * r_part + double(j) * 2.0 * r_part;
            x0.emplace_back(xx, y, 0.0);
            x0.emplace_back(xx, y, -2.0 * r_part);
            x0.emplace_back(xx, y, -4.0 * r_part);
            V3 vel = j == 0 ? V3(0.0, 0.0, 5.0) : (j == 2 ? V3(0.0, 5.0, 0.0) : V3());
            for (int q = 0; q < 3; q++) v0.push_back(vel);
        }
    }
    VV th0(x0.size()), w0(x0.size());
    const double dt = 5e-14;
    const double k_bond = 1000.0, g_bond = 2.0 * std::sqrt(2.0 * mass * k_bond);
    rotational_step_handler<VV, V3> handler;
    contact_force_functor<V3, double> contact(x0.size(), k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o, phi, k,
                                              gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0);
    alt_sinter_functor<V3, double> sinter(x0.size(), x0, k_bond, g_bond, k_bond, 0.2 * g_bond, k_bond, 0.05 * g_bond,
                                          k_bond, 0.05 * g_bond, r_part, mass, inertia, dt, V3(), 0.0, 1.0e-9,
                                          contact);
    // the public bond table reads like the reference's dense one
    size_t nbonded = 0;
    for (size_t i = 0; i < x0.size(); i++)
        for (size_t j = 0; j < x0.size(); j++) nbonded += sinter.bonded_contacts[i * x0.size() + j] ? 1 : 0;
    if (nbonded != 2 * 26) {
        std::fprintf(stderr, "expected 52 ordered bonded pairs, found %zu\n", nbonded);
        return 4;
    }
    binary_force_functor_container<V3, double, alt_sinter_functor<V3, double>> binary{sinter};
    unary_force_functor_container<V3, double> unary;
    granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                    decltype(unary)>
        system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
    for (size_t n = 0; n < steps; n++) system.do_step(dt);
    auto springs = sinter.get_contact_springs();
    double s = 0.0;
    for (auto &t : springs) s += std::get<0>(t).norm() + std::get<1>(t).norm() + std::get<2>(t).norm();
    if (!(s > 0.0)) {
        std::fprintf(stderr, "bonded springs read back as zero\n");
        return 5;
    }
    // the final state and the necks as VTK, as the reference drivers write
    // them (main.cpp:150), beside the output file
    const std::string path(out);
    const std::string dir = path.rfind('/') == std::string::npos ? "." : path.substr(0, path.rfind('/'));
    write_particles(dir, system.get_x(), system.get_theta(), r_part);
    std::vector<bool> enabled(x0.size() * x0.size());
    for (size_t q = 0; q < enabled.size(); q++) enabled[q] = !sinter.bonded_contacts[q];
    write_necks(dir, system.get_x(), r_part, enabled);
    finish(system, out);
    return 0;
}

// test/sintered.cpp: the legacy sinter_functor with its step handler on the
// 3x3x3 lattice, unit vectors from the reference's own generator
// (sintered.cpp:30-39). The vectors handed out are written to <out>.vecs
// (count as uint64, then 3 doubles each) for the oracle to draw the same.
template <typename field_container_t, typename field_value_t>
using sinter_step_handler_double = sinter_step_handler<field_container_t, field_value_t, double>;

int run_sintered(size_t steps, const char *out) {
    VV x0, v0;
    for (size_t i = 0; i < 3; i++) {
        auto y = -1.0 * r_part + double(i) * 2.0 * r_part;
        for (size_t j = 0; j < 3; j++) {
            auto x = -1.0 * r_part + double(j) * 2.0 * r_part;
            x0.emplace_back(x, y, 0.0);
            x0.emplace_back(x, y, -2.0 * r_part);
            x0.emplace_back(x, y, -4.0 * r_part);
            V3 vel = j == 0 ? V3(0.0, 0.0, 5.0) : (j == 2 ? V3(0.0, 5.0, 0.0) : V3());
            for (size_t q = 0; q < 3; q++) v0.push_back(vel);
        }
    }
    VV th0(x0.size()), w0(x0.size());
    const double dt = 1e-13, gamma_d = 0.0;
    std::mt19937_64 mt(0);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    VV drawn;
    auto generate_random_unit_vector = [&]() {
        V3 vec;
        do {
            vec = V3{dist(mt), dist(mt), dist(mt)};
        } while (vec.norm() == 0);
        drawn.push_back(vec.normalized());
        return drawn.back();
    };
    sinter_functor<V3, double> sinter_model(x0.size(), k, gamma_d, k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o,
                                            phi, k, gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0,
                                            x0.cbegin(), generate_random_unit_vector, 1.0e-9);
    binary_force_functor_container<V3, double, sinter_functor<V3, double>> binary{sinter_model};
    unary_force_functor_container<V3, double> unary;
    auto step_handler_instance = sinter_model.get_step_handler<VV>();
    granular_system<V3, double, rotational_velocity_verlet_half, sinter_step_handler_double, decltype(binary),
                    decltype(unary)>
        system(x0, v0, th0, w0, 0.0, V3(), 0.0, step_handler_instance, binary, unary);
    for (size_t n = 0; n < steps; n++) system.do_step(dt);
    FILE *f = std::fopen((std::string(out) + ".vecs").c_str(), "wb");
    if (!f) return 3;
    const uint64_t nv = drawn.size();
    std::fwrite(&nv, sizeof(nv), 1, f);
    for (const V3 &q : drawn) std::fwrite(q.c, sizeof(double), 3, f);
    std::fclose(f);
    finish(system, out);
    return 0;
}

// Neck strengths 1e-8 N * (1 + b % 7) for the b-th draw.
struct cycling_strength {
    size_t b = 0;
    double operator()() { return 1.0e-8 * double(1 + (b++ % 7)); }
};

int run_breaking(size_t steps, const char *out) {
    VV x0, v0;
    for (int i = 0; i < 3; i++) {
        double y = -1.0 * r_part + double(i) * 2.0 * r_part;
        for (int j = 0; j < 3; j++) {
            double xx = -1.0 * r_part + double(j) * 2.0 * r_part;
            x0.emplace_back(xx, y, 0.0);
            x0.emplace_back(xx, y, -2.0 * r_part);
            x0.emplace_back(xx, y, -4.0 * r_part);
            V3 vel = j == 0 ? V3(0.0, 0.0, 5.0) : (j == 2 ? V3(0.0, 5.0, 0.0) : V3());
            for (int q = 0; q < 3; q++) v0.push_back(vel);
        }
    }
    VV th0(x0.size()), w0(x0.size());
    const double dt = 5e-14;
    const double k_bond = 1000.0, g_bond = 2.0 * std::sqrt(2.0 * mass * k_bond);
    rotational_step_handler<VV, V3> handler;
    contact_force_functor<V3, double> contact(x0.size(), k, gamma_n, k, gamma_t, mu, phi, k, gamma_r, mu_o, phi, k,
                                              gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0);
    alt_breaking_sinter_functor<V3, double, cycling_strength> sinter(
        x0.size(), x0, k_bond, g_bond, k_bond, 0.2 * g_bond, k_bond, 0.05 * g_bond, k_bond, 0.05 * g_bond, r_part,
        mass, inertia, dt, V3(), 0.0, 1.0e-9, cycling_strength{}, contact);
    binary_force_functor_container<V3, double, decltype(sinter)> binary{sinter};
    unary_force_functor_container<V3, double> unary;
    granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                    decltype(unary)>
        system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
    size_t bonded_before = 0;
    for (size_t idx = 0; idx < x0.size() * x0.size(); idx++) bonded_before += sinter.bonded_contacts[idx] ? 1 : 0;
    for (size_t n = 0; n < steps; n++) system.do_step(dt);
    size_t bonded_after = 0;
    for (size_t idx = 0; idx < x0.size() * x0.size(); idx++) bonded_after += sinter.bonded_contacts[idx] ? 1 : 0;
    std::fprintf(stderr, "bonded ordered pairs: %zu -> %zu\n", bonded_before, bonded_after);
    if (bonded_before != 2 * 54 || !(bonded_after < bonded_before)) return 6;
    finish(system, out);
    return 0;
}

int run_surface(size_t steps, const char *out, bool moving) {
    using sc_t = surface_contact_force_functor<V3, double>;
    using sh_t = surface_hamaker_functor<V3, double>;
    using facet_t = triangular_facet<V3, double, sc_t, sh_t>;
    const double dt = 1e-13, a = 50.0 * r_part, mu_s = 0.5;
    // test/surface.cpp's facet; one particle meets its hypotenuse edge, one
    // overlaps its face while sliding and spinning
    VV x0{V3(0.5 * r_part, 0.5 * r_part, 0.71 * r_part), V3(-10.0 * r_part, -10.0 * r_part, 0.995 * r_part)};
    VV v0{V3(0.0, 0.0, -0.5), V3(0.3, 0.0, -0.5)}, th0(2), w0{V3(), V3(0.0, 1.0e8, 0.0)};
    rotational_step_handler<VV, V3> handler;
    contact_force_functor<V3, double> contact(x0.size(), k, gamma_n, k, gamma_t, mu_s, phi, k, gamma_r, mu_o, phi, k,
                                              gamma_o, mu_o, phi, r_part, mass, inertia, dt, V3(), 0.0);
    sc_t sc(x0.size(), k, gamma_n, k, gamma_t, mu_s, phi, k, gamma_r, mu_o, phi, k, gamma_o, mu_o, phi, r_part, mass,
            inertia, dt, V3(), 0.0);
    sh_t sh(moving ? 1.0e-20 : 0.0, 1.0e-9, r_part, mass, V3(), 0.0);
    facet_t facet(moving ? V3(0.5, -0.2, 1.0) : V3(), std::make_tuple(V3(-a, -a, 0), V3(-a, a, 0), V3(a, -a, 0)), sc,
                  sh);
    binary_force_functor_container<V3, double, contact_force_functor<V3, double>> binary{contact};
    unary_force_functor_container<V3, double, facet_t> unary{facet};
    granular_system<V3, double, rotational_velocity_verlet_half, rotational_step_handler, decltype(binary),
                    decltype(unary)>
        system(x0, v0, th0, w0, 0.0, V3(), 0.0, handler, binary, unary);
    for (size_t n = 0; n < steps; n++) {
        system.do_step(dt);
        facet.update_positions(dt);
    }
    finish(system, out);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <scenario> <steps> <output>\n", argv[0]);
        return 2;
    }
    const std::string s = argv[1];
    const size_t steps = std::strtoull(argv[2], nullptr, 10);
    try {
        if (s == "contact") return run_contact(steps, argv[3], false, false);
        if (s == "hamaker") return run_contact(steps, argv[3], true, false);
        if (s == "hamaker_neighbor") return run_contact(steps, argv[3], true, true);
        if (s == "simplified") return run_simplified(steps, argv[3]);
        if (s == "alt_sintered") return run_alt_sintered(steps, argv[3]);
        if (s == "sintered") return run_sintered(steps, argv[3]);
        if (s == "breaking") return run_breaking(steps, argv[3]);
        if (s == "surface") return run_surface(steps, argv[3], false);
        if (s == "plate") return run_surface(steps, argv[3], true);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "unknown scenario %s\n", s.c_str());
    return 2;
}
