// Drop-in for the reference drivers' VTK writer (writer.h, writer.cpp:51-172)
// over hg_write_vtk_particles / hg_write_vtk_necks. It keeps the reference's
// function names and arguments, its file names (dir/particles_<k>.vtk and
// dir/necks_<k>.vtk, k counting from 1 in each program), and its failure
// behaviour: a file that cannot be created prints "ERROR: Unable to create dump
// file" and exits with EXIT_FAILURE. The vector type is any type with
// operator[]: libgran_amd::vec3 in these headers, Eigen::Vector3d in the
// reference. write_spring_connectors (writer.cpp:13-49) draws the legacy
// sinter_functor's spring connectors; that functor has no device path, so
// the function is not provided.
#ifndef LIBGRAN_WRITER_H
#define LIBGRAN_WRITER_H

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hipgran.h"

namespace libgran_amd::vtk_detail {

inline size_t &particles_counter() {
    static size_t k = 1;
    return k;
}

inline size_t &necks_counter() {
    static size_t k = 1;
    return k;
}

template <typename V>
std::vector<double> flatten(std::vector<V> const &v) {
    std::vector<double> out(3 * v.size());
    for (size_t i = 0; i < v.size(); i++)
        for (int c = 0; c < 3; c++) out[3 * i + c] = v[i][c];
    return out;
}

inline void check(hg_status st) {
    if (st == HG_OK) return;
    if (st == HG_ERR_IO) {
        std::cerr << "ERROR: Unable to create dump file" << "\n";
        std::exit(EXIT_FAILURE);
    }
    throw std::runtime_error(std::string("VTK writer: ") + hg_last_error());
}

}  // namespace libgran_amd::vtk_detail

// Particle positions / r_part and the body axes of theta (writer.cpp:51-118).
template <typename V>
void write_particles(std::string const &dir, std::vector<V> const &x, std::vector<V> const &theta, double r_part) {
    namespace d = libgran_amd::vtk_detail;
    const std::string path = dir + "/particles_" + std::to_string(d::particles_counter()++) + ".vtk";
    if (x.size() != theta.size()) throw std::runtime_error("write_particles: x and theta differ in length");
    const std::vector<double> xf = d::flatten(x), tf = d::flatten(theta);
    d::check(hg_write_vtk_particles(path.c_str(), x.size(), xf.data(), tf.data(), r_part));
}

// Necks: the pairs i < j whose enabled_contacts[i * n + j] is false, row by
// row (writer.cpp:120-172).
template <typename V>
void write_necks(std::string const &dir, std::vector<V> const &x, double r_part,
                 std::vector<bool> const &enabled_contacts) {
    namespace d = libgran_amd::vtk_detail;
    const std::string path = dir + "/necks_" + std::to_string(d::necks_counter()++) + ".vtk";
    const size_t n = x.size();
    if (enabled_contacts.size() < n * n) throw std::runtime_error("write_necks: enabled_contacts is not n*n");
    std::vector<uint32_t> pairs;
    for (size_t i = 0; i + 1 < n; i++)
        for (size_t j = i + 1; j < n; j++)
            if (!enabled_contacts[i * n + j]) {
                pairs.push_back(uint32_t(i));
                pairs.push_back(uint32_t(j));
            }
    const std::vector<double> xf = d::flatten(x);
    d::check(hg_write_vtk_necks(path.c_str(), n, xf.data(), pairs.size() / 2, pairs.data(), r_part));
}

#endif  // LIBGRAN_WRITER_H
