// A driver calling the VTK writer as the reference's drivers do
// (main.cpp:150, test/hamaker_2.cpp:131), compiled against the drop-in
// writer.h and libhipgran.so. It writes particles_1.vtk and particles_2.vtk
// (the state, then the state with every orientation turned by 0.5 rad about
// x) and necks_1.vtk into <dir> for tests/test_vtk_writer.py to compare with
// the oracle. Host code only: no GPU is used.
//
//   writer_driver <dir> <n>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <libgran/vec3.h>
#include <writer.h>

using V3 = libgran_amd::vec3;

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <dir> <n>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const size_t n = std::strtoull(argv[2], nullptr, 10);
    const double r_part = 1.4e-8;
    // a 5 x 5 x k lattice of touching spheres with varied orientations
    std::vector<V3> x(n), theta(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = V3(double(i % 5) * 2.0 * r_part, double((i / 5) % 5) * 2.0 * r_part, double(i / 25) * 2.0 * r_part);
        theta[i] = V3(0.1 * double(i), -0.2 * double(i), 0.05 * double(i) * double(i));
    }
    // necked: neighbours along x
    std::vector<bool> enabled(n * n, true);
    for (size_t i = 0; i + 1 < n; i++)
        if (i % 5 != 4) enabled[i * n + i + 1] = false;
    write_particles(dir, x, theta, r_part);
    for (auto &t : theta) t = V3(t[0] + 0.5, t[1], t[2]);
    write_particles(dir, x, theta, r_part);
    write_necks(dir, x, r_part, enabled);
    return 0;
}
