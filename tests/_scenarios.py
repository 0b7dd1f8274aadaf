"""Initial conditions and parameter sets of the reference's scenario drivers
(test/*.cpp) and of the aggregate workloads, shared by the oracle and the GPU
tests so both sides see identical inputs."""
import math

import numpy as np

import _oracle as O

R = O.R_PART


def lattice_with_impactor():
    """test/contact.cpp:54-77 (also test/hamaker.cpp:59-83): 18 particles in two
    3x3 layers and one particle above moving at (1, 0, -10) m/s."""
    x = []
    for i in range(3):
        y = -3.0 * R + i * 2.0 * R
        for j in range(3):
            xx = -3.0 * R + j * 2.0 * R
            x.append([xx, y, 0.0])
            x.append([xx, y, -2.0 * R])
    x.append([0.0, 0.0, 3.0 * R])
    x = np.array(x)
    v = np.zeros_like(x)
    v[-1] = [1.0, 0.0, -10.0]
    return x, v


def sinter_lattice():
    """test/alt_sintered.cpp:69-90: 3x3x3 touching lattice; the x=x_min column
    moves at +5 m/s in z and the x=x_max column at +5 m/s in y."""
    x, v = [], []
    for i in range(3):
        y = -1.0 * R + i * 2.0 * R
        for j in range(3):
            xx = -1.0 * R + j * 2.0 * R
            x += [[xx, y, 0.0], [xx, y, -2.0 * R], [xx, y, -4.0 * R]]
            vel = [0.0, 0.0, 5.0] if j == 0 else ([0.0, 5.0, 0.0] if j == 2 else [0.0, 0.0, 0.0])
            v += [vel] * 3
    return np.array(x), np.array(v)


def surface_drop():
    """test/surface.cpp:227-242: one particle at (0.5r, 0.5r, 2.5r) moving at
    -0.5 m/s in z over a facet of half-width a = 50 r."""
    a = 50.0 * R
    verts = np.array([[[-a, -a, 0.0], [-a, a, 0.0], [a, -a, 0.0]]])
    x = np.array([[0.5 * R, 0.5 * R, 2.5 * R]])
    v = np.array([[0.0, 0.0, -0.5]])
    return x, v, verts


def surface_contacts():
    """Two particles in contact with the facet of surface_drop() from the
    start of a short run: one over the facet's hypotenuse edge (the
    reference geometry's x, y, lowered to z = 0.71 r so it meets the edge
    within a few dozen steps of 1e-13 s) and one overlapping the face, sliding
    and spinning."""
    a = 50.0 * R
    verts = np.array([[[-a, -a, 0.0], [-a, a, 0.0], [a, -a, 0.0]]])
    x = np.array([[0.5 * R, 0.5 * R, 0.71 * R], [-10.0 * R, -10.0 * R, 0.995 * R]])
    v = np.array([[0.0, 0.0, -0.5], [0.3, 0.0, -0.5]])
    w = np.array([[0.0, 0.0, 0.0], [0.0, 1.0e8, 0.0]])
    return x, v, w, verts


def aggregate_params(k=1.0e4, zeta=0.5, dt=1e-13, mu=1.0, A=1.0e-20, h0=1.0e-10, cutoff=3.0 * R, d_crit=1e-3 * R):
    """Aggregate workload functor set: alt_sinter (bonds with k, damping ratio
    zeta) embedding a contact functor with friction mu, plus a limited-range
    Hamaker functor. gamma = 2 zeta sqrt(k m/2) (reduced mass m/2)."""
    g = 2.0 * zeta * math.sqrt(k * O.MASS / 2.0)
    contact = np.array([k, g, k, g, mu, 1.0, k, g, mu, 1.0, k, g, mu, 1.0, R, O.MASS, O.INERTIA, dt])
    sinter = np.array([k, g, k, g, k, g, k, g, R, O.MASS, O.INERTIA, dt, d_crit])
    ham = np.array([A, h0, cutoff, R, O.MASS])
    return contact, sinter, ham
