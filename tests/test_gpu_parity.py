"""HIP path against the CPU oracle on identical inputs, through the C ABI.

The FP64 path evaluates every pair in the reference's order with the
reference's expressions and no FMA contraction, so the expected result is
bit-for-bit equality of positions, velocities, orientations, angular
velocities, accelerations, neighbour lists, bonds and springs.
"""
import math
import sys

import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART


def _gpu():
    import libgran_amd
    return libgran_amd


def both(binary, hamaker, x0, v0=None, w0=None, pair_mode=0, r_verlet=0.0, contact=None, sinter=None, ham=None,
         facets=None, facet_contact=0, facet_hamaker=0, surface_contact=None, surface_hamaker=None):
    L = _gpu()
    op = O.system_params(binary, hamaker, neighbor_mode=pair_mode, r_verlet=r_verlet, contact=contact,
                         sinter=sinter, hamaker=ham, facet_contact=facet_contact, facet_hamaker=facet_hamaker,
                         surface_contact=surface_contact, surface_hamaker=surface_hamaker)
    osys = O.OracleSystem(op, x0, v0, w0=w0)
    gp = L.make_params(binary, hamaker, pair_mode=pair_mode, r_verlet=r_verlet,
                       contact=contact if contact is not None else O.contact_params(),
                       sinter=sinter if sinter is not None else O.sinter_params(),
                       hamaker=ham if ham is not None else O.hamaker_params(),
                       facet_contact=facet_contact, facet_hamaker=facet_hamaker,
                       surface_contact=surface_contact if surface_contact is not None else O.contact_params(),
                       surface_hamaker=surface_hamaker if surface_hamaker is not None else O.hamaker_params())
    gsys = L.System(gp, x0, v0, omega0=w0)
    if facets is not None:
        verts, vel = facets
        osys.set_facets(verts, vel)
        gsys.set_facets(verts, vel)
    return osys, gsys


def assert_same_state(osys, gsys, what=""):
    a, b = osys.state(), gsys.state()
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(b[k], a[k], err_msg=f"{what}: field {k}")


def announce(name):
    print(f"[gpu-parity] {name}", flush=True)
    sys.stdout.flush()


def test_single_acceleration_pass_contact_lattice():
    announce("single pass, contact lattice")
    x, v = S.lattice_with_impactor()
    x[-1, 2] = 1.98 * R  # start the impactor overlapping the top layer
    o, g = both(O.BIN_CONTACT, O.HAM_FULL, x, v, ham=O.hamaker_params())
    o.compute_accelerations()
    g.compute_accelerations()
    assert_same_state(o, g, "one pass")


def test_contact_lattice_impact():
    """test/contact.cpp geometry, contact functor only, all pairs."""
    announce("contact lattice impact 32000 steps")
    x, v = S.lattice_with_impactor()
    o, g = both(O.BIN_CONTACT, O.HAM_NONE, x, v)
    for chunk in range(4):
        o.step(1e-13, 8000)
        g.step(1e-13, 8000)
        assert_same_state(o, g, f"chunk {chunk}")
    # the impactor reached the lattice: contact springs were active at some point
    assert np.abs(g.state()["v"][:-1]).max() > 0.0


def test_hamaker_lattice_all_pairs():
    """test/hamaker.cpp geometry: contact + all-pairs Hamaker (A=1e-20, h0=1e-9)."""
    announce("hamaker lattice all pairs")
    x, v = S.lattice_with_impactor()
    o, g = both(O.BIN_CONTACT, O.HAM_FULL, x, v, ham=O.hamaker_params(A=1e-20, h0=1e-9))
    o.step(1e-13, 6000)
    g.step(1e-13, 6000)
    assert_same_state(o, g)
    ij1, s1 = o.springs(0)
    ij2, s2 = g.springs(0)
    np.testing.assert_array_equal(ij2, ij1)
    np.testing.assert_array_equal(s2, s1)


def test_hamaker_lattice_neighbor_list():
    """test/hamaker_neighbor.cpp: r_verlet = 5r, list rebuilt every 20 steps."""
    announce("hamaker lattice neighbour list")
    x, v = S.lattice_with_impactor()
    x[-1, 2] = 2.3 * R
    o, g = both(O.BIN_CONTACT, O.HAM_FULL, x, v, pair_mode=1, r_verlet=5.0 * R,
                ham=O.hamaker_params(A=1e-20, h0=1e-9))
    o.step(1e-13, 3000, rebuild_every=20)
    g.step(1e-13, 3000, rebuild_every=20)
    assert_same_state(o, g)
    off1, idx1 = o.neighbors()
    off2, idx2 = g.neighbors()
    np.testing.assert_array_equal(off2, off1)
    np.testing.assert_array_equal(idx2, idx1)


def test_alt_sinter_lattice():
    """test/alt_sintered.cpp: 3x3x3 bonded lattice (spanning tree of 26 bonds)."""
    announce("alt_sinter lattice")
    x, v = S.sinter_lattice()
    k_bond = 1000.0
    sp = O.sinter_params(k_bond=k_bond, d_crit=1.0e-9, dt=5e-14)
    cp = O.contact_params(dt=5e-14)
    o, g = both(O.BIN_SINTER, O.HAM_NONE, x, v, contact=cp, sinter=sp)
    pb, skb = g.bonds()
    po, sko = o.bonds()
    np.testing.assert_array_equal(pb, po)
    assert skb == sko == 28
    o.step(5e-14, 5000)
    g.step(5e-14, 5000)
    assert_same_state(o, g)
    ij1, s1 = o.springs(1)
    ij2, s2 = g.springs(1)
    # the device table holds every bond from the start; the oracle creates an
    # entry at first evaluation. All bonds were evaluated, so the sets agree.
    np.testing.assert_array_equal(ij2, ij1)
    np.testing.assert_array_equal(s2, s1)


def test_surface_facet_drop():
    """test/surface.cpp facet and functors (surface contact mu 0.5, surface
    Hamaker with A = 0): one particle meeting the facet's edge, one sliding
    and spinning on its face."""
    announce("surface facet")
    x, v, w, verts = S.surface_contacts()
    cp = O.contact_params(mu=0.5)
    o, g = both(O.BIN_CONTACT, O.HAM_NONE, x, v, w0=w, contact=cp, facets=(verts, np.zeros((1, 3))),
                facet_contact=1, facet_hamaker=1, surface_contact=cp,
                surface_hamaker=O.hamaker_params(A=0.0, h0=1e-9))
    o.step(1e-13, 3000)
    g.step(1e-13, 3000)
    assert_same_state(o, g)
    st = g.state()
    assert st["v"][0, 2] > -0.5 and st["v"][1, 2] > -0.5  # both met the facet
    assert abs(st["omega"][1, 1] - 1.0e8) > 1.0  # friction acted on the spin


def test_moving_facets_and_tilted_plate():
    """Two moving triangles (a plate pair compressing a small aggregate) with
    surface contact + surface Hamaker."""
    announce("moving facets")
    L = _gpu()
    x = L.generate_aggregate(60, R, 1.8, 1.3, 4)
    zmin, zmax = x[:, 2].min(), x[:, 2].max()
    a = 40 * R
    verts = np.array([[[-a, -a, zmin - 1.02 * R], [a, -a, zmin - 1.02 * R], [-a, a, zmin - 0.9 * R]],
                      [[-a, -a, zmax + 1.02 * R], [-a, a, zmax + 1.02 * R], [a, -a, zmax + 1.02 * R]]])
    vel = np.array([[0.0, 0.0, 3.0], [0.0, 0.0, -3.0]])
    contact, sinter, ham = S.aggregate_params()
    rng = np.random.default_rng(3)
    v0 = rng.normal(0.0, 1.0, x.shape)
    o, g = both(O.BIN_SINTER, O.HAM_LIMITED, x, v0, contact=contact, sinter=sinter, ham=ham, facets=(verts, vel),
                facet_contact=1, facet_hamaker=1, surface_contact=contact,
                surface_hamaker=O.hamaker_params(A=1e-20, h0=1e-10))
    o.step(1e-13, 2000)
    g.step(1e-13, 2000)
    assert_same_state(o, g)
    gv, _ = g.facets(2)
    np.testing.assert_allclose(gv[0, :, 2] - verts[0, :, 2], 3.0 * 2000 * 1e-13, rtol=1e-9)


def test_aggregate_all_pairs_config1():
    """N=200 aggregate (seed 0, Df 1.8, kf 1.3): alt_sinter + limited-range
    Hamaker (cutoff 3r, h0 0.1 nm), all pairs, Gaussian v0 (1 m/s)."""
    announce("aggregate N=200 all pairs")
    L = _gpu()
    x = L.generate_aggregate(200, R, 1.8, 1.3, 0)
    rng = np.random.default_rng(0)
    v0 = rng.normal(0.0, 1.0, x.shape)
    contact, sinter, ham = S.aggregate_params()
    o, g = both(O.BIN_SINTER, O.HAM_LIMITED, x, v0, contact=contact, sinter=sinter, ham=ham)
    o.step(1e-13, 1000)
    g.step(1e-13, 1000)
    assert_same_state(o, g)
    for which in (0, 1):
        ij1, s1 = o.springs(which)
        ij2, s2 = g.springs(which)
        np.testing.assert_array_equal(ij2, ij1)
        np.testing.assert_array_equal(s2, s1)


def test_aggregate_neighbor_list_config2_small():
    """Neighbour-list workload at reduced size: N=3000 (seed 1), r_list 3.2r,
    rebuilt every 50 steps."""
    announce("aggregate N=3000 neighbour list")
    L = _gpu()
    x = L.generate_aggregate(3000, R, 1.8, 1.3, 1)
    rng = np.random.default_rng(1)
    v0 = rng.normal(0.0, 1.0, x.shape)
    contact, sinter, ham = S.aggregate_params()
    o, g = both(O.BIN_SINTER, O.HAM_LIMITED, x, v0, pair_mode=1, r_verlet=3.2 * R, contact=contact, sinter=sinter,
                ham=ham)
    o.step(1e-13, 400, rebuild_every=50)
    g.step(1e-13, 400, rebuild_every=50)
    assert_same_state(o, g)
    off1, idx1 = o.neighbors()
    off2, idx2 = g.neighbors()
    np.testing.assert_array_equal(off2, off1)
    np.testing.assert_array_equal(idx2, idx1)


def test_neighbor_list_full_size_bit_exact():
    """N=20000 aggregate (seed 1): device cell-list rebuild equals the
    oracle's brute-force list (same predicate, ascending j)."""
    announce("neighbour list N=20000")
    L = _gpu()
    x = L.generate_aggregate(20000, R, 1.8, 1.3, 1)
    contact, sinter, ham = S.aggregate_params()
    op = O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=contact,
                         sinter=sinter, hamaker=ham)
    osys = O.OracleSystem(op, x)
    g = L.System(L.make_params(2, 2, pair_mode=1, r_verlet=3.2 * R, contact=contact, sinter=sinter, hamaker=ham), x)
    off1, idx1 = osys.neighbors()
    off2, idx2 = g.neighbors()
    np.testing.assert_array_equal(off2, off1)
    np.testing.assert_array_equal(idx2, idx1)
    pb, _ = g.bonds()
    po, _ = osys.bonds()
    np.testing.assert_array_equal(pb, po)
