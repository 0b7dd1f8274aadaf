"""The legacy sinter_functor with its sinter_step_handler
(HG_BINARY_SINTER_LEGACY; sinter_bridge.h:13-272) on the GPU against the
oracle, through the C ABI: positions, velocities, orientations, spins,
accelerations and the moving connectors bit-exact, all pairs and neighbour
list; and the reference's own known answer for it, test/sintered.cpp at full
length (10^6 steps): radius of gyration 3.95981e-08 m and centre-of-mass
drift 2.35702e-07 m within 1 %."""
import math

import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART


def _L():
    import libgran_amd
    return libgran_amd


def announce(name):
    print(f"[gpu-legacy] {name}", flush=True)


def unit_vectors(seed, count=20000):
    v = np.random.default_rng(seed).uniform(-1.0, 1.0, (count, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def sintered_params(dt=1e-13, gamma_d=0.0):
    """test/sintered.cpp:52-73: k_d = k = 1e4, undamped bridges, the contact
    functor of the other scenarios, critical separation 1 nm."""
    s13 = O.sinter_params(k_bond=1.0e4, d_crit=1.0e-9, dt=dt)
    s13[1] = gamma_d
    return O.contact_params(dt=dt), s13


def make_pair(x, v, contact, s13, ham_model=O.HAM_NONE, ham=None, pair_mode=0, r_verlet=0.0, seed=3):
    L = _L()
    lib_ham = {O.HAM_NONE: L.HAMAKER_NONE, O.HAM_FULL: L.HAMAKER_FULL, O.HAM_LIMITED: L.HAMAKER_LIMITED}
    g = L.System(L.make_params(L.BINARY_SINTER_LEGACY, lib_ham[ham_model], pair_mode=pair_mode, r_verlet=r_verlet,
                               contact=contact, sinter=s13, hamaker=ham if ham is not None else O.hamaker_params()),
                 x, v)
    vecs = unit_vectors(seed)
    pairs, conn, _ = O.legacy_frames(x, contact[14], s13[12], vecs)
    g.set_legacy_connectors(pairs, conn)
    o = O.OracleSystem(O.system_params(O.BIN_SINTER_LEGACY, ham_model, neighbor_mode=pair_mode, r_verlet=r_verlet,
                                       contact=contact, sinter=s13, hamaker=ham), x, v)
    o.set_legacy(pairs, conn)
    assert np.array_equal(g.bonds()[0], pairs)
    return g, o


def same(g, o, what):
    a, b = g.state(), o.state()
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    np.testing.assert_array_equal(g.legacy_connectors(), o.legacy_connectors(), err_msg=f"{what}: connectors")
    # the energy ledger (connector springs in the normal term): the oracle's
    # to 1e-12 of its scale, as in test_gpu_energy.py
    L = _L()
    dev = np.array([g.energy(O.MASS, O.INERTIA)[t] for t in L.ENERGY_TERMS])
    orc = o.energy()
    err = np.abs(dev - orc).max() / np.abs(orc[:7]).sum()
    print(f"[gpu-legacy] {what}: ledger max |dev - oracle| / scale {err:.2e}", flush=True)
    assert err < 1e-12 and dev[2] > 0.0


def test_legacy_lattice_all_pairs():
    """test/sintered.cpp set-up (the 3x3x3 lattice, 54 bonds, two columns
    launched at 5 m/s), all pairs, 20,000 steps."""
    announce("lattice all pairs 20000 steps")
    x, v = S.sinter_lattice()
    c, s13 = sintered_params()
    g, o = make_pair(x, v, c, s13)
    assert len(g.bonds()[0]) == 54
    g.step(1e-13, 20000)
    o.step(1e-13, 20000)
    same(g, o, "lattice")


@pytest.mark.parametrize("ham_model", [O.HAM_NONE, O.HAM_LIMITED])
def test_legacy_aggregate_neighbor_list(ham_model):
    """A 400-particle aggregate bonded by the legacy rule (every touching
    pair, cycles kept), damped bridges, contacts between unbonded pairs,
    optionally limited-range Hamaker, over a 3.2r neighbour list rebuilt every
    50 steps: 400 steps."""
    announce(f"aggregate N=400 neighbour list, hamaker {ham_model}")
    L = _L()
    x = L.generate_aggregate(400, R, 1.8, 1.3, 23)
    v = np.random.default_rng(23).normal(0.0, 2.0, x.shape)
    c, s13 = sintered_params(gamma_d=2e-9)
    s13[12] = 1e-3 * R
    g, o = make_pair(x, v, c, s13, ham_model=ham_model, ham=O.hamaker_params(A=1e-20, h0=1e-10),
                     pair_mode=1, r_verlet=3.2 * R, seed=4)
    g.step(1e-13, 400, rebuild_every=50)
    o.step(1e-13, 400, rebuild_every=50)
    same(g, o, "aggregate")
    assert np.abs(g.state()["omega"]).max() > 0.0


def test_legacy_errors():
    """Stepping or reading the ledger before the connectors are set, mixed
    precision, a pool and the wrong bond list are refused."""
    announce("error paths")
    L = _L()
    x, v = S.sinter_lattice()
    c, s13 = sintered_params()
    p = L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_NONE, contact=c, sinter=s13)
    g = L.System(p, x, v)
    with pytest.raises(L.HipgranError, match="connectors"):
        g.step(1e-13, 1)
    with pytest.raises(L.HipgranError, match="connectors"):
        g.energy(O.MASS, O.INERTIA)
    pairs, conn, _ = O.legacy_frames(x, R, 1e-9, unit_vectors(1))
    with pytest.raises(L.HipgranError):
        g.set_legacy_connectors(pairs[::-1], conn[::-1])
    g.set_legacy_connectors(pairs, conn)
    g.step(1e-13, 1)
    pm = L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_NONE, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                       contact=c, sinter=s13, precision=L.PRECISION_MIXED)
    with pytest.raises(L.HipgranError, match="FP64"):
        L.System(pm, x, v)
    with pytest.raises(L.HipgranError, match="single systems"):
        L.Pool(p, max_particles=64, max_systems=1)


def test_legacy_sintered_reference_kat():
    """test/sintered.cpp:126-152 at its full length, 10^6 steps of 1e-13 s:
    the bonded lattice stays rigid (Rg 3.95981e-08 m) while its centre of mass
    drifts 2.35702e-07 m, both within the reference's 1 %."""
    announce("test/sintered.cpp known answer, 10^6 steps")
    x, v = S.sinter_lattice()
    c, s13 = sintered_params()
    g, _ = make_pair(x, v, c, s13)
    c0 = x.mean(axis=0)
    for k in range(10):
        g.step(1e-13, 100000)
        print(f"[gpu-legacy] {100000 * (k + 1)} steps", flush=True)
    xf = g.state()["x"]
    cm = xf.mean(axis=0)
    rg = math.sqrt(float(((xf - cm) ** 2).sum()) / len(xf))
    drift = float(np.linalg.norm(cm - c0))
    print(f"[gpu-legacy] Rg {rg:.6g} (target 3.95981e-08), drift {drift:.6g} (target 2.35702e-07)", flush=True)
    assert abs(rg - 3.95981e-08) / 3.95981e-08 <= 0.01
    assert abs(drift - 2.35702e-07) / 2.35702e-07 <= 0.01
