"""alt_breaking_sinter on the device against the oracle (FP64, bit for bit).

Every touching pair is bonded (no cycle check); each bond gets a neck
strength; an ordered bonded pair whose force |f_n n + f_t| exceeds it becomes a
contact pair with zero springs (alt_breaking_sinter_bridge.h:63-75, 117-119).
Strengths are drawn log-uniformly around the bonded forces of the scenario so
that part of the bonds break during the run.
"""
import sys

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
    print(f"[gpu-breaking] {name}", flush=True)
    sys.stdout.flush()


def same(a, b, what):
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(b[k], a[k], err_msg=f"{what}: field {k}")


def strengths(nbonds, seed, lo=3e-9, hi=3e-7):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), nbonds))


def test_breaking_lattice_all_pairs():
    """The alt_sintered lattice (54 bonds once cycles are allowed), shearing
    columns, all pairs."""
    announce("breaking lattice all pairs")
    L = _L()
    x, v = S.sinter_lattice()
    sp = O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14)
    cp = O.contact_params(dt=5e-14)
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_NONE, contact=cp, sinter=sp, sinter_breaking=1), x, v)
    g = L.System(L.make_params(L.BINARY_SINTER, L.HAMAKER_NONE, contact=cp, sinter=sp, sinter_breaking=1), x, v)
    pb, _ = g.bonds()
    np.testing.assert_array_equal(pb, o.bonds()[0])
    assert len(pb) == 54
    st = strengths(len(pb), 1)
    o.set_neck_strengths(st)
    g.set_neck_strengths(st)
    bonded0 = len(g.springs(1)[0])
    for chunk in range(4):
        o.step(5e-14, 1500)
        g.step(5e-14, 1500)
        same(o.state(), g.state(), f"chunk {chunk}")
    bonded1 = len(g.springs(1)[0])
    assert bonded1 < bonded0 == 2 * 54, "no bond broke"
    assert bonded1 > 0, "every bond broke"
    ij_o, s_o = o.springs(0)
    ij_g, s_g = g.springs(0)
    np.testing.assert_array_equal(ij_g, ij_o)
    np.testing.assert_array_equal(s_g, s_o)


def test_breaking_pool_aggregates_neighbor_list():
    """Two aggregates with limited-range Hamaker in one pool (one of them on
    the wide launches), against the oracle per system."""
    announce("breaking pool")
    L = _L()
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h, sinter_breaking=1)
    pool = L.Pool(p, max_particles=4000, max_systems=4)
    pool.set_wide_threshold(1000)
    runs = []
    for k, n in enumerate((400, 1500)):
        x = L.generate_aggregate(n, R, 1.8, 1.3, 60 + k)
        v0 = np.random.default_rng(60 + k).normal(0.0, 3.0, x.shape)
        slot = pool.add(x, v0)
        nb = len(pool.bonds(slot)[0])
        st = strengths(nb, 10 + k, lo=1e-9, hi=1e-7)
        pool.set_neck_strengths(slot, st)
        runs.append((x, v0, st, slot))
    pool.step(1e-13, 1000, rebuild_every=50)
    for x, v0, st, slot in runs:
        o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                           sinter=s, hamaker=h, sinter_breaking=1), x, v0)
        np.testing.assert_array_equal(pool.bonds(slot)[0], o.bonds()[0])
        o.set_neck_strengths(st)
        o.step(1e-13, 1000, rebuild_every=50)
        same(o.state(), pool.state(slot), f"slot {slot}")
        assert len(pool.springs(slot, 1)[0]) < 2 * len(pool.bonds(slot)[0]), "no bond broke"


def test_breaking_mixed_precision_runs_and_breaks():
    """Mixed precision compares |f| in FP32, so a bond near its strength may
    break a step apart from FP64; the run must break bonds and stay finite."""
    announce("breaking mixed")
    L = _L()
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h, sinter_breaking=1, precision=L.PRECISION_MIXED)
    x = L.generate_aggregate(800, R, 1.8, 1.3, 71)
    v0 = np.random.default_rng(71).normal(0.0, 3.0, x.shape)
    g = L.System(p, x, v0)
    nb = len(g.bonds()[0])
    g.set_neck_strengths(strengths(nb, 3, lo=1e-9, hi=1e-7))
    g.step(1e-13, 1000, rebuild_every=50)
    st = g.state()
    assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["v"]))
    assert 0 < len(g.springs(1)[0]) < 2 * nb


def test_neck_strengths_need_breaking_mode():
    L = _L()
    x, v = S.sinter_lattice()
    g = L.System(L.make_params(L.BINARY_SINTER, L.HAMAKER_NONE, sinter=O.sinter_params(d_crit=1e-9)), x, v)
    with pytest.raises(L.HipgranError):
        g.set_neck_strengths(np.ones(26))
