"""Pools stepped by any number of steps, and systems that stop on their own.

hg_pool_step takes any step count: a launch runs every system up to the
nearest step at which one of them is due for a list rebuild, so calls of 20
and then 30 steps equal one call of 50, and systems added at different times
keep their own rebuild cadence. A system whose state goes non-finite, or whose
pair-state table overflows, stops alone: the other systems finish the call and
stay bit-identical to their runs without it.
"""
import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART
DT = 1e-13


def _L():
    import libgran_amd
    return libgran_amd


def announce(name):
    print(f"[gpu-pool-faults] {name}", flush=True)


def agg_case(n, seed, df=1.8):
    L = _L()
    x = L.generate_aggregate(n, R, df, 1.3, seed)
    v0 = np.random.default_rng(seed).normal(0.0, 1.0, x.shape)
    return x, v0


def nl_params(precision=0, **kw):
    L = _L()
    c, s, h = S.aggregate_params()
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                         contact=c, sinter=s, hamaker=h, precision=precision, **kw)


def same(a, b, what):
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def alone(params, x, v, calls, wide_min=None):
    """The system alone in a pool of its own, stepped by `calls`."""
    L = _L()
    pool = L.Pool(params, max_particles=x.shape[0] + 16, max_systems=1)
    if wide_min is not None:
        pool.set_wide_threshold(wide_min)
    slot = pool.add(x, v)
    for k in calls:
        pool.step(DT, k, rebuild_every=50)
    st = pool.state(slot)
    pool.close()
    return st


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_partial_intervals_equal_whole_ones(precision):
    """20 + 30 steps == 50 steps and 7 + 61 + 12 == 80, on a pool of mixed sizes
    with one system on the wide launches (threshold 2048)."""
    announce(f"partial intervals, precision {precision}")
    L = _L()
    cases = [agg_case(n, seed) for n, seed in ((60, 31), (333, 32), (1500, 33), (2600, 34))]
    total = sum(x.shape[0] for x, _ in cases) + 64

    def run(calls):
        pool = L.Pool(nl_params(precision), max_particles=total, max_systems=8)
        pool.set_wide_threshold(2048)
        slots = [pool.add(x, v) for x, v in cases]
        for k in calls:
            pool.step(DT, k, rebuild_every=50)
        out = [pool.state(s) for s in slots]
        steps = [pool.info(s)[2] for s in slots]
        pool.close()
        return out, steps

    whole, k1 = run([50])
    parts, k2 = run([20, 30])
    assert k1 == k2 == [50] * 4
    for i, (a, b) in enumerate(zip(whole, parts)):
        same(a, b, f"50 vs 20+30, system {i}")
    whole, _ = run([50, 30])
    parts, k3 = run([7, 61, 12])
    assert k3 == [80] * 4
    for i, (a, b) in enumerate(zip(whole, parts)):
        same(a, b, f"80 vs 7+61+12, system {i}")
    if precision == 0:
        for i, (x, v) in enumerate(cases):
            o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R,
                                               contact=S.aggregate_params()[0], sinter=S.aggregate_params()[1],
                                               hamaker=S.aggregate_params()[2]), x, v)
            o.step(DT, 80, rebuild_every=50)
            same(o.state(), parts[i], f"oracle, system {i}")


def test_systems_at_different_rebuild_phases():
    """A system added after 20 steps rebuilds on its own cadence (its steps 0,
    50, ...), the first one on its own: each equals its run alone."""
    announce("different rebuild phases")
    L = _L()
    (xa, va), (xb, vb) = agg_case(400, 41), agg_case(700, 42)
    p = nl_params(0)
    pool = L.Pool(p, max_particles=1200, max_systems=4)
    sa = pool.add(xa, va)
    pool.step(DT, 20, rebuild_every=50)
    sb = pool.add(xb, vb)
    pool.step(DT, 95, rebuild_every=50)
    assert pool.info(sa)[2] == 115 and pool.info(sb)[2] == 95
    same(alone(p, xa, va, [115]), pool.state(sa), "first system")
    same(alone(p, xb, vb, [95]), pool.state(sb), "second system")
    c, s, h = S.aggregate_params()
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                       sinter=s, hamaker=h), xb, vb)
    o.step(DT, 95, rebuild_every=50)
    same(o.state(), pool.state(sb), "second system vs oracle")


def overflowing(n, seed):
    """An aggregate with one particle at 1e300 m/s: its bonds' damping force
    overflows to Inf in the first acceleration pass (what FE_OVERFLOW traps in
    the reference's tests, test/contact.cpp:9-18), and NaN follows."""
    x, v = agg_case(n, seed)
    v = v.copy()
    v[3] = [1.0e300, 0.0, 0.0]
    return x, v


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_nonfinite_system_stops_alone(precision):
    """A system whose state overflows: hg_pool_step reports HG_ERR_NUMERIC,
    that system stops with HG_FAULT_NUMERIC, and the others complete every
    step bit-identical to their runs alone, before and after it is removed."""
    announce(f"non-finite system, precision {precision}")
    L = _L()
    good = [agg_case(n, seed) for n, seed in ((150, 51), (900, 52), (2500, 53))]
    xb, vb = overflowing(200, 54)
    p = nl_params(precision)
    total = sum(x.shape[0] for x, _ in good) + xb.shape[0] + 64
    pool = L.Pool(p, max_particles=total, max_systems=8)
    pool.set_wide_threshold(2048)
    slots = [pool.add(x, v) for x, v in good[:2]]
    bad = pool.add(xb, vb)
    slots.append(pool.add(*good[2]))
    with pytest.raises(L.HipgranError) as e:
        pool.step(DT, 120, rebuild_every=50)
    assert e.value.status == L._abi.HG_ERR_NUMERIC
    assert f"system {bad} " in str(e.value)
    assert pool.fault(bad) & L._abi.HG_FAULT_NUMERIC
    assert pool.info(bad)[2] == 50  # stopped at the end of its first interval
    for s, (x, v) in zip(slots, good):
        assert pool.fault(s) == 0
        assert pool.info(s)[2] == 120
        same(alone(p, x, v, [120], wide_min=2048), pool.state(s), f"slot {s} beside the faulted system")
    st = pool.state(bad)
    assert not np.all(np.isfinite(st["v"]))
    pool.remove(bad)
    pool.step(DT, 30, rebuild_every=50)
    for s, (x, v) in zip(slots, good):
        same(alone(p, x, v, [150], wide_min=2048), pool.state(s), f"slot {s} after the removal")


def test_pair_capacity_overflow_stops_one_system():
    """A pair-state table of 2 entries per particle: an overlapping 3x3x3
    lattice needs 6 and stops with HG_FAULT_PAIR_CAPACITY; head-on pairs beside
    it (one contact each) finish their steps as they do alone."""
    announce("pair capacity overflow isolation")
    L = _L()
    ham = O.hamaker_params(A=1e-20, h0=1e-9)
    p = L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=O.contact_params(), hamaker=ham, pair_capacity=2)
    lat = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], dtype=float) * 1.9 * R
    heads = []
    for q, speed in enumerate((3.0, 5.0)):
        x = np.array([[0.0, 0.0, 0.0], [2.05 * R, 0.1 * R * q, 0.0]])
        v = np.array([[speed, 0.0, 0.0], [-speed, 0.0, 0.0]])
        heads.append((x, v))
    pool = L.Pool(p, max_particles=64, max_systems=4)
    s0 = pool.add(*heads[0])
    sl = pool.add(lat, np.zeros_like(lat))
    s1 = pool.add(*heads[1])
    with pytest.raises(L.HipgranError) as e:
        pool.step(DT, 3000)
    assert e.value.status == L._abi.HG_ERR_CAPACITY
    assert "pair-state table needs 6" in str(e.value)
    assert pool.fault(sl) == L._abi.HG_FAULT_PAIR_CAPACITY
    for s, (x, v) in zip((s0, s1), heads):
        assert pool.fault(s) == 0 and pool.info(s)[2] == 3000
        ref = alone(p, x, v, [3000])
        same(ref, pool.state(s), f"head-on pair slot {s}")
        o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_FULL, hamaker=ham), x, v)
        o.step(DT, 3000)
        same(o.state(), pool.state(s), f"head-on pair slot {s} vs oracle")
    pool.remove(sl)
    pool.step(DT, 500)
    for s, (x, v) in zip((s0, s1), heads):
        same(alone(p, x, v, [3500]), pool.state(s), f"head-on pair slot {s} after the removal")


def test_single_system_nonfinite_reports_numeric():
    """hg_step reports HG_ERR_NUMERIC when the state overflows: FP64 per-step
    path at N=400, the one-launch small-system path at N=20, mixed precision
    at N=400."""
    announce("single system non-finite")
    L = _L()
    for n, seed, prec in ((400, 61, 0), (20, 62, 0), (400, 63, 1)):
        x, v = overflowing(n, seed)
        s = L.System(nl_params(prec), x, v)
        with pytest.raises(L.HipgranError) as e:
            s.step(DT, 10, rebuild_every=50)
        assert e.value.status == L._abi.HG_ERR_NUMERIC
        s.close()
