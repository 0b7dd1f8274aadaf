"""The energy ledger on the GPU (hg_energy, hg_pool_energy) against the
oracle's, through the C ABI. The device's per-particle terms are summed on the
host in particle order, as the oracle sums its own; the trajectories are
bit-exact in FP64, so the ledgers differ only where the device's log, pow and
sqrt round differently from the host's libm: the stated tolerance is 1e-12 of
the ledger's scale (the sum of the magnitudes of its seven components) per
component. On the cases below the two ledgers have measured equal to the last
bit. test_oracle_energy.py holds the ledger to the forces."""
import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART
DT = 1e-13
TOL = 1e-12


def _L():
    import libgran_amd
    return libgran_amd


def announce(name):
    print(f"[gpu-energy] {name}", flush=True)


def as_array(d):
    return np.array([d[k] for k in _L().ENERGY_TERMS])


def same_ledger(dev, orc, what):
    dev, orc = as_array(dev), np.asarray(orc)
    scale = np.abs(orc[:7]).sum()
    err = np.abs(dev - orc) / scale
    print(f"[gpu-energy] {what}: max |dev - oracle| / scale {err.max():.2e}", flush=True)
    assert err.max() < TOL, f"{what}: {dev} vs {orc}"


def agg_case(n, seed, df=1.8):
    x = _L().generate_aggregate(n, R, df, 1.3, seed)
    return x, np.random.default_rng(seed).normal(0.0, 1.0, x.shape)


def agg_params(ham, precision=0):
    L = _L()
    c, s, h = S.aggregate_params()
    lib_ham = {O.HAM_FULL: L.HAMAKER_FULL, O.HAM_LIMITED: L.HAMAKER_LIMITED, O.HAM_SIMPLIFIED: L.HAMAKER_SIMPLIFIED}
    return L.make_params(L.BINARY_SINTER, lib_ham[ham], pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R, contact=c,
                         sinter=s, hamaker=h, precision=precision)


def agg_oracle(ham, x, v):
    c, s, h = S.aggregate_params()
    return O.OracleSystem(O.system_params(O.BIN_SINTER, ham, neighbor_mode=1, r_verlet=3.2 * R, contact=c, sinter=s,
                                          hamaker=h), x, v)


@pytest.mark.parametrize("ham", [O.HAM_FULL, O.HAM_LIMITED, O.HAM_SIMPLIFIED])
def test_energy_lattice_all_pairs(ham):
    """test/hamaker.cpp geometry: contact + Hamaker over all pairs, ledger at
    the start and after 3000 steps (contacts with springs by then)."""
    announce(f"lattice all pairs, hamaker model {ham}")
    L = _L()
    x, v = S.lattice_with_impactor()
    hp = O.hamaker_params(A=1e-20, h0=1e-9)
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, ham, hamaker=hp), x, v)
    lib_ham = {O.HAM_FULL: L.HAMAKER_FULL, O.HAM_LIMITED: L.HAMAKER_LIMITED, O.HAM_SIMPLIFIED: L.HAMAKER_SIMPLIFIED}
    g = L.System(L.make_params(L.BINARY_CONTACT, lib_ham[ham], contact=O.contact_params(), hamaker=hp), x, v)
    same_ledger(g.energy(O.MASS, O.INERTIA), o.energy(), "start")
    o.step(DT, 3000)
    g.step(DT, 3000)
    e = g.energy(O.MASS, O.INERTIA)
    same_ledger(e, o.energy(), "after 3000 steps")
    assert e["u_hamaker"] != 0.0


@pytest.mark.parametrize("ham", [O.HAM_LIMITED, O.HAM_SIMPLIFIED])
def test_energy_aggregate_neighbor_list(ham):
    """A 400-particle aggregate, alt_sinter + Hamaker over a 3.2r neighbour
    list rebuilt every 50 steps: every ledger term is live (bonds, contact
    springs, spin) after 300 steps."""
    announce(f"aggregate N=400 neighbour list, hamaker model {ham}")
    x, v = agg_case(400, 41)
    o = agg_oracle(ham, x, v)
    g = _L().System(agg_params(ham), x, v)
    o.step(DT, 300, rebuild_every=50)
    g.step(DT, 300, rebuild_every=50)
    e = g.energy(O.MASS, O.INERTIA)
    same_ledger(e, o.energy(), "after 300 steps")
    for k in ("ke_trans", "ke_rot", "u_normal", "u_tangential", "u_rolling", "u_hamaker"):
        assert e[k] != 0.0, k


@pytest.mark.parametrize("wide_min", [0, 1000])
def test_energy_pool(wide_min):
    """Three aggregates in one pool (chunk kernel; with wide threshold 1000
    the largest takes the wide launches): each system's ledger equals the
    oracle's of that system run alone, and a system's own ledger does not
    change when another is removed."""
    announce(f"pool, wide threshold {wide_min}")
    L = _L()
    cases = [agg_case(n, seed) for n, seed in ((120, 42), (450, 43), (1500, 44))]
    pool = L.Pool(agg_params(O.HAM_LIMITED), max_particles=4000, max_systems=4)
    slots = [pool.add(x, v) for x, v in cases]
    pool.set_wide_threshold(wide_min)
    pool.step(DT, 150, rebuild_every=50)
    ledgers = []
    for (x, v), slot in zip(cases, slots):
        o = agg_oracle(O.HAM_LIMITED, x, v)
        o.step(DT, 150, rebuild_every=50)
        ledgers.append(pool.energy(slot, O.MASS, O.INERTIA))
        same_ledger(ledgers[-1], o.energy(), f"system of {len(x)}")
    pool.remove(slots[0])
    for slot, before in zip(slots[1:], ledgers[1:]):
        np.testing.assert_array_equal(as_array(pool.energy(slot, O.MASS, O.INERTIA)), as_array(before))


def test_energy_pool_needs_a_listed_system():
    """A neighbour-list system added to a pool is listed by its first step;
    its ledger before that is refused, not read from a stale list."""
    announce("pool ledger before the first step")
    L = _L()
    x, v = agg_case(100, 45)
    pool = L.Pool(agg_params(O.HAM_LIMITED), max_particles=200, max_systems=2)
    slot = pool.add(x, v)
    with pytest.raises(L.HipgranError):
        pool.energy(slot, O.MASS, O.INERTIA)
    pool.step(DT, 50, rebuild_every=50)
    pool.energy(slot, O.MASS, O.INERTIA)


def test_energy_mixed_pool_tracks_fp64():
    """The ledger of a MIXED pool is computed in FP64 from the FP64 state and
    springs the mixed steps leave; over 150 steps of a 1500-particle aggregate
    it stays within 1e-6 of the FP64 oracle's ledger scale per component."""
    announce("mixed pool ledger N=1500")
    L = _L()
    x, v = agg_case(1500, 46)
    pool = L.Pool(agg_params(O.HAM_LIMITED, precision=1), max_particles=2000, max_systems=2)
    slot = pool.add(x, v)
    pool.step(DT, 150, rebuild_every=50)
    o = agg_oracle(O.HAM_LIMITED, x, v)
    o.step(DT, 150, rebuild_every=50)
    dev, orc = as_array(pool.energy(slot, O.MASS, O.INERTIA)), o.energy()
    err = np.abs(dev - orc) / np.abs(orc[:7]).sum()
    print(f"[gpu-energy] mixed vs FP64 ledger: max {err.max():.2e} per component {np.array2string(err, precision=2)}",
          flush=True)
    assert err.max() < 1e-6


@pytest.mark.parametrize("n,seed", [(777, 11), (20000, 1)])
def test_diagnostics_equal_the_reference_helpers(n, seed):
    """hg_diagnostics (one device reduction) against the oracle's sequential
    restatement of test/compute_energy.cpp:7-58 and
    test/mass_distribution.cpp:7-27 on a stepped aggregate (n not a multiple
    of the 1024-thread block). The device sums in another order, so each
    quantity is held to 1e-12 of its own scale: the kinetic energies and Rg
    (sums of positive terms) relative to themselves, |P|, |L| and the centre
    of mass relative to the sums of the magnitudes they are made of."""
    print(f"[gpu-energy] diagnostics n={n}", flush=True)
    L = _L()
    x = L.generate_aggregate(n, R, 1.8, 1.3, seed)
    v0 = np.random.default_rng(seed).normal(0.0, 1.0, x.shape)
    c, s, h = S.aggregate_params()
    g = L.System(L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                               r_verlet=3.2 * R, contact=c, sinter=s, hamaker=h), x, v0)
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                       sinter=s, hamaker=h), x, v0)
    g.step(DT, 300, rebuild_every=50)
    o.step(DT, 300, rebuild_every=50)
    st = o.state()
    np.testing.assert_array_equal(g.state()["v"], st["v"])
    assert np.abs(st["omega"]).max() > 0.0
    dg, do = g.diagnostics(O.MASS, O.INERTIA), o.diagnostics()
    for k in ("ke", "ke_trans", "ke_rot", "rg"):
        assert dg[k] == pytest.approx(do[k], rel=1e-12, abs=0.0), k
    p_scale = O.MASS * np.abs(st["v"]).sum()
    l_scale = O.INERTIA * np.abs(st["omega"]).sum() + O.MASS * (np.linalg.norm(st["x"], axis=1)
                                                               * np.linalg.norm(st["v"], axis=1)).sum()
    assert abs(dg["P"] - do["P"]) <= 1e-12 * p_scale
    assert abs(dg["L"] - do["L"]) <= 1e-12 * l_scale
    np.testing.assert_allclose(dg["cm"], do["cm"], rtol=0, atol=1e-12 * np.abs(st["x"]).max())
    # the quantities are what their names say, from the state itself
    assert do["ke_trans"] == pytest.approx(0.5 * O.MASS * (st["v"] ** 2).sum(), rel=1e-12)
    assert do["rg"] == pytest.approx(np.sqrt(((st["x"] - st["x"].mean(0)) ** 2).sum(1).mean()), rel=1e-12)


def test_thermo_series_as_the_reference_driver(tmp_path):
    """System.run writes energies.dat as main.cpp:142-159 does (header, one
    line every thermo period before the step, t = n dt, six significant
    digits) and VTK dumps as writer.cpp:51-54 names them; every line equals
    the oracle's diagnostics of the same state to the printed precision."""
    print("[gpu-energy] thermo series", flush=True)
    L = _L()
    x = L.generate_aggregate(300, R, 1.8, 1.3, 12)
    v0 = np.random.default_rng(12).normal(0.0, 1.0, x.shape)
    c, s, h = S.aggregate_params()
    g = L.System(L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                               r_verlet=3.2 * R, contact=c, sinter=s, hamaker=h), x, v0)
    path = tmp_path / "energies.dat"
    lines = g.run(DT, 1000, O.MASS, O.INERTIA, thermo_period=150, thermo_path=str(path), rebuild_every=50,
                  dump_period=400, dump_dir=str(tmp_path), r_part=R)
    assert lines == 7  # n = 0, 150, ..., 900
    assert sorted(p.name for p in tmp_path.glob("particles_*.vtk")) == [f"particles_{k}.vtk" for k in range(3)]
    text = path.read_text().splitlines()
    assert text[0] == "t\tE_trs\tE_rot\tE_tot\tP\tL"
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                       sinter=s, hamaker=h), x, v0)
    for k, line in enumerate(text[1:]):
        n = 150 * k
        if n:
            o.step(DT, 150, rebuild_every=50)
        d = o.diagnostics()
        cols = line.split("\t")
        assert cols[0] == f"{float(n) * DT:g}"
        got = [float(q) for q in cols[1:]]
        want = [d["ke_trans"], d["ke_rot"], d["ke"], d["P"], d["L"]]
        for a, b in zip(got, want):
            assert a == pytest.approx(b, rel=6e-6), f"line {k}: {line}"
    o.step(DT, 100, rebuild_every=50)
    np.testing.assert_array_equal(g.state()["x"], o.state()["x"])


def test_dissipated_work_equals_the_oracle():
    """hg_get_dissipated_work against the oracle's ledger: a damped glancing
    contact with spin (all pairs), and a damped aggregate through neighbour
    lists (bonds and contacts). The per-pair terms are the same FP64
    arithmetic and both sum per particle in pair order, then over particles in
    particle order: equal to the last bit, and the run itself unchanged."""
    print("[gpu-energy] dissipated work", flush=True)
    L = _L()
    dt = 1e-14
    g = 0.4e-9
    con = np.array([1.0e4, 2.0e-9, 1.0e4, g, 0.3, 0.8, 1.0e4, g, 0.3, 0.8, 1.0e4, g, 0.3, 0.8, R, O.MASS, O.INERTIA,
                    dt])
    x = np.array([[0.0, 0.0, 0.0], [np.sqrt((2.0 * R + 1e-11) ** 2 - (0.4 * R) ** 2), 0.4 * R, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    w = np.array([[0.0, 0.0, 3e7], [2e7, 0.0, -1e7]])
    cases = [("glancing contact", L.make_params(L.BINARY_CONTACT, L.HAMAKER_NONE, contact=con),
              O.system_params(O.BIN_CONTACT, O.HAM_NONE, contact=con), x, v, w, 3000, 0)]
    xa = L.generate_aggregate(500, R, 1.8, 1.3, 31)
    va = np.random.default_rng(31).normal(0.0, 2.0, xa.shape)
    c, s, h = S.aggregate_params()
    cases.append(("aggregate, lists", L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                                                     r_verlet=3.2 * R, contact=c, sinter=s, hamaker=h),
                  O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                  sinter=s, hamaker=h), xa, va, None, 300, 50))
    for name, pg, po, x0, v0, w0, steps, every in cases:
        gs = L.System(pg, x0, v0, omega0=w0)
        o = O.OracleSystem(po, x0, v0, w0=w0)
        gs.track_dissipation()
        o.track_dissipation()
        ddt = 1e-14 if every == 0 else 1e-13
        gs.step(ddt, steps, rebuild_every=every)
        o.step(ddt, steps, rebuild_every=every)
        wg, wo = gs.dissipated_work(), o.dissipated_work()
        print(f"[gpu-energy] {name}: {wg}", flush=True)
        assert wo["total"] > 0.0
        for k in wo:
            assert wg[k] == wo[k], (name, k, wg[k], wo[k])
        st = o.state()
        for f in ("x", "v", "omega"):
            np.testing.assert_array_equal(gs.state()[f], st[f], err_msg=f"{name}: {f}")
    bad = L.System(L.make_params(L.BINARY_CONTACT, L.HAMAKER_NONE, contact=con, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                                 r_verlet=3.2 * R, precision=L.PRECISION_MIXED), xa, va)
    with pytest.raises(L.HipgranError):
        bad.track_dissipation()
