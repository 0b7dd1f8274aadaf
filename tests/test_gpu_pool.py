"""Pools (batched / serving path) and mixed precision on the GPU.

FP64 pools must reproduce the oracle bit for bit per system, whatever else
shares the pool. Mixed precision must stay within the stated force tolerance
of the FP64 oracle (1e-5 relative) and give bit-identical results for a system
alone and in a mix.
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
    print(f"[gpu-pool] {name}", flush=True)


def agg_case(n, seed, df=1.8):
    L = _L()
    x = L.generate_aggregate(n, R, df, 1.3, seed)
    v0 = np.random.default_rng(seed).normal(0.0, 1.0, x.shape)
    return x, v0


def nl_params(precision=0):
    L = _L()
    c, s, h = S.aggregate_params()
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                         contact=c, sinter=s, hamaker=h, precision=precision)


def oracle_nl(x, v0):
    c, s, h = S.aggregate_params()
    return O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                          sinter=s, hamaker=h), x, v0)


def same(a, b, what):
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def test_pool_fp64_matches_oracle_per_system():
    announce("pool FP64 vs oracle, 5 aggregates")
    L = _L()
    cases = [agg_case(n, seed) for n, seed in ((50, 10), (137, 11), (400, 12), (1200, 13), (3000, 14))]
    pool = L.Pool(nl_params(), max_particles=6000, max_systems=8)
    slots = [pool.add(x, v) for x, v in cases]
    pool.step(DT, 200, rebuild_every=50)
    for (x, v), slot in zip(cases, slots):
        o = oracle_nl(x, v)
        o.step(DT, 200, rebuild_every=50)
        same(o.state(), pool.state(slot), f"slot {slot} n={x.shape[0]}")
        off1, idx1 = o.neighbors()
        off2, idx2 = pool.neighbors(slot)
        np.testing.assert_array_equal(off2, off1)
        np.testing.assert_array_equal(idx2, idx1)
        ij1, s1 = o.springs(1)
        ij2, s2 = pool.springs(slot, 1)
        np.testing.assert_array_equal(ij2, ij1)
        np.testing.assert_array_equal(s2, s1)
        assert pool.info(slot)[2] == 200


def test_pool_all_pairs_reference_lattices():
    """All-pairs pool: the contact+Hamaker lattice (test/hamaker.cpp) twice
    with different impactor speeds, next to each other."""
    announce("pool all pairs lattices")
    L = _L()
    x, v = S.lattice_with_impactor()
    v2 = v.copy()
    v2[-1] = [0.0, 2.0, -20.0]
    ham = O.hamaker_params(A=1e-20, h0=1e-9)
    p = L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=O.contact_params(), hamaker=ham)
    pool = L.Pool(p, max_particles=64, max_systems=4)
    s1, s2 = pool.add(x, v), pool.add(x, v2)
    pool.step(DT, 12000)
    for vv, slot in ((v, s1), (v2, s2)):
        o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_FULL, hamaker=ham), x, vv)
        o.step(DT, 12000)
        same(o.state(), pool.state(slot), f"lattice slot {slot}")


def test_pool_with_facets_matches_oracle():
    announce("pool facets vs oracle")
    L = _L()
    c, s, h = S.aggregate_params()
    sh = O.hamaker_params(A=1e-20, h0=1e-10)
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h, facet_contact=1, facet_hamaker=1, surface_contact=c,
                      surface_hamaker=sh)
    pool = L.Pool(p, max_particles=2000, max_systems=4, max_facets=4)
    runs = []
    for seed, speed in ((21, 4.0), (22, -3.0)):
        x, v0 = agg_case(300, seed)
        zmin, zmax = x[:, 2].min(), x[:, 2].max()
        a = 80 * R
        # plates 0.97 r from the outermost centres: in contact from the start
        lo, hi = zmin - 0.97 * R, zmax + 0.97 * R
        verts = np.array([[[-a, -a, lo], [a, -a, lo], [-a, a, lo]],
                          [[a, a, lo], [-a, a, lo], [a, -a, lo]],
                          [[-a, -a, hi], [-a, a, hi], [a, -a, hi]],
                          [[a, a, hi], [a, -a, hi], [-a, a, hi]]])
        vel = np.array([[0, 0, speed], [0, 0, speed], [0, 0, -speed], [0, 0, -speed]], dtype=float)
        slot = pool.add(x, v0)
        pool.set_facets(slot, verts, vel)
        runs.append((x, v0, verts, vel, slot))
    pool.step(DT, 500, rebuild_every=50)
    for x, v0, verts, vel, slot in runs:
        o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                           sinter=s, hamaker=h, facet_contact=1, facet_hamaker=1, surface_contact=c,
                                           surface_hamaker=sh), x, v0)
        o.set_facets(verts, vel)
        o.step(DT, 500, rebuild_every=50)
        same(o.state(), pool.state(slot), f"facet slot {slot}")


@pytest.mark.parametrize("precision", [0, 1])
def test_pool_result_alone_equals_in_mix(precision):
    announce(f"pool alone vs mix, precision {precision}")
    L = _L()
    x, v0 = agg_case(700, 31)
    alone = L.Pool(nl_params(precision), max_particles=1000, max_systems=2)
    sa = alone.add(x, v0)
    alone.step(DT, 300, rebuild_every=50)
    mix = L.Pool(nl_params(precision), max_particles=20000, max_systems=16)
    others = [agg_case(n, 40 + k) for k, n in enumerate((2500, 60, 900, 4000, 333))]
    for xo, vo in others[:3]:
        mix.add(xo, vo)
    sm = mix.add(x, v0)
    for xo, vo in others[3:]:
        mix.add(xo, vo)
    mix.step(DT, 300, rebuild_every=50)
    same(alone.state(sa), mix.state(sm), "alone vs mix")


@pytest.mark.parametrize("shape", [(512, 1), (256, 2), (256, 8)])
@pytest.mark.parametrize("precision", [0, 1])
def test_pool_packed_small_systems_equal_alone(precision, shape):
    """Small systems stepped side by side in one chunk workgroup (a pack) give
    the results each gives alone, with packing on and off, over whole and
    partial rebuild intervals. At 256 x 8 a workgroup's LDS share is too small
    for a pack's state, so its members are stepped one at a time."""
    announce(f"packed small systems, precision {precision}, shape {shape}")
    L = _L()
    sizes = (50, 64, 90, 120, 150, 200, 260, 333, 75, 410)
    cases = [agg_case(n, 70 + k) for k, n in enumerate(sizes)]
    alone = []
    for x, v0 in cases:
        P = L.Pool(nl_params(precision), max_particles=1000, max_systems=1)
        s = P.add(x, v0)
        P.step(DT, 120, rebuild_every=50)
        alone.append(P.state(s))
    if precision == 0:
        for k in (0, 7):
            o = oracle_nl(*cases[k])
            o.step(DT, 120, rebuild_every=50)
            same(o.state(), alone[k], f"oracle vs alone, n {sizes[k]}")
    for packing in (True, False):
        mix = L.Pool(nl_params(precision), max_particles=5000, max_systems=16)
        mix.set_chunk_shape(*shape)
        mix.set_packing(packing)
        slots = [mix.add(x, v0) for x, v0 in cases]
        mix.step(DT, 120, rebuild_every=50)
        for k, s in enumerate(slots):
            same(alone[k], mix.state(s), f"packing {packing}, n {sizes[k]}")


def test_pool_many_packs_equal_alone():
    """Thirty-two small systems packed best fit decreasing into several packs,
    among them packs held to the eight-member cap (eight 50-particle systems
    fill only 400 of 512 threads) beside systems too large to pack: each
    result equals the system alone, FP64, and one packed system equals the
    oracle."""
    announce("many packs")
    L = _L()
    sizes = [50] * 12 + [64, 90, 120, 150, 200, 260, 333, 410, 470, 500, 511] + [520, 700] + [75] * 7
    cases = [agg_case(n, 200 + k) for k, n in enumerate(sizes)]
    alone = []
    for x, v0 in cases:
        P = L.Pool(nl_params(0), max_particles=1000, max_systems=1)
        s = P.add(x, v0)
        P.step(DT, 60, rebuild_every=50)
        alone.append(P.state(s))
    o = oracle_nl(*cases[0])
    o.step(DT, 60, rebuild_every=50)
    same(o.state(), alone[0], "oracle vs alone")
    mix = L.Pool(nl_params(0), max_particles=int(sum(sizes)) + 64, max_systems=len(sizes))
    slots = [mix.add(x, v0) for x, v0 in cases]
    mix.step(DT, 60, rebuild_every=50)
    for k, s in enumerate(slots):
        same(alone[k], mix.state(s), f"n {sizes[k]} (system {k})")


def test_pool_continuous_admission_and_retirement():
    """Systems join and leave between steps; each result equals its own
    alone-run with the same number of steps."""
    announce("pool continuous batching")
    L = _L()
    pool = L.Pool(nl_params(1), max_particles=5000, max_systems=8)
    a = agg_case(800, 51)
    b = agg_case(1500, 52)
    c = agg_case(1200, 53)
    sa, sb = pool.add(*a), pool.add(*b)
    pool.step(DT, 100, rebuild_every=50)
    pool.remove(sa)
    sc = pool.add(*c)  # reuses a's slot and part of its range
    pool.step(DT, 150, rebuild_every=50)
    res_b, res_c = pool.state(sb), pool.state(sc)
    assert pool.info(sb)[2] == 250 and pool.info(sc)[2] == 150
    for (x, v), steps, res in ((b, 250, res_b), (c, 150, res_c)):
        solo = L.Pool(nl_params(1), max_particles=2000, max_systems=1)
        s = solo.add(x, v)
        solo.step(DT, steps, rebuild_every=50)
        same(solo.state(s), res, "continuous vs alone")


def _force_errors(a_gpu, a_ref, mags):
    diff = np.linalg.norm(a_gpu - a_ref, axis=1)
    normwise = np.linalg.norm(a_gpu - a_ref) / np.linalg.norm(a_ref)
    denom = np.where(mags > 0, mags, 1.0)
    per = diff / denom
    return normwise, per.max(), np.sqrt(np.mean(per ** 2))


def test_mixed_precision_force_error_full_size():
    """N=20000 aggregate (seed 1): one acceleration pass in MIXED against the
    FP64 oracle from the same evolved state. Tolerance (stated): 1e-5,
    normwise and per particle relative to sum_j |a_ij|."""
    announce("mixed precision force error N=20000")
    L = _L()
    x, v0 = agg_case(20000, 1)
    o = oracle_nl(x, v0)
    o.step(DT, 100, rebuild_every=50)
    st = o.state()
    # fresh systems on the evolved state: springs zero on both sides
    o2 = oracle_nl(st["x"], st["v"])
    o2.set_state(omega=st["omega"])
    g = L.System(nl_params(1), st["x"], st["v"], omega0=st["omega"])
    o2.compute_accelerations()
    g.compute_accelerations()
    ref = o2.state()
    got = g.state()
    mags = o2.pair_magnitudes()
    nw, pmax, prms = _force_errors(got["a"], ref["a"], mags[:, 0])
    nwa, pmaxa, prmsa = _force_errors(got["alpha"], ref["alpha"], mags[:, 1])
    print(f"[mixed] linear normwise {nw:.3e} per-particle max {pmax:.3e} rms {prms:.3e}; "
          f"angular normwise {nwa:.3e} max {pmaxa:.3e} rms {prmsa:.3e}", flush=True)
    assert nw < 1e-5 and pmax < 1e-5
    assert pmaxa < 1e-5


def test_mixed_precision_short_trajectory():
    """200 steps with springs building up: MIXED vs the FP64 oracle. Force
    error stays within 1e-5 (per particle, relative to sum_j |a_ij|) at every
    compared step; positions deviate by far less than a nanometre."""
    announce("mixed precision trajectory N=3000")
    L = _L()
    x, v0 = agg_case(3000, 2)
    o = oracle_nl(x, v0)
    g = L.System(nl_params(1), x, v0)
    worst = 0.0
    for k in range(4):
        o.step(DT, 50, rebuild_every=50)
        g.step(DT, 50, rebuild_every=50)
        ref, got = o.state(), g.state()
        mags = o.pair_magnitudes()[:, 0]
        _, pmax, _ = _force_errors(got["a"], ref["a"], mags)
        worst = max(worst, pmax)
        dx = np.abs(got["x"] - ref["x"]).max()
        print(f"[mixed] step {50 * (k + 1)}: per-particle force error max {pmax:.3e}, max |dx| {dx:.3e} m",
              flush=True)
    assert worst < 1e-5
    assert dx < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_pool_wide_launches_equal_chunk_kernel(precision):
    """The same systems stepped by the chunk kernel (wide threshold 0) and by
    the wide per-step launches (threshold 1): identical results, with facets,
    a system added mid-run (initial acceleration pass on the wide path) and a
    mixed pool where both paths run side by side."""
    announce(f"pool wide vs chunk, precision {precision}")
    L = _L()
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      precision=precision, contact=c, sinter=s, hamaker=h, facet_contact=1, facet_hamaker=1,
                      surface_contact=c, surface_hamaker=O.hamaker_params(A=1e-20, h0=1e-10))
    from libgran_amd import workloads as W
    cases = [agg_case(n, 70 + k) for k, n in enumerate((90, 700, 2600, 300))]
    late = agg_case(1500, 79)
    results = []
    for wide_min in (0, 1, 1000):
        pool = L.Pool(p, max_particles=8000, max_systems=8, max_facets=4)
        pool.set_wide_threshold(wide_min)
        slots = []
        for k, (x, v0) in enumerate(cases):
            slot = pool.add(x, v0)
            pool.set_facets(slot, *W.plates(x, R, 6.0, compress=k % 2 == 0))
            slots.append(slot)
        pool.step(DT, 150, rebuild_every=50)
        slots.append(pool.add(*late))
        pool.step(DT, 200, rebuild_every=50)
        results.append([pool.state(sl) for sl in slots])
        pool.close()
    for other in results[1:]:
        for k, (a, b) in enumerate(zip(results[0], other)):
            for f in ("x", "v", "theta", "omega", "a", "alpha"):
                np.testing.assert_array_equal(b[f], a[f], err_msg=f"system {k} field {f}")


def test_pool_wide_all_pairs_matches_oracle():
    """All-pairs mode on the wide path (threshold 1) against the oracle."""
    announce("pool wide all pairs vs oracle")
    L = _L()
    x, v = S.lattice_with_impactor()
    x[-1, 2] = 2.05 * R
    p = L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=O.contact_params(), hamaker=O.hamaker_params())
    pool = L.Pool(p, max_particles=64, max_systems=2)
    pool.set_wide_threshold(1)
    slot = pool.add(x, v)
    pool.step(DT, 3000)
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_FULL), x, v)
    o.step(DT, 3000)
    same(o.state(), pool.state(slot), "wide all pairs")


def test_pool_autotune_and_graph_do_not_change_results():
    """Results with the online autotuner running (its candidates switch the
    split every other interval) and with graphs off equal the plain run."""
    announce("pool autotune / graph invariance")
    L = _L()
    cases = [agg_case(n, 90 + k) for k, n in enumerate((3000, 1300, 150, 2200, 90))]
    outs = []
    for mode in ("plain", "tuned", "nograph"):
        pool = L.Pool(nl_params(1), max_particles=8000, max_systems=8)
        slots = [pool.add(x, v) for x, v in cases]
        if mode == "tuned":
            pool.autotune(True)
        if mode == "nograph":
            pool.set_wide_threshold(1000)
            pool.set_wide_graph(False)
        pool.step(DT, 1500, rebuild_every=50)  # the tuner's 26 intervals on these systems, and more
        if mode == "tuned":
            t = pool.tuning()
            assert t["tuned"], t
        outs.append([pool.state(s) for s in slots])
        pool.close()
    for other in outs[1:]:
        for k, (a, b) in enumerate(zip(outs[0], other)):
            for f in ("x", "v", "omega", "a"):
                np.testing.assert_array_equal(b[f], a[f], err_msg=f"system {k} field {f}")
