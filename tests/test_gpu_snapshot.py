"""Snapshots (hg_save / hg_load, hg_pool_save / hg_pool_load).

A system saved at step k and reloaded continues exactly as the uninterrupted
run: state, accelerations, both spring tables, neighbour lists, bonds, time
and step count equal bit for bit at the end. Covered: a single FP64 system
with moving plates saved mid-interval, the same in mixed precision, breaking
bonds with finite neck strengths, the legacy sinter_functor with its
connectors, a pool system pre-empted (saved, removed, reloaded into the pool
while the others kept stepping), and a pool system reloaded as a single
system. Damaged bytes and a snapshot taken with other parameters are
refused.
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
    print(f"[gpu-snapshot] {name}", flush=True)


def agg(n, seed):
    L = _L()
    x = L.generate_aggregate(n, R, 1.8, 1.3, seed)
    return x, np.random.default_rng(seed).normal(0.0, 1.0, x.shape)


def plate_params(precision=0, **kw):
    L = _L()
    c, s, h = S.aggregate_params()
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                         contact=c, sinter=s, hamaker=h, precision=precision, facet_contact=1, facet_hamaker=1,
                         surface_contact=c, surface_hamaker=O.hamaker_params(A=1e-20, h0=1e-10), **kw)


def everything(sys_):
    st = sys_.state()
    off, idx = sys_.neighbors() if sys_.params is None or sys_.params.pair_mode == 1 else (None, None)
    out = dict(st)
    out["springs0"] = sys_.springs(0)
    out["springs1"] = sys_.springs(1)
    out["bonds"] = sys_.bonds()[0]
    out["time"] = sys_.time()
    if off is not None:
        out["nl"] = (off, idx)
    return out


def assert_same(a, b, what):
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    for k in ("springs0", "springs1"):
        np.testing.assert_array_equal(a[k][0], b[k][0], err_msg=f"{what}: {k} pairs")
        np.testing.assert_array_equal(a[k][1], b[k][1], err_msg=f"{what}: {k}")
    np.testing.assert_array_equal(a["bonds"], b["bonds"])
    assert a["time"] == b["time"], what
    if "nl" in a and "nl" in b:
        np.testing.assert_array_equal(a["nl"][0], b["nl"][0])
        np.testing.assert_array_equal(a["nl"][1], b["nl"][1])


def with_plates(sys_, x, speed=6.0):
    from libgran_amd import workloads as W
    verts, vel = W.plates(x, R, speed, compress=True)
    sys_.set_facets(verts, vel)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_single_system_resumes_bit_for_bit(precision):
    """Plates pressing a 1,500-particle aggregate: saved after 130 steps (in
    the middle of a rebuild interval), reloaded, 170 more steps; equals 300
    uninterrupted steps."""
    announce(f"single system, precision {precision}")
    L = _L()
    x, v = agg(1500, 81)
    p = plate_params(precision)
    ref = L.System(p, x, v)
    with_plates(ref, x)
    ref.step(DT, 300, rebuild_every=50)
    a = L.System(p, x, v)
    with_plates(a, x)
    a.step(DT, 130, rebuild_every=50)
    snap = a.save()
    a.close()
    b = L.System.load(snap)
    b.params = p
    assert b.time()[1] == 130
    b.step(DT, 170, rebuild_every=50)
    assert_same(everything(ref), everything(b), "resumed")
    assert b.time()[1] == 300
    np.testing.assert_array_equal(b.facets(4)[0], ref.facets(4)[0])


def test_breaking_bonds_resume_with_their_strengths():
    """alt_breaking_sinter with finite neck strengths. Saved before the first
    step, every break happens after the reload and depends on the strengths
    it carried; saved after 3 steps, the bonds broken by then (the weak ones
    break in the first force evaluation) must stay broken. Both reloads equal
    the uninterrupted 400 steps."""
    announce("breaking bonds")
    L = _L()
    x, v = agg(600, 82)
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h, sinter_breaking=1)
    ref = L.System(p, x, v)
    nb = len(ref.bonds()[0])
    strengths = np.random.default_rng(5).uniform(1e-8, 2e-7, nb)
    ref.set_neck_strengths(strengths)
    ref.step(DT, 400, rebuild_every=50)
    broken = 2 * nb - len(ref.springs(1)[0])
    for at in (0, 3):
        a = L.System(p, x, v)
        a.set_neck_strengths(strengths)
        a.step(DT, at, rebuild_every=50)
        broken_at = 2 * nb - len(a.springs(1)[0]) if at else 0
        b = L.System.load(a.save())
        b.params = p
        b.step(DT, 400 - at, rebuild_every=50)
        assert_same(everything(ref), everything(b), f"breaking, saved at step {at}")
        print(f"[gpu-snapshot] saved at step {at}: ordered bonds broken then {broken_at}, by step 400 {broken}",
              flush=True)
    assert broken > 0


def test_legacy_sinter_resumes_with_its_connectors():
    """The legacy sinter_functor lattice of test/sintered.cpp (27 particles,
    connectors moved by every step): saved after 333 steps, reloaded; the
    connectors and the state equal the uninterrupted run after 1,000."""
    announce("legacy sinter_functor")
    L = _L()
    xl, vl = [], []
    for i in range(3):
        for j in range(3):
            xx, yy = (-1.0 + 2.0 * j) * R, (-1.0 + 2.0 * i) * R
            xl += [[xx, yy, 0.0], [xx, yy, -2.0 * R], [xx, yy, -4.0 * R]]
            vl += [[0.0, 0.0, 5.0] if j == 0 else ([0.0, 5.0, 0.0] if j == 2 else [0.0, 0.0, 0.0])] * 3
    xl, vl = np.array(xl), np.array(vl)
    vecs = np.random.default_rng(3).uniform(-1.0, 1.0, (2000, 3))
    vecs /= np.linalg.norm(vecs, axis=1, keepdims=True)
    it = iter(vecs)
    pairs, conn = L.legacy_sinter_frames(xl, R, 1.0e-9, lambda: next(it))
    kd = 1.0e4
    sinter = [kd, 0.0, kd, 0.0, kd, 0.0, kd, 0.0, R, O.MASS, O.INERTIA, DT, 1.0e-9]
    p = L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_NONE, contact=O.contact_params(), sinter=sinter)
    ref = L.System(p, xl, vl)
    ref.set_legacy_connectors(pairs, conn)
    ref.step(DT, 1000)
    a = L.System(p, xl, vl)
    a.set_legacy_connectors(pairs, conn)
    a.step(DT, 333)
    b = L.System.load(a.save())
    b.params = p
    b.step(DT, 667)
    for k in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(b.state()[k], ref.state()[k], err_msg=k)
    np.testing.assert_array_equal(b.legacy_connectors(), ref.legacy_connectors())


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_pool_preemption_leaves_results_unchanged(precision):
    """A pool of three aggregates: after 70 steps the middle one is saved and
    removed, the others step 30 more, it is reloaded (into another slot and
    range) and all step 100 more. Each system equals itself run alone
    without interruption for its own step count."""
    announce(f"pool pre-emption, precision {precision}")
    L = _L()
    cases = [agg(n, seed) for n, seed in ((300, 83), (900, 84), (2500, 85))]
    p = plate_params(precision)
    pool = L.Pool(p, max_particles=6000, max_systems=4, max_facets=4)
    pool.set_wide_threshold(2048)
    slots = []
    from libgran_amd import workloads as W
    for x, v in cases:
        s = pool.add(x, v)
        pool.set_facets(s, *W.plates(x, R, 6.0, compress=True))
        slots.append(s)
    pool.step(DT, 70, rebuild_every=50)
    snap = pool.save(slots[1])
    pool.remove(slots[1])
    pool.step(DT, 30, rebuild_every=50)
    slots[1] = pool.load(snap)
    assert pool.info(slots[1])[2] == 70
    pool.step(DT, 100, rebuild_every=50)
    for k, ((x, v), s) in enumerate(zip(cases, slots)):
        steps = 170 if k == 1 else 200
        alone = L.Pool(p, max_particles=x.shape[0] + 16, max_systems=1, max_facets=4)
        alone.set_wide_threshold(2048)
        a = alone.add(x, v)
        alone.set_facets(a, *W.plates(x, R, 6.0, compress=True))
        alone.step(DT, steps, rebuild_every=50)
        got, want = pool.state(s), alone.state(a)
        for f in ("x", "v", "theta", "omega", "a", "alpha"):
            np.testing.assert_array_equal(got[f], want[f], err_msg=f"system {k}: {f}")
        np.testing.assert_array_equal(pool.springs(s, 1)[1], alone.springs(a, 1)[1])
        assert pool.info(s)[2] == steps
        alone.close()
    pool.close()


def test_pool_snapshot_loads_as_single_system():
    """A pool system saved after 120 steps continues as a single system and
    equals the oracle stepped 200 steps from the start (FP64)."""
    announce("pool snapshot to single system")
    L = _L()
    x, v = agg(800, 86)
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h)
    pool = L.Pool(p, max_particles=1000, max_systems=2)
    slot = pool.add(x, v)
    pool.step(DT, 120, rebuild_every=50)
    b = L.System.load(pool.save(slot))
    b.params = p
    b.step(DT, 80, rebuild_every=50)
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                       sinter=s, hamaker=h), x, v)
    o.step(DT, 200, rebuild_every=50)
    for f in ("x", "v", "theta", "omega", "a", "alpha"):
        np.testing.assert_array_equal(b.state()[f], o.state()[f], err_msg=f)
    np.testing.assert_array_equal(b.springs(1)[1], o.springs(1)[1])
    pool.close()


def test_damaged_or_foreign_snapshots_are_refused():
    announce("damaged and foreign snapshots")
    L = _L()
    x, v = agg(200, 87)
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h)
    a = L.System(p, x, v)
    a.step(DT, 20, rebuild_every=50)
    snap = bytearray(a.save())
    bad = bytearray(snap)
    bad[len(bad) // 2] ^= 0x10
    with pytest.raises(L.HipgranError) as e:
        L.System.load(bytes(bad))
    assert e.value.status == L._abi.HG_ERR_INVALID and "checksum" in str(e.value)
    with pytest.raises(L.HipgranError):
        L.System.load(bytes(snap[:-9]) + bytes(snap[-8:]))
    other = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.3 * R,
                          contact=c, sinter=s, hamaker=h)
    pool = L.Pool(other, max_particles=400, max_systems=2)
    with pytest.raises(L.HipgranError) as e:
        pool.load(bytes(snap))
    assert e.value.status == L._abi.HG_ERR_INVALID
    pool.close()
