"""Small single systems step through one launch per rebuild interval
(k_tiny_steps: pair-parallel, up to 27 particles; k_small_steps: one particle
per thread, up to 256) instead of the per-step kernels, which hg_step uses
while launch profiling is on or when hg_set_per_step_launches asks for them.
All three must give the same state, bit for bit, for every functor mix:
positions, velocities, orientations, spins, accelerations and the pair-state
springs."""
import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART


def _L():
    import libgran_amd
    return libgran_amd


def run_both(params, x, v, steps, rebuild_every=0, legacy=False, facets=None, necks=None):
    L = _L()
    out = []
    for path in ("profiled", "per_step", "one_launch"):
        s = L.System(params, x, v)
        if legacy:
            vecs = np.random.default_rng(5).uniform(-1.0, 1.0, (5000, 3))
            pairs, conn, _ = O.legacy_frames(x, R, params.sinter.d_crit, vecs / np.linalg.norm(vecs, axis=1)[:, None])
            s.set_legacy_connectors(pairs, conn)
        if facets is not None:
            s.set_facets(*facets)
        if necks is not None:
            s.set_neck_strengths(necks)
        s.profile(path == "profiled")
        s.per_step_launches(path == "per_step")
        s.step(1e-13, steps, rebuild_every=rebuild_every)
        st = s.state()
        st["springs0"] = s.springs(0)[1]
        st["springs1"] = s.springs(1)[1]
        out.append(st)
        s.close()
    for other in out[1:]:
        for k in out[0]:
            np.testing.assert_array_equal(other[k], out[0][k], err_msg=k)
    return out[2]


@pytest.mark.parametrize("ham", ["FULL", "LIMITED", "SIMPLIFIED"])
@pytest.mark.parametrize("pairs", ["all", "list"])
def test_tiny_contact_lattice(ham, pairs):
    """The 19-particle lattice with the impactor (test/contact.cpp,
    test/hamaker*.cpp), contact + Hamaker, 3000 steps."""
    L = _L()
    x, v = S.lattice_with_impactor()
    kw = dict(pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=5.0 * R) if pairs == "list" else {}
    p = L.make_params(L.BINARY_CONTACT, getattr(L, f"HAMAKER_{ham}"), contact=O.contact_params(),
                      hamaker=O.hamaker_params(A=1e-20, h0=1e-9), **kw)
    st = run_both(p, x, v, 3000, rebuild_every=20 if pairs == "list" else 0)
    assert len(st["springs0"]) > 0  # contacts carried springs


def test_tiny_sinter_lattice_and_hamaker_only():
    """The bonded 27-particle lattice (alt_sinter), and Hamaker with no binary
    functor."""
    L = _L()
    x, v = S.sinter_lattice()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_NONE, contact=O.contact_params(dt=5e-14),
                      sinter=O.sinter_params(k_bond=1000.0, d_crit=1e-9, dt=5e-14))
    st = run_both(p, x, v, 2000)
    assert len(st["springs1"]) == 2 * 26
    x, v = S.lattice_with_impactor()
    run_both(L.make_params(L.BINARY_NONE, L.HAMAKER_FULL, hamaker=O.hamaker_params(A=1e-19, h0=1e-9)), x, v, 2000)


@pytest.mark.parametrize("pairs", ["all", "list"])
def test_tiny_breaking_sinter_lattice(pairs):
    """alt_breaking_sinter on the bonded 27-particle lattice with neck
    strengths drawn log-uniformly around its bond forces: part of the bonds
    break, their entries leave the table and the pairs go on as contacts."""
    L = _L()
    x, v = S.sinter_lattice()
    kw = dict(pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=4.0 * R) if pairs == "list" else {}
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_NONE, contact=O.contact_params(dt=5e-14),
                      sinter=O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14), sinter_breaking=1, **kw)
    necks = np.exp(np.random.default_rng(1).uniform(np.log(3e-9), np.log(3e-7), 54))
    st = run_both(p, x, v, 6000, rebuild_every=50 if pairs == "list" else 0, necks=necks)
    assert 0 < len(st["springs1"]) < 2 * 54  # some bonds broke, not all


def test_small_paths_legacy_facets_and_larger_systems():
    """The legacy sinter_functor lattice and a 200-particle aggregate
    (one particle per thread), and a particle on a facet."""
    L = _L()
    x, v = S.sinter_lattice()
    s13 = O.sinter_params(k_bond=1.0e4, d_crit=1.0e-9)
    s13[1] = 0.0
    run_both(L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_NONE, contact=O.contact_params(), sinter=s13), x, v,
             3000, legacy=True)
    xa = L.generate_aggregate(200, R, 1.8, 1.3, 0)
    va = np.random.default_rng(0).normal(0.0, 1.0, xa.shape)
    c, s, h = S.aggregate_params()
    for mode in ("all", "list"):
        kw = dict(pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R) if mode == "list" else {}
        run_both(L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, contact=c, sinter=s, hamaker=h, **kw), xa, va, 500,
                 rebuild_every=50 if mode == "list" else 0)
    xs, vs, verts = S.surface_drop()
    run_both(L.make_params(L.BINARY_CONTACT, L.HAMAKER_NONE, contact=O.contact_params(mu=0.5), facet_contact=1,
                           facet_hamaker=1, surface_contact=O.contact_params(mu=0.5),
                           surface_hamaker=O.hamaker_params(A=0.0, h0=1e-9)), xs, vs, 3000,
             facets=(verts, np.zeros((1, 3))))


def test_step_async_equals_step_and_defers_errors():
    """hg_step_async queued once per step, as a drop-in driver's do_step
    does, gives the state of one hg_step over the same steps; a table
    overflow that hg_step reports at once surfaces at the next state read."""
    L = _L()
    x, v = S.lattice_with_impactor()
    p = L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=O.contact_params(),
                      hamaker=O.hamaker_params(A=1e-20, h0=1e-9))
    a, b = L.System(p, x, v), L.System(p, x, v)
    a.step(1e-13, 400)
    for _ in range(400):
        b.step_async(1e-13, 1)
    sa, sb = a.state(), b.state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    x3 = np.array([[0.0, 0.0, 0.0], [1.9 * R, 0.0, 0.0], [-1.9 * R, 0.0, 0.0]])  # particle 0 touches two
    p1 = L.make_params(L.BINARY_CONTACT, L.HAMAKER_NONE, contact=O.contact_params(), pair_capacity=1)
    g = L.System(p1, x3, np.zeros_like(x3))
    with pytest.raises(L.HipgranError) as e:
        g.step(1e-13, 1)
    assert e.value.status == 5  # HG_ERR_CAPACITY
    g2 = L.System(p1, x3, np.zeros_like(x3))
    g2.step_async(1e-13, 1)
    with pytest.raises(L.HipgranError) as e:
        g2.state()
    assert e.value.status == 5
