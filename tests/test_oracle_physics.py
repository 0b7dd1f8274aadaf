"""Closed-form and brute-force checks of the oracle where the reference's own
code cannot be run: the integrator and system loop (libtimestep is not part of
the reference tree), the neighbour-list predicate, and the physics the
functors encode (checked independently of their expressions).

Each check is chosen so that a plausible mistake fails it: a kick-drift order
other than kick-drift-kick fails the Stormer recurrence or the central-
difference velocity; a wrong reduced mass or a flipped damping sign moves the
coefficient of restitution; a Hamaker expression with a dropped term departs
from -dU/dh of the sphere-sphere potential; '<=' instead of '<' in the
neighbour predicate admits the lattice ties.
"""
import math

import numpy as np
import pytest

import _oracle as O
import _scenarios as S

R = O.R_PART
M = O.MASS
EPS = np.finfo(np.float64).eps


def system(binary_model, hamaker_model, x, v=None, w=None, **kw):
    return O.OracleSystem(O.system_params(binary_model, hamaker_model, **kw), x, v, w0=w)


def trajectory(sys_, dt, nsteps):
    """States after each of nsteps single steps (index 0 = before stepping,
    with the accelerations of the initial state)."""
    sys_.compute_accelerations()
    out = [sys_.state()]
    for _ in range(nsteps):
        sys_.step(dt, 1)
        out.append(sys_.state())
    return {k: np.array([s[k] for s in out]) for k in out[0]}


def test_free_flight_and_free_rotation():
    """Particles out of range of every force move on straight lines and spin
    at constant angular velocity: x = x0 + v0 t, theta = omega0 t."""
    x = np.array([[0.0, 0.0, 0.0], [20 * R, 0.0, 0.0], [0.0, 20 * R, 0.0]])
    v = np.array([[1.0, -2.0, 0.5], [0.0, 3.0, -1.0], [-0.25, 0.0, 0.0]])
    w = np.array([[1e9, 0.0, -2e9], [0.0, 5e8, 0.0], [3e9, 3e9, 3e9]])
    s = system(O.BIN_CONTACT, O.HAM_NONE, x, v, w)
    dt, n = 1e-13, 400
    s.step(dt, n)
    st = s.state()
    t = n * dt
    np.testing.assert_array_equal(st["v"], v)
    np.testing.assert_array_equal(st["omega"], w)
    np.testing.assert_allclose(st["x"], x + v * t, rtol=0, atol=1e3 * EPS * 20 * R)
    np.testing.assert_allclose(st["theta"], w * t, rtol=1e-12, atol=0)
    assert s.t == pytest.approx(t, rel=1e-12)


def test_kick_drift_kick_recurrences():
    """Velocity Verlet (kick-drift-kick, forces at the drifted positions):
    x[n+1] - 2 x[n] + x[n-1] = dt^2 a[n] and v[n] = (x[n+1] - x[n-1]) / (2 dt),
    and the same for theta / omega / alpha. Symplectic Euler satisfies the
    first but not the second; drift-kick-drift satisfies neither. The
    bonded lattice with moving columns exercises every force and torque."""
    x, v = S.sinter_lattice()
    s = system(O.BIN_SINTER, O.HAM_NONE, x, v, contact=O.contact_params(dt=5e-14),
               sinter=O.sinter_params(k_bond=1000.0, d_crit=1e-9, dt=5e-14))
    dt = 5e-14
    tr = trajectory(s, dt, 300)
    for pos, vel, acc in (("x", "v", "a"), ("theta", "omega", "alpha")):
        X, V, A = tr[pos], tr[vel], tr[acc]
        assert np.abs(A[1:-1]).max() > 0.0
        second = X[2:] - 2.0 * X[1:-1] + X[:-2]
        scale = np.abs(X).max()
        np.testing.assert_allclose(second, dt * dt * A[1:-1], rtol=1e-6, atol=64 * EPS * scale)
        central = (X[2:] - X[:-2]) / (2.0 * dt)
        np.testing.assert_allclose(central, V[1:-1], rtol=1e-6, atol=64 * EPS * scale / dt)


def test_first_step_uses_initial_accelerations():
    """The first step kicks with a(x0): x1 = x0 + v0 dt + a0 dt^2 / 2 and
    v1 = v0 + (a0 + a1) dt / 2."""
    x, v = S.sinter_lattice()
    kw = dict(contact=O.contact_params(dt=5e-14), sinter=O.sinter_params(k_bond=1000.0, d_crit=1e-9, dt=5e-14))
    dt = 5e-14
    s0 = system(O.BIN_SINTER, O.HAM_NONE, x + 0.01 * R * np.sin(np.arange(x.size)).reshape(x.shape), v, **kw)
    st0 = s0.state()
    s0.compute_accelerations()
    a0 = s0.state()["a"]
    s1 = system(O.BIN_SINTER, O.HAM_NONE, st0["x"], v, **kw)
    s1.step(dt, 1)
    st1 = s1.state()
    scale = np.abs(st0["x"]).max()
    np.testing.assert_allclose(st1["x"], st0["x"] + v * dt + 0.5 * a0 * dt * dt, rtol=0, atol=8 * EPS * scale)
    np.testing.assert_allclose(st1["v"], v + 0.5 * (a0 + st1["a"]) * dt, rtol=1e-12, atol=1e-12 * np.abs(v).max())


def test_bonded_pair_is_a_discrete_harmonic_oscillator():
    """An undamped bond stretched by s0 and released from rest: the gap
    s = d - 2r obeys s'' = -(2k/m) s, whose velocity-Verlet solution is
    s[n] = s0 cos(n theta) with cos(theta) = 1 - (omega0 dt)^2 / 2 exactly."""
    k = 1000.0
    sp = np.array([k, 0.0, k, 0.0, k, 0.0, k, 0.0, R, M, O.INERTIA, 5e-14, 1e-9])
    s0 = 2e-10
    x = np.array([[-(R + s0 / 2), 0.0, 0.0], [R + s0 / 2, 0.0, 0.0]])
    s = system(O.BIN_SINTER, O.HAM_NONE, x, sinter=sp, contact=O.contact_params(dt=5e-14))
    assert len(s.bonds()[0]) == 1
    dt, n = 5e-14, 4000
    tr = trajectory(s, dt, n)
    gap = (tr["x"][:, 1, 0] - tr["x"][:, 0, 0]) - 2.0 * R
    omega0 = math.sqrt(2.0 * k / M)
    theta = math.acos(1.0 - 0.5 * (omega0 * dt) ** 2)
    expect = s0 * np.cos(np.arange(n + 1) * theta)
    assert theta * n > 4 * math.pi  # at least two periods
    np.testing.assert_allclose(gap, expect, rtol=0, atol=1e-6 * s0)
    # energy of the shadow Hamiltonian stays bounded (no drift)
    e = 0.5 * M * (tr["v"][:, :, 0] ** 2).sum(axis=1) + 0.5 * k * gap ** 2
    assert np.ptp(e) < 2.0 * (omega0 * dt) ** 2 * e[0]


def _collide(s, dt, axis_gap, chunk=20, max_steps=200000):
    touched, steps = False, 0
    while steps < max_steps:
        s.step(dt, chunk)
        steps += chunk
        g = axis_gap(s.state())
        if g < 0:
            touched = True
        elif touched:
            return s.state()
    raise AssertionError("the collision did not end")


@pytest.mark.parametrize("zeta", [0.1, 0.3, 0.6])
def test_head_on_restitution_coefficient(zeta):
    """Linear spring-dashpot normal contact between equal spheres (reduced
    mass m/2): e = exp(-pi zeta / sqrt(1 - zeta^2)), zeta = gamma_n / sqrt(2 k m)."""
    k = 1.0e4
    gamma_n = zeta * math.sqrt(2.0 * k * M)
    dt = 1e-15
    cp = O.contact_params(k=k, gamma_n=gamma_n, dt=dt)
    g0 = 1e-12
    x = np.array([[-(R + g0 / 2), 0.0, 0.0], [R + g0 / 2, 0.0, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    s = system(O.BIN_CONTACT, O.HAM_NONE, x, v, contact=cp)
    st = _collide(s, dt, lambda q: (q["x"][1, 0] - q["x"][0, 0]) - 2.0 * R)
    e = -st["v"][0, 0] / 1.0
    expect = math.exp(-math.pi * zeta / math.sqrt(1.0 - zeta * zeta))
    assert e == pytest.approx(expect, rel=5e-3)
    np.testing.assert_allclose(st["v"].sum(axis=0), 0.0, atol=1e-12)


def test_facet_restitution_coefficient():
    """Sphere on an immobile facet (effective mass m):
    e = exp(-pi zeta / sqrt(1 - zeta^2)), zeta = gamma_n / (2 sqrt(k m))."""
    k, zeta = 1.0e4, 0.3
    gamma_n = 2.0 * zeta * math.sqrt(k * M)
    dt = 1e-15
    cp = O.contact_params(k=k, gamma_n=gamma_n, dt=dt)
    x, _, verts = S.surface_drop()
    x[0] = [-10.0 * R, -10.0 * R, R + 1e-12]  # over the face of the triangle
    v = np.array([[0.0, 0.0, -1.0]])
    s = system(O.BIN_CONTACT, O.HAM_NONE, x, v, contact=cp, facet_contact=1, facet_hamaker=0, surface_contact=cp)
    s.set_facets(verts, np.zeros((1, 3)))
    st = _collide(s, dt, lambda q: q["x"][0, 2] - R)
    expect = math.exp(-math.pi * zeta / math.sqrt(1.0 - zeta * zeta))
    assert st["v"][0, 2] == pytest.approx(expect, rel=5e-3)


def test_linear_momentum_is_conserved():
    """Pair forces are equal and opposite (every functor evaluated as (i, j)
    and (j, i)), so with equal masses sum(v) is constant."""
    import libgran_amd as L
    x = L.generate_aggregate(40, R, 1.8, 1.3, 11)
    v0 = np.random.default_rng(5).normal(0.0, 2.0, x.shape)
    contact, sinter, ham = S.aggregate_params()
    s = system(O.BIN_SINTER, O.HAM_LIMITED, x, v0, contact=contact, sinter=sinter, hamaker=ham)
    s.step(1e-13, 1500)
    st = s.state()
    assert np.abs(st["v"] - v0).max() > 1e-3  # the forces did act
    np.testing.assert_allclose(st["v"].sum(axis=0), v0.sum(axis=0), rtol=0, atol=1e-11 * np.abs(v0).sum())


def _hamaker_potential(A, r, h):
    """Sphere-sphere Hamaker energy for equal radii r at surface gap h."""
    d2 = (2.0 * r + h) ** 2
    return -A / 6.0 * (2.0 * r * r / (d2 - 4.0 * r * r) + 2.0 * r * r / d2 + math.log((d2 - 4.0 * r * r) / d2))


def test_hamaker_force_is_the_potential_gradient():
    """F = -dU/dd along the line of centres (attractive), with the gap
    clamped at h0 below it; the limited-range functor is the same force
    within the cut-off (strictly beyond it: zero)."""
    A, h0, cutoff = 1e-20, 1e-10, 3.0 * R
    ham = O.hamaker_params(A=A, h0=h0, r_cutoff=cutoff)
    gaps = np.geomspace(1.5 * h0, 2.0 * R, 40)
    st = np.zeros((gaps.size, 18))
    st[:, 3] = 2.0 * R + gaps
    full = O.oracle_hamaker_seq(O.HAM_FULL, ham, st)
    lim = O.oracle_hamaker_seq(O.HAM_LIMITED, ham, st)
    for g, out, out_l in zip(gaps, full, lim):
        step = 1e-4 * g
        dudh = (_hamaker_potential(A, R, g + step) - _hamaker_potential(A, R, g - step)) / (2.0 * step)
        fx = out[0] * M  # force on particle 0, which sits at the origin: +x is towards particle 1
        assert fx == pytest.approx(dudh, rel=1e-6)
        np.testing.assert_array_equal(out[6:9], -out[0:3])
        if 2.0 * R + g > cutoff:
            assert np.all(out_l[:3] == 0.0)
        else:
            np.testing.assert_array_equal(out_l, out)
    # clamp: any gap below h0 gives the force at h0
    st2 = np.zeros((3, 18))
    st2[:, 3] = 2.0 * R + np.array([h0, 0.5 * h0, -0.2 * h0])
    out2 = O.oracle_hamaker_seq(O.HAM_FULL, ham, st2)
    np.testing.assert_allclose(out2[1:, 0], out2[0, 0], rtol=1e-12)


def _brute_neighbours(x, r_verlet):
    d = x[:, None, :] - x[None, :, :]
    dist = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    off, idx = [0], []
    for i in range(x.shape[0]):
        js = [j for j in np.nonzero(dist[i] < r_verlet)[0] if j != i]
        idx += js
        off.append(len(idx))
    return np.array(off, dtype=np.uint64), np.array(idx, dtype=np.uint32)


@pytest.mark.parametrize("brute", [True, False])
def test_neighbour_list_predicate(brute):
    """norm(x_i - x_j) < r_verlet, ascending j, i excluded; both the brute
    force and the cell-list builds. The lattice (spacing 2r, r_verlet = 2r)
    puts every nearest-neighbour pair exactly on the boundary, which the
    strict predicate excludes."""
    rng = np.random.default_rng(8)
    cloud = rng.uniform(-12 * R, 12 * R, (300, 3))
    g = np.arange(4) * 2.0 * R
    lattice = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    for x, rv in ((cloud, 3.0 * R), (lattice, 2.0 * R), (lattice, 2.0 * R * 1.0000001)):
        s = system(O.BIN_CONTACT, O.HAM_NONE, x, neighbor_mode=1, r_verlet=rv)
        s.rebuild(brute=brute)
        off, idx = s.neighbors()
        off_b, idx_b = _brute_neighbours(x, rv)
        np.testing.assert_array_equal(off, off_b)
        np.testing.assert_array_equal(idx, idx_b)
    assert off_b[-1] == 0 or rv > 2.0 * R  # lattice at exactly 2r: no neighbours


def _closest_point_ericson(p, a, b, c):
    """Closest point of triangle abc to p by Voronoi regions (Ericson, Real-Time
    Collision Detection, 2005, sec. 5.1.5); independent of the facet code."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def _closest_on_segment(p, a, b):
    t = np.clip((p - a) @ (b - a) / ((b - a) @ (b - a)), 0.0, 1.0)
    return a + t * (b - a)


def _check_closest(verts, pts):
    """Returns (interior projections, edge fallbacks) after checking every point."""
    a, b, c = verts[0]
    got = O.oracle_closest_point(verts, pts)
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm)
    scale = np.abs(verts).max()
    inside = fallback = 0
    for p, q in zip(pts, got):
        exact = _closest_point_ericson(p, a, b, c)
        proj = p - ((p - a) @ nrm) * nrm
        if np.allclose(exact, proj, rtol=0, atol=1e-12 * scale):
            inside += 1
            if np.allclose(q, proj, rtol=0, atol=1e-12 * scale):
                continue
            edges = [_closest_on_segment(p, u, w) for u, w in ((a, b), (b, c), (c, a))]
            edge = min(edges, key=lambda e: (p - e) @ (p - e))
            np.testing.assert_allclose(q, edge, rtol=0, atol=1e-12 * scale)
            fallback += 1
        else:
            np.testing.assert_allclose(q, exact, rtol=0, atol=1e-12 * scale)
    return inside, fallback


def test_closest_point_on_triangle():
    """Against Ericson's region-based closest point. Where the plane
    projection falls outside the triangle the facet returns the closest
    point. Where it falls inside, triangular_facet.h:86-97 compares the sum
    of three sub-triangle areas with the facet area; rounding can make the
    sum exceed the area, and the facet then returns the nearest edge point
    (about one interior point in ten, also on the reference's own horizontal
    facet). The oracle reproduces that, so inside both outcomes are accepted
    and any other point fails."""
    rng = np.random.default_rng(4)
    inside = fallback = 0
    for _ in range(40):
        i, f = _check_closest(rng.normal(0.0, 1.0, (1, 3, 3)), rng.normal(0.0, 1.5, (50, 3)))
        inside, fallback = inside + i, fallback + f
    assert inside > 100 and fallback > 0
    # the reference's facet (test/surface.cpp:75-80)
    a_ = 50.0 * R
    verts = np.array([[[-a_, -a_, 0.0], [-a_, a_, 0.0], [a_, -a_, 0.0]]])
    pts = np.column_stack([rng.uniform(-2 * a_, 2 * a_, 800), rng.uniform(-2 * a_, 2 * a_, 800),
                           rng.uniform(-2 * R, 2 * R, 800)])
    i, f = _check_closest(verts, pts)
    assert i > 100
    # the face-contact points of the facet scenarios take the projection
    got = O.oracle_closest_point(verts, np.array([[-10.0 * R, -10.0 * R, 0.995 * R], [-10.0 * R, -10.0 * R, R]]))
    np.testing.assert_array_equal(got, [[-10.0 * R, -10.0 * R, 0.0], [-10.0 * R, -10.0 * R, 0.0]])
