"""Independent pins of the oracle's friction, surface, bond-break, fold and
topology logic (CPU only).

Each check compares the oracle with a closed form, a symmetry or a brute-force
computation that does not restate the oracle's expressions, and is chosen so
that a plausible mistake fails it:

* Luding's slip model (contact_force.h:105-147) driven at a constant normal
  load and a constant sliding, rolling or twisting velocity produces an exact
  stick-slip sequence: a dynamic step returns mu*phi*f_n; the next step, with
  the spring just reset to -(mu*phi*f_n t + gamma v)/k, returns mu*phi*f_n
  again; each further static step adds k |v| dt; the first value above
  mu_s f_n resets. A branch swap, a dropped gamma*v in the reset or a cap at
  mu_s instead of mu*phi breaks the sequence.
* The spring identities of every active step of random contact sequences:
  static steps advance the rotated spring by v dt, dynamic ones reset it as
  above, and no tangential force exceeds mu_s f_n.
* Exchanging i and j negates the tangential and torsion springs and leaves the
  rolling spring unchanged (bitwise for the last two); forces obey Newton's
  third law.
* The sphere-plane Hamaker force equals -dU/dh of the textbook sphere-plane
  potential, and is constant below h0.
* A bond breaks at the first evaluation whose force exceeds its strength;
  that evaluation is still the bonded one, and the contact functor with zero
  springs takes every later step.
* The binary container's result is the sum of its functors.
* The bond graph equals an independent union-find scan in the reference's
  lexicographic order.
* The legacy sinter_functor's five connector springs give 5 k_d delta for an
  axial stretch and k_d u (torque r k_d u) for a transverse shift.
"""
import os

import numpy as np
import pytest

import _oracle as O
import _scenarios as S
import _seq

R = O.R_PART
M = O.MASS
I = O.INERTIA
DT = 1e-13
FIXTURE = np.load(os.path.join(os.path.dirname(__file__), "golden", "aggregate150.npz"))


# ----------------------------------------------------------- helpers --------
def contact_p(k=1.0e4, gamma_n=5e-9, k_t=1.0e4, gamma_t=1e-9, mu=0.5, phi=0.6, k_r=1.0e4, gamma_r=1e-9,
              mu_r=0.3, phi_r=0.5, k_o=1.0e4, gamma_o=1e-9, mu_o=0.2, phi_o=0.4, dt=DT):
    """contact_force_functor ctor order (contact_force.h:10-30)."""
    return np.array([k, gamma_n, k_t, gamma_t, mu, phi, k_r, gamma_r, mu_r, phi_r, k_o, gamma_o, mu_o, phi_o,
                     R, M, I, dt])


def pair_states(nsteps, xi, xj, vi, vj, wi, wj):
    row = np.concatenate([xi, xj, vi, vj, wi, wj])
    return np.tile(row, (nsteps, 1))


def stick_slip(k, gamma, v, f_n, mu, phi, dt, nsteps):
    """The magnitudes Luding's slip model returns at constant f_n and relative
    velocity v (zero spring at the start), derived by hand: static steps
    return |k xi + gamma v| with xi growing by v dt; the first value above
    mu f_n is a dynamic step returning mu phi f_n and resetting the spring so
    that the next step's force is exactly mu phi f_n again."""
    out, kind = [], []
    m = 0  # static steps since the spring was zero (ramp) or reset (cycle)
    base = gamma * v  # |f0| on the first step of the ramp
    for _ in range(nsteps):
        f0 = base + m * k * v * dt
        if f0 <= mu * f_n:
            out.append(f0)
            kind.append("static")
            m += 1
        else:
            out.append(mu * phi * f_n)
            kind.append("dynamic")
            base, m = mu * phi * f_n, 0
    return np.array(out), kind


def assert_stick_slip(got, want, kind):
    assert "dynamic" in kind and kind.count("static") > kind.count("dynamic")
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=0)


# ------------------------------------------------------------- Luding -------
def test_sliding_contact_stick_slip_closed_form():
    """Two spheres at a fixed overlap of 0.1 nm sliding past each other at
    7 m/s: the tangential force magnitude follows the closed-form stick-slip
    sequence, opposes the sliding and is never above mu_s f_n; the normal
    force is k delta."""
    p = contact_p()
    delta, v = 1e-10, 7.0
    xi, xj = np.zeros(3), np.array([0.0, 0.0, 2 * R - delta])
    st = pair_states(200, xi, xj, np.array([v, 0.0, 0.0]), np.zeros(3), np.zeros(3), np.zeros(3))
    out, springs = O.oracle_contact_seq(p, st)
    f = out[:, 0:3] * M
    n = np.array([0.0, 0.0, -1.0])
    f_n = p[0] * delta
    np.testing.assert_allclose(f @ n, f_n, rtol=1e-9)
    ft = f - np.outer(f @ n, n)
    mag = np.linalg.norm(ft, axis=1)
    want, kind = stick_slip(p[2], p[3], v, f_n, p[4], p[5], DT, 200)
    assert_stick_slip(mag, want, kind)
    assert np.all(ft[:, 0] < 0) and np.all(np.abs(ft[:, 1:]) == 0.0)
    assert mag.max() <= p[4] * f_n * (1 + 1e-12)
    # no rolling or twisting: no torque beyond the tangential force's arm
    rpp = R - delta / 2
    np.testing.assert_allclose(out[:, 3:6] * I, -rpp * np.cross(n, ft), rtol=1e-9, atol=1e-40)


def test_rolling_contact_stick_slip_closed_form():
    """Counter-rotating spheres (w_i = -w_j normal to the contact) at a fixed
    overlap: no sliding or twisting velocity, a rolling velocity 2 r' |w|, and
    the rolling resistance torque r n x f_r whose magnitude follows the
    stick-slip sequence of the rolling spring's constants."""
    p = contact_p()
    delta = 1e-10
    rpp = R - delta / 2
    w = 7.0 / (2 * rpp)
    xi, xj = np.zeros(3), np.array([0.0, 0.0, 2 * R - delta])
    st = pair_states(200, xi, xj, np.zeros(3), np.zeros(3), np.array([0.0, w, 0.0]), np.array([0.0, -w, 0.0]))
    out, springs = O.oracle_contact_seq(p, st)
    assert np.all(out[:, 0:2] == 0.0)  # no tangential force
    tau = out[:, 3:6] * I
    mag = np.linalg.norm(tau, axis=1) / R
    f_n = p[0] * delta
    want, kind = stick_slip(p[6], p[7], 2 * rpp * w, f_n, p[8], p[9], DT, 200)
    assert_stick_slip(mag, want, kind)
    # rolling and sliding springs stay separate: the sliding spring stays zero
    assert np.all(springs[:, 0:3] == 0.0)


@pytest.mark.parametrize("w", [1.0e6, 1.0e11])
def test_twisting_contact_closed_form(w):
    """Spheres spinning oppositely about their normal: the torsion spring lies
    along n, so the rotation into the contact plane (contact_force.h:115-122)
    zeroes it before every evaluation and the torque is r f_o with |f_o| =
    gamma_o 2 r w while that is at most mu_o f_n (stored spring: v_o dt),
    else mu_o phi_o f_n (stored spring: the dynamic reset)."""
    p = contact_p()
    delta = 1e-10
    n = np.array([0.0, 0.0, -1.0])
    xi, xj = np.zeros(3), np.array([0.0, 0.0, 2 * R - delta])
    st = pair_states(50, xi, xj, np.zeros(3), np.zeros(3), w * n, -w * n)
    out, springs = O.oracle_contact_seq(p, st)
    tau = out[:, 3:6] * I
    f_n = p[0] * delta
    v_o = 2 * R * w
    k_o, g_o, mu_o, phi_o = p[10], p[11], p[12], p[13]
    static = g_o * v_o <= mu_o * f_n
    want = g_o * v_o if static else mu_o * phi_o * f_n
    np.testing.assert_allclose(np.linalg.norm(tau, axis=1) / R, want, rtol=1e-12)
    np.testing.assert_allclose(tau / np.linalg.norm(tau, axis=1)[:, None], np.tile(-n, (50, 1)), atol=1e-12)
    xi_o = springs[:, 6:9]
    if static:
        np.testing.assert_allclose(xi_o, np.tile(v_o * n * DT, (50, 1)), rtol=1e-12)
    else:
        np.testing.assert_allclose(k_o * xi_o, np.tile(-(mu_o * phi_o * f_n * -n + g_o * v_o * n), (50, 1)),
                                   rtol=1e-12)
    assert static == (w < 1e10)


def test_facet_sliding_stick_slip_closed_form():
    """A sphere pressed 0.1 nm into a resting facet and sliding along it at
    7 m/s (surface_contact_force.h:57-147): the same closed-form stick-slip
    sequence, with f_n = k delta."""
    p = contact_p()
    a = 50 * R
    verts = np.array([[-a, -a, 0.0], [a, -a, 0.0], [-a, a, 0.0]])
    delta, v = 1e-10, 7.0
    st = np.tile(np.array([0.0, 0.0, R - delta, v, 0.0, 0.0, 0.0, 0.0, 0.0]), (200, 1))
    out = O.oracle_facet_seq(verts.ravel(), np.zeros(3), p, O.hamaker_params(A=0.0), 0.0, st)
    f = out[:, 0:3] * M
    f_n = p[0] * delta
    np.testing.assert_allclose(f[:, 2], f_n, rtol=1e-9)
    mag = np.linalg.norm(f[:, 0:2], axis=1)
    want, kind = stick_slip(p[2], p[3], v, f_n, p[4], p[5], DT, 200)
    assert_stick_slip(mag, want, kind)
    assert np.all(f[:, 0] < 0)


def kinematics(row, p):
    """Contact kinematics of one ordered pair state (contact_force.h:65-84):
    normal, f_n and the three relative velocities."""
    xi, xj, vi, vj, wi, wj = (row[3 * q:3 * q + 3] for q in range(6))
    d = xi - xj
    n = d / np.linalg.norm(d)
    ov = 2 * R - d @ n
    rpp = R - ov / 2
    f_n = p[0] * ov + p[1] * (-(vi - vj) @ n)
    v_ij = (vi - vj) + rpp * np.cross(n, wi) + rpp * np.cross(n, wj)
    v_t = v_ij - (v_ij @ n) * n
    v_r = -rpp * (np.cross(n, wi) - np.cross(n, wj))
    v_o = R * (n @ wi - n @ wj) * n
    return n, ov, rpp, f_n, (v_t, v_r, v_o)


def branch_violations(p, states, out, springs):
    """Checks every active step of the ordered pair (0, 1) against the slip
    model's branch identities: the stored springs of all three models, and
    the returned force and its cap for the sliding one (the rolling and
    torsion forces share the torque and are not separable in a general
    state; the closed-form tests above isolate them). Returns (violations,
    static count, dynamic count) per spring (0 sliding, 1 rolling, 2
    torsion)."""
    consts = [(p[2], p[3], p[4], p[5]), (p[6], p[7], p[8], p[9]), (p[10], p[11], p[12], p[13])]
    bad, nstat, ndyn = [], [0, 0, 0], [0, 0, 0]
    prev = np.zeros(9)
    for k, row in enumerate(states):
        n, ov, rpp, f_n, vel = kinematics(row, p)
        new = springs[k, 0:9]
        if ov <= 0:
            prev = np.zeros(9)
            continue
        # the tangential force recovered from the linear output
        f = out[k, 0:3] * M
        f_t = f - (f @ n) * n
        for q, (kq, gq, muq, phiq) in enumerate(consts):
            xp, xn, v = prev[3 * q:3 * q + 3], new[3 * q:3 * q + 3], vel[q]
            proj = xp - (xp @ n) * n
            rot = proj / np.linalg.norm(proj) * np.linalg.norm(xp) if np.linalg.norm(proj) > 0 else proj
            f0 = -kq * rot - gq * v
            scale = max(np.linalg.norm(f0), abs(f_n), 1e-30)
            that = f0 / np.linalg.norm(f0) if np.linalg.norm(f0) > 0 else np.zeros(3)
            fq = np.linalg.norm(f_t) if q == 0 else None
            if np.linalg.norm(f0) <= muq * f_n:
                nstat[q] += 1
                ok = np.allclose(xn, rot + v * DT, rtol=0, atol=1e-12 * (np.linalg.norm(xn) + np.linalg.norm(v) * DT))
                if fq is not None:
                    ok &= abs(fq - np.linalg.norm(f0)) <= 1e-9 * scale
            else:
                ndyn[q] += 1
                ok = np.allclose(kq * xn, -(muq * phiq * f_n * that + gq * v), rtol=0, atol=1e-9 * scale)
                if fq is not None:
                    ok &= abs(fq - abs(muq * phiq * f_n)) <= 1e-9 * scale
            if fq is not None and f_n > 0:
                ok &= fq <= muq * f_n * (1 + 1e-9) + 1e-9 * scale
            if not ok:
                bad.append((k, q))
        prev = new.copy()
    return bad, nstat, ndyn


def slip_fixture():
    p = contact_p(k=2.0e4, mu=0.05, phi=0.7, mu_r=0.02, mu_o=0.01)
    runs = []
    for seed in range(3):
        st = _seq.pair_sequence(100 + seed, v_sigma=20.0)
        out, springs = O.oracle_contact_seq(p, st)
        runs.append((st, out, springs))
    return p, runs


def test_contact_branch_identities_on_random_sequences():
    """Contact formation, sliding, rolling, twisting and separation with
    random velocities: every active step satisfies the static or the dynamic
    identity, both occur for every spring, and a separated step resets the
    springs and the force (contact_force.h:68-71)."""
    p, runs = slip_fixture()
    tot_s, tot_d = np.zeros(3, int), np.zeros(3, int)
    for st, out, springs in runs:
        bad, ns, nd = branch_violations(p, st, out, springs)
        assert not bad, bad[:5]
        tot_s += ns
        tot_d += nd
        d = np.linalg.norm(st[:, 3:6] - st[:, 0:3], axis=1)
        sep = d > 2 * R * (1 + 1e-9)
        assert sep.sum() > 20 and np.all(out[sep] == 0.0) and np.all(springs[sep] == 0.0)
    assert np.all(tot_s > 10) and np.all(tot_d > 10), (tot_s, tot_d)


@pytest.mark.parametrize("plant", ["swap", "no_gamma_v", "cap_mu_s"])
def test_branch_identities_catch_planted_mistakes(plant):
    """The identity check above fails when the recorded steps carry one of
    three planted mistakes: the static update on dynamic steps, a reset
    without the gamma*v term, or dynamic forces capped at mu_s f_n."""
    p, runs = slip_fixture()
    st, out, springs = runs[0]
    out, springs = out.copy(), springs.copy()
    consts = [(p[2], p[3], p[4], p[5])]
    prev = np.zeros(9)
    planted = 0
    for k, row in enumerate(st):
        n, ov, rpp, f_n, vel = kinematics(row, p)
        if ov <= 0:
            prev = np.zeros(9)
            continue
        kq, gq, muq, phiq = consts[0]
        xp, v = prev[0:3], vel[0]
        proj = xp - (xp @ n) * n
        rot = proj / np.linalg.norm(proj) * np.linalg.norm(xp) if np.linalg.norm(proj) > 0 else proj
        f0 = -kq * rot - gq * v
        nxt = springs[k, 0:9].copy()
        if np.linalg.norm(f0) > muq * f_n and f_n > 0:
            that = f0 / np.linalg.norm(f0)
            planted += 1
            if plant == "swap":
                springs[k, 0:3] = rot + v * DT
            elif plant == "no_gamma_v":
                springs[k, 0:3] = -(muq * phiq * f_n * that) / kq
            else:
                f = out[k, 0:3] * M
                ft = f - (f @ n) * n
                out[k, 0:3] = ((f @ n) * n + ft / np.linalg.norm(ft) * muq * f_n) / M
        prev = nxt
    assert planted > 0
    bad, _, _ = branch_violations(p, st, out, springs)
    assert bad


def test_exchange_symmetry_of_springs_and_forces():
    """(i, j) and (j, i) of the same contact: the sliding and torsion springs
    are opposite and the rolling spring equal (bitwise for rolling and
    torsion, whose relative velocities are computed without reordering),
    and the linear forces obey Newton's third law."""
    p = contact_p(k=2.0e4, mu=0.3, phi=0.7)
    for seed in range(4):
        st = _seq.pair_sequence(300 + seed, v_sigma=5.0)
        out, sp = O.oracle_contact_seq(p, st)
        scale = np.abs(sp[:, 0:3]).max()
        assert scale > 0
        np.testing.assert_allclose(sp[:, 0:3], -sp[:, 9:12], rtol=0, atol=1e-11 * scale)
        np.testing.assert_array_equal(sp[:, 3:6], sp[:, 12:15])
        np.testing.assert_array_equal(sp[:, 6:9], -sp[:, 15:18])
        fs = np.abs(out[:, 0:3]).max()
        np.testing.assert_allclose(out[:, 0:3], -out[:, 6:9], rtol=0, atol=1e-11 * fs)


# ---------------------------------------------------- surface Hamaker -------
def sphere_plane_u(A, r, h):
    """Hamaker sphere-plane potential (Hamaker 1937): the pairwise r^-6
    attraction integrated over a sphere of radius r at gap h and a half-space."""
    return -A / 6.0 * (r / h + r / (h + 2 * r) + np.log(h / (h + 2 * r)))


def test_surface_hamaker_is_the_sphere_plane_gradient():
    """surface_hamaker_force.h:18-30: the force on a sphere at gap h above a
    facet equals -dU/dh of the sphere-plane potential (central differences),
    points at the facet, and is constant below the saturation gap h0."""
    A, h0 = 1.0e-20, 1.0e-10
    a = 50 * R
    verts = np.array([[-a, -a, 0.0], [a, -a, 0.0], [-a, a, 0.0]])
    gaps = np.concatenate([np.geomspace(0.2 * h0, 0.9 * h0, 5), np.geomspace(1.5 * h0, 4 * R, 40)])
    st = np.zeros((len(gaps), 9))
    st[:, 2] = R + gaps
    out = O.oracle_facet_seq(verts.ravel(), np.zeros(3), contact_p(), O.hamaker_params(A=A, h0=h0), 0.0, st)
    f = out[:, 0:3] * M
    assert np.all(f[:, 0:2] == 0.0) and np.all(f[:, 2] < 0)
    for h, fz in zip(gaps, f[:, 2]):
        hh = max(h, h0)
        e = 1e-5 * hh
        want = -(sphere_plane_u(A, R, hh + e) - sphere_plane_u(A, R, hh - e)) / (2 * e)
        assert fz == pytest.approx(want, rel=2e-6), h
    below = f[gaps < h0, 2]
    assert np.all(below == below[0])


# ----------------------------------------------------------- bond break -----
def test_bond_breaks_at_first_force_above_strength():
    """alt_breaking_sinter_bridge.h:115-119: the ordered pair stays bonded
    while |f_n n + f_t| <= strength; the first evaluation above it is still
    the bonded one; from the next step on the contact functor evaluates the
    pair with zero springs."""
    sinter, contact = O.sinter_params(), O.contact_params()
    for seed in range(3):
        st = _seq.bonded_sequence(250 + seed)
        bonded, _, b = O.oracle_sinter_seq(sinter, contact, st)
        assert b == 1
        fmag = np.linalg.norm(bonded[:, 0:3] * M, axis=1)
        strength = 0.9 * fmag.max()
        first = int(np.argmax(fmag > strength))
        assert 0 < first < len(st) - 1
        out, broken = O.oracle_breaking_seq(sinter, contact, strength, st)
        assert np.all(broken[:first] == 0) and np.all(broken[first:] == 1)
        np.testing.assert_array_equal(out[:first + 1], bonded[:first + 1, 0:6])
        after, _ = O.oracle_contact_seq(contact, st[first + 1:])
        np.testing.assert_array_equal(out[first + 1:], after[:, 0:6])
        # never breaking is the plain bonded functor
        o_inf, b_inf = O.oracle_breaking_seq(sinter, contact, float("inf"), st)
        assert b_inf.sum() == 0
        np.testing.assert_array_equal(o_inf, bonded[:, 0:6])


# ---------------------------------------------------------------- fold ------
def test_container_is_the_sum_of_its_functors():
    """binary_force_container.h:20-32 with operator_overloads.h:10-13: the
    container's (a, alpha) of each ordered pair is the bonded/contact
    functor's plus the limited-range Hamaker functor's, bit for bit."""
    sinter, contact, ham = O.sinter_params(), O.contact_params(), O.hamaker_params(h0=1.0e-10)
    for seed, seq in ((0, _seq.bonded_sequence(410)), (1, _seq.pair_sequence(411, amp=0.2))):
        fold = O.oracle_fold_seq(sinter, contact, ham, seq)
        b, _, _ = O.oracle_sinter_seq(sinter, contact, seq)
        h = O.oracle_hamaker_seq(O.HAM_LIMITED, ham, seq)
        np.testing.assert_array_equal(fold, b + h, err_msg=f"sequence {seed}")
        assert np.abs(h).max() > 0 and np.abs(b).max() > 0


def test_simplified_hamaker_closed_form():
    """simplified_hamaker_force.h:18-30: 2 A r / h^2 along x_j - x_i with
    h = |x_j - x_i| - 2r clamped at h0, no torque; the two ordered pairs are
    equal and opposite."""
    hp = O.hamaker_params(h0=1.0e-10)
    A, h0 = hp[0], hp[1]
    clamped = 0
    for seed in range(2):
        s = _seq.far_sequence(500 + seed)
        out = O.oracle_hamaker_seq(O.HAM_SIMPLIFIED, hp, s)
        rv = s[:, 3:6] - s[:, 0:3]
        d = np.linalg.norm(rv, axis=1)
        h = np.maximum(d - 2.0 * R, h0)
        clamped += int(np.sum(d - 2.0 * R < h0))
        want = (2.0 * A * R / h ** 2 / M)[:, None] * rv / d[:, None]
        np.testing.assert_allclose(out[:, 0:3], want, rtol=1e-13, atol=0.0)
        assert np.all(out[:, 3:6] == 0.0) and np.all(out[:, 9:12] == 0.0)
        np.testing.assert_array_equal(out[:, 6:9], -out[:, 0:3])
    assert clamped > 0


# ------------------------------------------------------------ topology ------
def kruskal_bonds(x, r, d_crit, prevent_cycles=True):
    """Independent bond search: pairs i < j in lexicographic order with
    |(|x_i - x_j| - 2r)| < d_crit, a union-find forest rejecting the pairs
    whose ends are already connected."""
    parent = list(range(len(x)))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    bonds, skipped = [], 0
    for i in range(len(x)):
        for j in range(i + 1, len(x)):
            if abs(np.sqrt(np.sum((x[i] - x[j]) ** 2)) - 2 * r) < d_crit:
                ri, rj = find(i), find(j)
                if prevent_cycles and ri == rj:
                    skipped += 1
                    continue
                bonds.append((i, j))
                parent[rj] = ri
    return np.array(bonds, dtype=np.uint32).reshape(-1, 2), skipped


def test_topology_matches_union_find_on_reference_geometries():
    """alt_sinter_bridge.h:56-90 on the reference's own aggregate
    (resources/aggregate.h: 150 touching spheres, a tree, so 149 bonds) and
    on the 3x3x3 lattice of test/alt_sintered.cpp (54 touching pairs: the
    forest keeps 26 and skips 28)."""
    x150 = FIXTURE["x_in_r"] * R
    pairs, skipped = O.oracle_topology(x150, R, 1e-3 * R)
    want, wsk = kruskal_bonds(x150, R, 1e-3 * R)
    np.testing.assert_array_equal(pairs, want)
    assert len(pairs) == 149 and skipped == wsk == 0
    xl, _ = S.sinter_lattice()
    pairs, skipped = O.oracle_topology(xl, R, 1.0e-9)
    want, wsk = kruskal_bonds(xl, R, 1.0e-9)
    np.testing.assert_array_equal(pairs, want)
    assert len(pairs) == 26 and skipped == wsk == 28


def test_topology_catches_a_planted_cycle_rule_mistake():
    """The union-find scan without the cycle check bonds every touching pair
    of the lattice: the oracle's forest differs from it, so the comparison
    above would catch a dropped cycle check."""
    xl, _ = S.sinter_lattice()
    pairs, _ = O.oracle_topology(xl, R, 1.0e-9)
    every, _ = kruskal_bonds(xl, R, 1.0e-9, prevent_cycles=False)
    assert len(every) == 54 and len(pairs) != len(every)


# ----------------------------------------------- legacy sinter_functor ------
def legacy_pair_after(dx_j, kd=1.0e4):
    """Two spheres bonded touching along x (frames from the unit vector y),
    j shifted by dx_j through the step handler, no velocities: the (0, 1)
    and (1, 0) evaluations of the legacy sinter_functor."""
    x0 = np.array([[0.0, 0.0, 0.0], [2 * R, 0.0, 0.0]])
    s13 = O.sinter_params(k_bond=kd, d_crit=1.0e-9)
    s13[1] = 0.0  # no connector damping
    incs = np.zeros((1, 2, 12))
    incs[0, 1, 0:3] = dx_j
    out, used = O.oracle_legacy_seq(2, x0, np.zeros((2, 3)), s13, O.contact_params(), np.array([[0.0, 1.0, 0.0]]),
                                    incs)
    assert used == 1
    return out[0, 0, 1], out[0, 1, 0]


def test_legacy_bond_axial_stretch_closed_form():
    """sinter_bridge.h:235-259: a stretch delta along the bond lengthens the
    four side springs (rest length 2r) and the contact-point spring (rest
    length 0) by delta each: force 5 k_d delta towards the partner, no
    torque, equal and opposite on the partner."""
    kd, delta = 1.0e4, 1.0e-11
    e01, e10 = legacy_pair_after([delta, 0.0, 0.0], kd)
    np.testing.assert_allclose(e01[0:3] * M, [5 * kd * delta, 0.0, 0.0], rtol=1e-9, atol=1e-12 * kd * delta)
    np.testing.assert_allclose(e01[3:6] * I, 0.0, atol=1e-12 * kd * delta * R)
    np.testing.assert_allclose(e10[0:3], -e01[0:3], rtol=1e-9)


def test_legacy_bond_transverse_shift_closed_form():
    """A shift u normal to the bond: the contact-point spring becomes u
    (force k_d u, torque r k_d u about the particle's centre), and the four
    side springs, now sqrt(4r^2 + u^2) long, pull along the centre line with
    4 k_d (sqrt(4r^2 + u^2) - 2r); their torques cancel in pairs."""
    kd, u = 1.0e4, 2.0e-11
    e01, _ = legacy_pair_after([0.0, u, 0.0], kd)
    d = np.array([2 * R, u, 0.0])
    side = 4 * kd * (np.linalg.norm(d) - 2 * R) * d / np.linalg.norm(d)
    want = side + np.array([0.0, kd * u, 0.0])
    np.testing.assert_allclose(e01[0:3] * M, want, rtol=1e-6, atol=1e-9 * kd * u)
    np.testing.assert_allclose(e01[3:6] * I, [0.0, 0.0, R * kd * u], rtol=1e-9, atol=1e-12 * kd * u * R)


def test_breaking_topology_bonds_every_touching_pair():
    """alt_breaking_sinter_bridge.h:63-75 bonds every pair within d_crit (no
    cycle check): the 3x3x3 lattice gets all 54 touching pairs."""
    x, v = S.sinter_lattice()
    sp = O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14)
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_NONE, sinter=sp, sinter_breaking=1), x, v)
    pairs, skipped = o.bonds()
    assert len(pairs) == 54 and skipped == 0
    o2 = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_NONE, sinter=sp), x, v)
    assert len(o2.bonds()[0]) == 26


def test_legacy_frames_geometry():
    """The connector frame of each bond (sinter_bridge.h:163-206): ds and db
    unit, normal to the bond and to each other; the side connectors at
    x +- r ds and x +- r db on both particles, so each side spring is the
    centre-to-centre vector; the fifth at the contact point of each sphere."""
    r = O.R_PART
    x0 = np.array([[0.0, 0.0, 0.0], [2.0 * r, 0.0, 0.0], [2.0 * r, 2.0 * r + 5e-10, 0.0],
                   [-(2.0 * r - 3e-9) * 0.6, (2.0 * r - 3e-9) * 0.8, 0.0]])
    vecs = np.array([[1.0, 0.0, 0.0], [0.0, 0.6, 0.8], [0.6, 0.0, -0.8], [0.0, 0.0, 1.0]])
    pairs, conn, used = O.legacy_frames(x0, r, 1.0e-9, vecs)
    assert pairs.tolist() == [[0, 1], [1, 2]] and used == 3
    for (i, j), c in zip(pairs, conn):
        d = x0[j] - x0[i]
        n = d / np.linalg.norm(d)
        ds, db = (c[0, 0] - x0[i]) / r, (c[0, 2] - x0[i]) / r
        for u in (ds, db):
            assert abs(np.linalg.norm(u) - 1.0) < 1e-9 and abs(u @ n) < 1e-9
        assert abs(ds @ db) < 1e-9
        np.testing.assert_allclose(c[0, 1] - x0[i], -ds * r, atol=1e-22)
        np.testing.assert_allclose(c[0, 3] - x0[i], -db * r, atol=1e-22)
        for q in range(4):
            np.testing.assert_allclose(c[1, q] - c[0, q], d, atol=1e-22)
        np.testing.assert_allclose(c[0, 4], x0[i] + r * n, atol=1e-22)
        np.testing.assert_allclose(c[1, 4], x0[j] - r * n, atol=1e-22)
    with pytest.raises(ValueError):
        O.legacy_frames(x0, r, 1.0e-9, vecs[:2])


def test_legacy_rigid_motion_leaves_bonds_unloaded():
    """Two touching bonded spheres turned rigidly about their centre, with
    consistent increments (dx = Omega x (x - c) dt, dtheta = Omega dt) and
    velocities, no bond damping (the functor damps the connectors' raw
    relative velocity, Omega x d in a rigid turn): a rigid motion stretches
    no spring, so the bond force stays at rounding level against its scale
    k_d r. Connectors not turned with the particle (the step handler's
    rotation left out, a planted mistake) load the bond at once."""
    r = O.R_PART
    x0 = np.array([[-r, 0.0, 0.0], [r, 0.0, 0.0]])
    s13 = O.sinter_params(k_bond=1000.0, d_crit=1.0e-9)
    s13[1] = 0.0
    omega = np.array([0.0, 0.0, 2e8])
    dt, steps = 1e-13, 200
    incs = np.zeros((steps, 2, 12))
    x = x0.copy()
    for k in range(steps):
        xn = x.copy()
        for p in range(2):
            ang = np.linalg.norm(omega) * dt
            c, s = np.cos(ang), np.sin(ang)
            xn[p] = [c * x[p, 0] - s * x[p, 1], s * x[p, 0] + c * x[p, 1], 0.0]
            incs[k, p, 0:3] = xn[p] - x[p]
            incs[k, p, 3:6] = omega * dt
            incs[k, p, 6:9] = np.cross(omega, xn[p])
            incs[k, p, 9:12] = omega
        x = xn
    vecs = np.array([[0.0, 1.0, 0.0]])
    out, used = O.oracle_legacy_seq(2, x0, np.zeros((2, 3)), s13, O.contact_params(), vecs, incs)
    scale = 1000.0 * r / O.MASS
    assert np.abs(out[:, 0, 1, 0:3]).max() < 1e-6 * scale
    wrong = incs.copy()
    wrong[:, :, 3:6] = 0.0  # dtheta withheld from the step handler: connectors only translate
    bad, _ = O.oracle_legacy_seq(2, x0, np.zeros((2, 3)), s13, O.contact_params(), vecs, wrong)
    assert np.abs(bad[-1, 0, 1, 0:3]).max() > 1e-3 * scale  # 4 mrad of lag: 0.8 % of the scale
