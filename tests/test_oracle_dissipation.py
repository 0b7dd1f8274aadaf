"""The dissipated-work ledger (the reference reports kinetic energy only,
test/compute_energy.cpp:7-58; the term is defined in this library) checked on
the oracle: with damping and Coulomb slip, the ledger's total energy plus the
cumulative dissipated work stays constant up to the integrator's first-order
error, while a large part of the energy is lost. A dissipation term off by a
factor or left out leaves a balance that does not shrink with dt. Oracle only
(CPU); tests/test_gpu_energy.py holds the device's term to the oracle's."""
import numpy as np

import _oracle as O

R = O.R_PART
M = O.MASS
I = O.INERTIA


def contact(gamma_n, mu, dt, k=1.0e4):
    """contact_force_functor parameters (contact_force.h:10-30), every shear
    term damped at 0.2 of gamma_n."""
    g = 0.2 * gamma_n
    return np.array([k, gamma_n, k, g, mu, 0.8, k, g, mu, 0.8, k, g, mu, 0.8, R, M, I, dt])


def sinter(gamma, dt, k=1000.0, d_crit=1.0e-9):
    g = 0.2 * gamma
    return np.array([k, gamma, k, g, k, g, k, g, R, M, I, dt, d_crit])


def balance(o, dt, nsteps, samples=200):
    """Samples of (ledger total, dissipated work by term and total)."""
    every = max(1, nsteps // samples)
    rows = []
    for k in range(nsteps // every + 1):
        if k:
            o.step(dt, every)
        e = o.energy()[7]
        w = o.dissipated_work()
        rows.append([e, w["normal"], w["tangential"], w["rolling"], w["torsion"], w["total"]])
    return np.array(rows)


def glancing(dt, gamma_n=2.0e-9, mu=0.3):
    """Two spheres meeting off-centre with spins that make their contact
    points slide: normal damping, tangential slip and the rolling and torsion
    springs all take energy."""
    x = np.array([[0.0, 0.0, 0.0], [np.sqrt((2.0 * R + 1e-11) ** 2 - (0.4 * R) ** 2), 0.4 * R, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    w = np.array([[0.0, 0.0, 3e7], [2e7, 0.0, -1e7]])
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_NONE, contact=contact(gamma_n, mu, dt)), x, v, w0=w)
    o.track_dissipation()
    return o


def drift(rows, weights=(1.0, 1.0, 1.0, 1.0)):
    """Largest change of total energy plus the weighted dissipated work,
    relative to the initial energy."""
    e, wt = rows[:, 0], rows[:, 1:5] @ np.array(weights)
    s = e + wt
    return np.max(np.abs(s - s[0])) / e[0]


def test_damped_glancing_contact_balances_first_order():
    T, runs = 6e-11, []
    for dt in (1e-14, 5e-15):
        runs.append(balance(glancing(dt), dt, int(round(T / dt))))
    a, b = runs
    lost = a[-1, 5] / a[0, 0]
    d = [drift(r) for r in runs]
    print(f"glancing contact: {lost:.3f} of the energy dissipated; balance drift {d[0]:.3e} at dt, {d[1]:.3e} "
          f"at dt/2; terms {a[-1, 1:5] / a[0, 0]}")
    assert lost > 0.2
    for c in range(1, 5):  # every term took part
        assert abs(a[-1, c]) > 1e-4 * a[0, 0]
    assert d[0] < 0.01 * lost
    assert 1.6 < d[0] / d[1] < 2.5
    # planted mistakes: a term doubled or left out does not balance
    for c in range(4):
        for wgt in (0.0, 2.0):
            wts = [1.0] * 4
            wts[c] = wgt
            dp = [drift(r, wts) for r in runs]
            if abs(a[-1, c + 1]) > 0.02 * a[0, 0]:
                assert dp[0] > 3.0 * d[0], (c, wgt)


def test_slip_and_separation_lose_the_spring_energy():
    """Without damping, a sliding contact still loses energy (Coulomb slip
    resets the spring) and a separating one drops its stretched springs; the
    balance holds through both."""
    dt = 1e-14
    x = np.array([[0.0, 0.0, 0.0], [np.sqrt((2.0 * R + 1e-11) ** 2 - (0.6 * R) ** 2), 0.6 * R, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_NONE, contact=contact(0.0, 0.1, dt)), x, v)
    o.track_dissipation()
    rows = balance(o, dt, 8000)
    w = O.OracleSystem.dissipated_work(o)
    print(f"slip: dissipated {w['total'] / rows[0, 0]:.3e} of the energy, normal {w['normal']:.1e}, balance "
          f"drift {drift(rows):.3e}")
    assert w["normal"] == 0.0  # no normal damping
    assert w["tangential"] > 1e-3 * rows[0, 0]
    assert drift(rows) < 0.05 * w["total"] / rows[0, 0]
    assert o.energy()[2] == 0.0  # separated


def test_damped_bond_balances_first_order():
    """A bonded pair (alt_sinter, damped) stretching, shearing and spinning:
    the bond's damping takes most of the energy, and the balance drift halves
    with dt."""
    v = np.array([[0.05, 0.02, 0.0], [-0.05, -0.02, 0.01]])
    w = np.array([[0.0, 0.0, 2e6], [1e6, 0.0, 0.0]])
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R, 0.0, 0.0]])
    T, runs = 1e-10, []
    for dt in (2.5e-15, 1.25e-15):
        g = 0.1 * np.sqrt(2.0 * M * 1000.0)
        o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_NONE, sinter=sinter(g, dt),
                                           contact=contact(0.0, 1.0, dt)), x, v, w0=w)
        assert len(o.bonds()[0]) == 1
        o.track_dissipation()
        runs.append(balance(o, dt, int(round(T / dt))))
    a = runs[0]
    lost = a[-1, 5] / a[0, 0]
    d = [drift(r) for r in runs]
    print(f"damped bond: {lost:.3f} dissipated; balance drift {d[0]:.3e} at dt, {d[1]:.3e} at dt/2")
    assert lost > 0.1
    assert d[0] < 0.05 * lost
    assert 1.6 < d[0] / d[1] < 2.5
    dp = [drift(r, (2.0, 1.0, 1.0, 1.0)) for r in runs]
    assert dp[0] > 3.0 * d[0]


def test_tracking_leaves_the_trajectory_unchanged():
    dt = 1e-14
    a, b = glancing(dt), glancing(dt)
    b.track_dissipation(False)
    a.step(dt, 3000)
    b.step(dt, 3000)
    for k in ("x", "v", "theta", "omega"):
        np.testing.assert_array_equal(a.state()[k], b.state()[k])
