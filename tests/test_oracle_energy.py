"""The energy ledger (the reference reports kinetic energy only,
test/compute_energy.cpp:7-58; the potential terms are defined in this library) checked
against the forces it must be the potential of: in undamped two-body motion
the ledger's total is conserved up to the velocity-Verlet error, and a
potential with a wrong sign, factor or saturation branch would drift by orders
of magnitude more. Oracle only (CPU); tests/test_gpu_energy.py holds the
device ledger to the oracle's."""
import numpy as np
import pytest

import _oracle as O

R = O.R_PART
M = O.MASS
I = O.INERTIA


def undamped_sinter(k=1000.0, d_crit=1.0e-9, dt=1e-14):
    return np.array([k, 0.0, k, 0.0, k, 0.0, k, 0.0, R, M, I, dt, d_crit])


def undamped_contact(k=1.0e4, mu=1.0e3, dt=1e-14):
    return np.array([k, 0.0, k, 0.0, mu, 1.0, k, 0.0, mu, 1.0, k, 0.0, mu, 1.0, R, M, I, dt])


def run(o, dt, nsteps, every):
    """Ledger components and total every `every` steps, and the smallest
    surface gap of particles 0 and 1 seen at those samples."""
    rows, gap = [o.energy()], np.inf
    for _ in range(nsteps // every):
        o.step(dt, every)
        rows.append(o.energy())
        xs = o.state()["x"]
        gap = min(gap, np.linalg.norm(xs[1] - xs[0]) - 2.0 * R)
    return np.array(rows), gap


@pytest.mark.parametrize("model", [O.HAM_FULL, O.HAM_LIMITED, O.HAM_SIMPLIFIED])
def test_hamaker_pair_conserves_ledger(model):
    """Two spheres attracting each other from a 5 nm gap through the h0
    saturation and on into overlap (no contact functor): kinetic plus Hamaker
    potential stays constant while a large part of it changes hands."""
    ham = O.hamaker_params(A=1e-20, h0=1e-9)
    # offset sideways by one radius, so that the centres never meet (where the
    # force direction would flip through a zero distance, a kink no integrator
    # steps across without error)
    x = np.array([[0.0, 0.0, 0.0], [np.sqrt((2.0 * R + 5e-9) ** 2 - R ** 2), R, 0.0]])
    v = np.array([[0.5, 0.0, 0.0], [-0.5, 0.0, 0.0]])
    o = O.OracleSystem(O.system_params(O.BIN_NONE, model, hamaker=ham), x, v)
    e, gap = run(o, 1e-13, 70000, 350)
    assert gap < 0.0  # went through the saturation region into overlap
    swing = np.ptp(e[:, 0] + e[:, 1])  # kinetic energy that changed hands
    drift = np.max(np.abs(e[:, 7] - e[0, 7]))
    print(f"model {model}: kinetic swing {swing:.3e} J, total drift {drift:.3e} J")
    assert swing > 0.05 * e[0, 0]
    assert drift < 1e-6 * swing


def bonded_pair(dt, v, w):
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R, 0.0, 0.0]])
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_NONE, sinter=undamped_sinter(dt=dt),
                                       contact=undamped_contact(dt=dt)), x, v, w0=w)
    assert len(o.bonds()[0]) == 1
    return o


def relative_drift(e, total):
    return np.max(np.abs(total - total[0])) / total[0]


def test_bonded_axial_mode_conserves_ledger():
    """A bond of two touching spheres (no damping) stretching along its axis:
    kinetic and normal elastic energy exchange, and their sum holds to the
    integrator's error."""
    v = np.array([[0.05, 0.0, 0.0], [-0.05, 0.0, 0.0]])
    e, _ = run(bonded_pair(1e-14, v, np.zeros((2, 3))), 1e-14, 20000, 100)
    assert e[:, 2].max() > 0.3 * e[0, 7]  # the bond took up a large part
    drift = relative_drift(e, e[:, 7])
    print(f"bonded axial: relative drift {drift:.3e}")
    assert drift < 1e-4
    assert relative_drift(e, e[:, 7] + e[:, 2]) > 0.1  # the normal term counted twice would show


def test_bonded_twist_torsion_term_closed_form():
    """A bond twisting about its axis. The functor's spring rotation removes
    the component along the normal (include/libgran/sinter_bridge/
    alt_sinter_bridge.h:144) before it takes the force, and the torsion spring
    lies along the normal, so the elastic torsion torque is zero: the spin
    keeps its kinetic energy exactly, and each ordered entry holds only the
    last step's increment r*w*dt. The torsion term is then k/4 (r w dt)^2 per
    entry, k/2 (r w dt)^2 for the pair."""
    k, dt, wx = 1000.0, 1e-14, 2e6
    w = np.zeros((2, 3))
    w[0, 0] = wx
    e, _ = run(bonded_pair(dt, np.zeros((2, 3)), w), dt, 2000, 100)
    assert np.all(e[:, 1] == e[0, 1])
    np.testing.assert_allclose(e[1:, 5], 0.5 * k * (R * wx * dt) ** 2, rtol=1e-12)


def test_bonded_shear_modes_drift_at_first_order():
    """Shear, rolling and spin together. The bonded functor advances each
    tangential spring after taking its force from the old value
    (include/libgran/sinter_bridge/alt_sinter_bridge.h:140-155), an explicit
    step whose energy error is first order in dt; the overlap term is a
    function of position and second order. So over a fixed time the ledger's
    drift halves with dt: the ledger is the potential of the forces. A ledger
    term off by a factor leaves a drift that does not shrink with dt, and the
    same run with the normal, tangential or rolling term doubled shows it."""
    v = np.array([[0.05, 0.02, 0.0], [-0.05, -0.02, 0.01]])
    w = np.array([[0.0, 0.0, 2e6], [1e6, 0.0, 0.0]])
    T, rows = 2e-10, []
    for dt in (2.5e-15, 1.25e-15):
        n = int(round(T / dt))
        e, _ = run(bonded_pair(dt, v, w), dt, n, n // 200)
        rows.append(e)
    for c in (2, 3, 4, 5):
        assert np.ptp(rows[0][:, c]) > 0.0  # every spring term took part
    d = [relative_drift(e, e[:, 7]) for e in rows]
    print(f"bonded shear: relative drift {d[0]:.3e} at dt, {d[1]:.3e} at dt/2")
    assert d[0] < 0.05
    assert 1.8 < d[0] / d[1] < 2.3
    for c in (2, 3, 4):  # planted mistakes, on the oracle's ledger
        dp = [relative_drift(e, e[:, 7] + e[:, c]) for e in rows]
        assert dp[0] / dp[1] < 1.6


def test_contact_collision_conserves_ledger():
    """Head-on collision of two spheres through the contact functor with no
    damping: kinetic energy turns into k/2 overlap^2 and back."""
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R + 1e-11, 0.0, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_NONE, contact=undamped_contact()), x, v)
    e, gap = run(o, 1e-14, 6000, 20)
    assert e[:, 2].max() > 0.5 * e[0, 0]  # most of the kinetic energy went into the overlap
    drift = np.max(np.abs(e[:, 7] - e[0, 7]))
    print(f"contact: total drift {drift:.3e} J of {e[0, 7]:.3e} J")
    assert drift < 1e-4 * e[0, 7]
    assert e[-1, 2] == 0.0 and abs(e[-1, 0] - e[0, 0]) < 1e-4 * e[0, 0]  # separated again, elastic


def legacy_pair_energy(dt, T, v, w):
    """Two touching spheres bonded by the legacy sinter_functor (undamped
    connector springs, k_d = 1000), the ledger sampled 200 times over T."""
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R, 0.0, 0.0]])
    s13 = undamped_sinter(dt=dt)
    pairs, conn, _ = O.legacy_frames(x, R, 1.0e-9, np.array([[0.0, 0.6, 0.8]]))
    o = O.OracleSystem(O.system_params(O.BIN_SINTER_LEGACY, O.HAM_NONE, contact=undamped_contact(dt=dt),
                                       sinter=s13), x, v, w0=w)
    o.set_legacy(pairs, conn)
    n = int(round(T / dt))
    e, _ = run(o, dt, n, n // 200)
    return e


def test_legacy_bond_ledger_converges_at_second_order():
    """The legacy bond's five connector springs (sinter_bridge.h:244-256) are
    forces of position only, moved rigidly with the particles by the step
    handler, so the ledger's k_d/2 (|s| - L0)^2 per spring is their potential
    and the total drifts only by the velocity-Verlet error: a quarter when dt
    halves, in stretching and in a general motion with spin. The normal term
    counted twice (a planted mistake) leaves a drift that does not shrink."""
    v = np.array([[0.05, 0.02, 0.0], [-0.05, -0.02, 0.01]])
    w = np.array([[0.0, 0.0, 2e6], [1e6, 0.0, 0.0]])
    rows = [legacy_pair_energy(dt, 2e-10, v, w) for dt in (1e-14, 5e-15)]
    assert rows[0][:, 2].max() > 0.5 * rows[0][0, 7]  # the bond took up most of the energy
    d = [relative_drift(e, e[:, 7]) for e in rows]
    print(f"legacy bond: relative drift {d[0]:.3e} at dt, {d[1]:.3e} at dt/2")
    assert d[0] < 2e-5
    assert 3.6 < d[0] / d[1] < 4.4
    dp = [relative_drift(e, e[:, 7] + e[:, 2]) for e in rows]
    assert dp[0] / dp[1] < 1.5
