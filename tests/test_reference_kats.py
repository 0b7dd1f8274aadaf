"""The reference's own known-answer tests, at their full lengths.

Each reference scenario driver (test/contact.cpp, test/hamaker.cpp,
test/hamaker_neighbor.cpp, test/alt_sintered.cpp, test/surface.cpp) runs
t_tot / dt steps and checks kinetic energy, linear momentum, radius of
gyration or centre-of-mass drift against a printed target within 1 %. The
same checks run here at full length on the device (FP64 pool path, part of
the GPU suite) and on the CPU oracle (minutes of host time: opt-in with
HG_SLOW=1). The device also runs the first tenth of every scenario beside the
oracle and must equal it bit for bit there. They pin the system loop and integrator, whose source
(libtimestep) is not part of the reference tree, against numbers the
reference's authors produced with it. Measured on MI355X at full length: every
target met within 1 % except alt_sintered's Rg (below).

One reference target is not reproduced: test/alt_sintered.cpp expects the
bonded 3x3x3 lattice to keep its initial radius of gyration (3.9598e-08 m is
2*sqrt(2)*r to six digits, the closed form of the undeformed lattice rather
than a simulation output; the driver's check uses unqualified abs(), which
under libstdc++ binds to abs(int) and passes whatever the value), while the
restated model lets it unfold by 3.1 %
over the run (4.08443e-08 m; the centre-of-mass drift target is met). The
bond topology is a spanning tree (alt_sinter_bridge.h:64-86 skips the 28
cycle-closing pairs of the 54 touching ones), and the spring rotation
(alt_sinter_bridge.h:141-147) turns the torsion spring into the tangent plane
once the lattice rotates, so the tree twists open under its own rotation. The
growth is the same, to four digits, under kick-drift-kick, textbook velocity
Verlet with forces at the old velocities, kick-drift-kick with forces at
predicted velocities, kick-drift-kick with accelerations recomputed at the
start of every step, symplectic Euler, and with FMA contraction on; so no
choice in the unpinned integrator layer closes the gap. For this quantity the
test asserts the oracle's own full-length value instead (device against
oracle), and records the reference's target and the deviation.
"""
import math
import sys

import numpy as np
import pytest

import _oracle as O
import _scenarios as S

R = O.R_PART
M = O.MASS
I = O.INERTIA


# test/compute_energy.cpp:7-30, test/mass_distribution.cpp:7-27
def kinetic_energy(v, w):
    return float((M * (v * v).sum(axis=1) / 2.0 + I * (w * w).sum(axis=1) / 2.0).sum())


def linear_momentum(v):
    return M * float(np.linalg.norm(v.sum(axis=0)))


def radius_of_gyration(x):
    c = x.mean(axis=0)
    return math.sqrt(float(((x - c) ** 2).sum()) / x.shape[0])


def _within(value, target, pct=1.0):
    return abs(value - target) / target <= pct / 100.0


def _surface_setup():
    x, v, verts = S.surface_drop()  # test/surface.cpp:75-88
    return x, v, verts


# name -> (binary, hamaker, pair_mode, r_verlet, rebuild_every, dt, t_tot, params, setup, checks)
def _cases():
    ham = O.hamaker_params(A=1.0e-20, h0=1.0e-9)
    lattice = S.lattice_with_impactor
    return {
        # test/contact.cpp:31-33, 110-121
        "contact": dict(binary=O.BIN_CONTACT, hamaker=O.HAM_NONE, dt=1e-13, t_tot=3.0e-7, setup=lattice,
                        checks=dict(P=1.96373e-19, KE=3.4925e-19)),
        # test/hamaker.cpp:33-35, 128-144
        "hamaker": dict(binary=O.BIN_CONTACT, hamaker=O.HAM_FULL, ham=ham, dt=1e-13, t_tot=3.0e-7, setup=lattice,
                        checks=dict(P=1.96373e-19, KE=8.33959e-20, Rg=3.56253e-08)),
        # test/hamaker_neighbor.cpp:33-35, 118-146: r_verlet = 5 r, list rebuilt every 20 steps
        "hamaker_neighbor": dict(binary=O.BIN_CONTACT, hamaker=O.HAM_FULL, ham=ham, pair_mode=1, r_verlet=5.0 * R,
                                 rebuild_every=20, dt=1e-13, t_tot=3.0e-7, setup=lattice,
                                 checks=dict(P=1.96373e-19, KE=8.33959e-20, Rg=3.56253e-08)),
        # test/alt_sintered.cpp:38-40, 124-152
        "alt_sintered": dict(binary=O.BIN_SINTER, hamaker=O.HAM_NONE, dt=5e-14, t_tot=1.0e-7,
                             contact=O.contact_params(dt=5e-14),
                             sinter=O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14),
                             setup=S.sinter_lattice, checks=dict(drift=2.35702e-07),
                             # (reference target, oracle value at full length): not reproduced, see above
                             deviations=dict(Rg=(3.9598e-08, 4.084433453446481e-08))),
        # test/surface.cpp:49-51, 60-72, 117-123: surface contact (mu 0.5) and
        # surface Hamaker with A = 0 on one facet
        "surface": dict(binary=O.BIN_CONTACT, hamaker=O.HAM_NONE, dt=1e-13, t_tot=5.0e-7,
                        contact=O.contact_params(mu=0.5), facet_contact=1, facet_hamaker=1,
                        surface_contact=O.contact_params(mu=0.5),
                        surface_hamaker=O.hamaker_params(A=0.0, h0=1.0e-9), setup=_surface_setup,
                        checks=dict(KE=1.0922e-21)),
    }


def _evaluate(case, x0, st):
    out = {}
    for name in list(case["checks"]) + list(case.get("deviations", {})):
        if name == "P":
            out[name] = linear_momentum(st["v"])
        elif name == "KE":
            out[name] = kinetic_energy(st["v"], st["omega"])
        elif name == "Rg":
            out[name] = radius_of_gyration(st["x"])
        elif name == "drift":
            out[name] = float(np.linalg.norm(st["x"].mean(axis=0) - x0.mean(axis=0)))
    return out


def _assert_checks(name, case, got):
    for k, target in case["checks"].items():
        print(f"[kat] {name}: {k} = {got[k]:.6g}, reference target {target:.6g}", flush=True)
        assert _within(got[k], target), f"{name}: {k} = {got[k]:.6g}, reference target {target:.6g} (1 %)"
    for k, (target, oracle_value) in case.get("deviations", {}).items():
        print(f"[kat] {name}: {k} = {got[k]:.6g}, reference target {target:.6g} not reproduced "
              f"({100.0 * (got[k] - target) / target:+.2f} %), oracle {oracle_value:.6g}", flush=True)
        assert got[k] == pytest.approx(oracle_value, rel=1e-9)


def _init(case):
    res = case["setup"]()
    x, v = res[0], res[1]
    facets = (res[2], np.zeros((1, 3))) if len(res) == 3 else None
    return x, v, facets


def _nsteps(case):
    return int(case["t_tot"] / case["dt"])  # size_t(t_tot / dt)


CASES = _cases()


def _device_run(case, progress, n=None, single=False):
    """Runs one scenario on its own pool (own stream), or as a single system,
    for n steps (default: the reference's full length) and returns the final
    state; progress[0] counts the steps done."""
    import libgran_amd as L
    x, v, facets = _init(case)
    n = _nsteps(case) if n is None else n
    p = L.make_params(case["binary"], case["hamaker"], pair_mode=case.get("pair_mode", 0),
                      r_verlet=case.get("r_verlet", 0.0),
                      contact=case.get("contact", O.contact_params()), sinter=case.get("sinter", O.sinter_params()),
                      hamaker=case.get("ham", O.hamaker_params()), facet_contact=case.get("facet_contact", 0),
                      facet_hamaker=case.get("facet_hamaker", 0),
                      surface_contact=case.get("surface_contact", O.contact_params()),
                      surface_hamaker=case.get("surface_hamaker", O.hamaker_params()))
    every = case.get("rebuild_every", 0)
    if single:
        # one launch per slice on the small-system paths
        s = L.System(p, x, v)
        if facets:
            s.set_facets(*facets)
        done = 0
        while done < n:
            k = min(200000, n - done)
            s.step(case["dt"], k, every)
            done += k
            progress[0] = done
        assert s.time()[1] == n
        return x, s.state()
    pool = L.Pool(p, max_particles=max(64, x.shape[0]), max_systems=1, max_facets=1 if facets else 0)
    slot = pool.add(x, v)
    if facets:
        pool.set_facets(slot, *facets)
    slice_len = 100000 - 100000 % every if every else 100000
    done = 0
    while done < n:
        k = min(slice_len, n - done)
        pool.step(case["dt"], k, every)
        done += k
        progress[0] = done
    assert pool.info(slot)[2] == n
    return x, pool.state(slot)


def _run_all(steps_of, oracle_too=False):
    """All five scenarios concurrently (one pool and stream each; a scenario
    of a few dozen particles occupies one workgroup), and optionally the
    oracle for the same steps on host threads. Returns {name: (x0, device
    state, oracle state or None)}."""
    import threading
    import time

    import libgran_amd  # noqa: F401  (load the library before the threads start)
    names = list(CASES)
    progress = {nm: [0] for nm in names}
    results, oracle, errors = {}, {}, {}

    def work(nm):
        try:
            results[nm] = _device_run(CASES[nm], progress[nm], steps_of(nm))
        except Exception as e:  # reported below
            errors[nm] = e

    def work_oracle(nm):
        try:
            oracle[nm] = _oracle_run(CASES[nm], steps_of(nm))
        except Exception as e:  # reported below
            errors[nm + " (oracle)"] = e

    threads = [threading.Thread(target=work, args=(nm,)) for nm in names]
    if oracle_too:
        threads += [threading.Thread(target=work_oracle, args=(nm,)) for nm in names]
    for t in threads:
        t.start()
    t0 = time.time()
    while any(t.is_alive() for t in threads):
        time.sleep(15.0)
        print(f"[kat] {time.time() - t0:6.0f} s: " +
              ", ".join(f"{nm} {progress[nm][0] / steps_of(nm):.0%}" for nm in names), flush=True)
        sys.stdout.flush()
    for t in threads:
        t.join()
    assert not errors, errors
    return {nm: (results[nm][0], results[nm][1], oracle.get(nm)) for nm in names}


def _oracle_run(case, n):
    x, v, facets = _init(case)
    sp = O.system_params(case["binary"], case["hamaker"], neighbor_mode=case.get("pair_mode", 0),
                         r_verlet=case.get("r_verlet", 0.0), contact=case.get("contact"), sinter=case.get("sinter"),
                         hamaker=case.get("ham"), facet_contact=case.get("facet_contact", 0),
                         facet_hamaker=case.get("facet_hamaker", 0), surface_contact=case.get("surface_contact"),
                         surface_hamaker=case.get("surface_hamaker"))
    s = O.OracleSystem(sp, x, v)
    if facets:
        s.set_facets(*facets)
    s.step(case["dt"], n, case.get("rebuild_every", 0))
    return s.state()


@pytest.mark.gpu
def test_reference_scenarios_device_equals_oracle():
    """The five scenarios for the first tenth of their runs (300,000 /
    200,000 / 500,000 steps) on the device and on the oracle: bit for bit."""
    runs = _run_all(lambda nm: _prefix_steps(CASES[nm]), oracle_too=True)
    for nm, (_, dev, ora) in runs.items():
        for k in ("x", "v", "theta", "omega", "a", "alpha"):
            np.testing.assert_array_equal(dev[k], ora[k], err_msg=f"{nm}: field {k}")


def _prefix_steps(case):
    n = _nsteps(case) // 10
    every = case.get("rebuild_every", 0)
    return n - n % every if every else n


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_reference_kats_device(name):
    """One scenario at full length against the reference's targets, as a
    single system (under a minute and a half of device time each)."""
    print(f"[kat] {name}: {_nsteps(CASES[name])} steps", flush=True)
    x0, st = _device_run(CASES[name], [0], single=True)
    _assert_checks(name, CASES[name], _evaluate(CASES[name], x0, st))


@pytest.mark.slow
@pytest.mark.parametrize("name", list(CASES))
def test_reference_kat_oracle(name):
    case = CASES[name]
    x, v, facets = _init(case)
    sp = O.system_params(case["binary"], case["hamaker"], neighbor_mode=case.get("pair_mode", 0),
                         r_verlet=case.get("r_verlet", 0.0), contact=case.get("contact"), sinter=case.get("sinter"),
                         hamaker=case.get("ham"), facet_contact=case.get("facet_contact", 0),
                         facet_hamaker=case.get("facet_hamaker", 0), surface_contact=case.get("surface_contact"),
                         surface_hamaker=case.get("surface_hamaker"))
    s = O.OracleSystem(sp, x, v)
    if facets:
        s.set_facets(*facets)
    s.step(case["dt"], _nsteps(case), case.get("rebuild_every", 0))
    _assert_checks(name, case, _evaluate(case, x, s.state()))
