"""The single-aggregate workloads at their full size on the device.

Single aggregate, FP64 (N=20,000, seed 1, neighbour lists 3.2r rebuilt
every 50 steps): every one of 100 steps gives the oracle's accelerations and
angular accelerations bit for bit, and the lists, bonds and springs agree.

Precision variant (the same aggregate, FP32 force arithmetic): the
trajectory over 5,000 steps against the FP64 oracle, with the deviation
reported and bounded at every 1,000 steps.
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
    print(f"[gpu-configs] {name}", flush=True)


def aggregate_20000():
    L = _L()
    x = L.generate_aggregate(20000, R, 1.8, 1.3, 1)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, x.shape)
    return x, v0


def nl_params(precision):
    L = _L()
    c, s, h = S.aggregate_params()
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                         contact=c, sinter=s, hamaker=h, precision=precision)


def oracle_nl(x, v0):
    c, s, h = S.aggregate_params()
    return O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                          sinter=s, hamaker=h), x, v0)


def test_single_aggregate_fp64_per_step_forces_n20000():
    """100 steps, one per call on both sides: a and alpha after every step
    equal the oracle's bit for bit (including the rebuilds at steps 0 and
    50); at the end the state, neighbour lists, bonds and both spring tables
    agree."""
    announce("FP64 N=20000, 100 steps, per-step accelerations")
    L = _L()
    x, v0 = aggregate_20000()
    o = oracle_nl(x, v0)
    g = L.System(nl_params(0), x, v0)
    for k in range(100):
        o.step(DT, 1, rebuild_every=50)
        g.step(DT, 1, rebuild_every=50)
        a, b = o.state(), g.state()
        np.testing.assert_array_equal(b["a"], a["a"], err_msg=f"step {k}: a")
        np.testing.assert_array_equal(b["alpha"], a["alpha"], err_msg=f"step {k}: alpha")
        assert np.all(np.isfinite(b["a"]))
    for f in ("x", "v", "theta", "omega"):
        np.testing.assert_array_equal(b[f], a[f], err_msg=f)
    off1, idx1 = o.neighbors()
    off2, idx2 = g.neighbors()
    np.testing.assert_array_equal(off2, off1)
    np.testing.assert_array_equal(idx2, idx1)
    np.testing.assert_array_equal(g.bonds()[0], o.bonds()[0])
    for which in (0, 1):
        ij1, s1 = o.springs(which)
        ij2, s2 = g.springs(which)
        np.testing.assert_array_equal(ij2, ij1)
        np.testing.assert_array_equal(s2, s1)
    assert len(o.springs(1)[0]) > 30000  # the bonds carry springs in both directions


def test_mixed_precision_5000_step_trajectory_n20000():
    """The precision variant against the FP64 oracle for 5,000 steps. Stated
    bounds at every 1,000 steps: max |dx| below 1e-14 m (7e-7 of a radius)
    and max |dv| below 1e-5 of the largest speed (measured on MI355X at 5,000
    steps: 1.4e-16 m and 4.4e-7)."""
    announce("mixed N=20000, 5000 steps vs the FP64 oracle")
    L = _L()
    x, v0 = aggregate_20000()
    o = oracle_nl(x, v0)
    g = L.System(nl_params(1), x, v0)
    for k in range(5):
        o.step(DT, 1000, rebuild_every=50)
        g.step(DT, 1000, rebuild_every=50)
        a, b = o.state(), g.state()
        dx = np.abs(b["x"] - a["x"]).max()
        vmax = np.abs(a["v"]).max()
        dv = np.abs(b["v"] - a["v"]).max() / vmax
        dth = np.abs(b["theta"] - a["theta"]).max()
        print(f"[mixed trajectory] step {1000 * (k + 1)}: max |dx| {dx:.3e} m, max |dv|/max|v| {dv:.3e}, "
              f"max |dtheta| {dth:.3e} rad", flush=True)
        assert dx < 1e-14
        assert dv < 1e-5
