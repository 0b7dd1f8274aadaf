"""The continuous-batching server (hg_server) on the device.

Requests are aggregate plate tests (two moving plates of two triangles each).
The arena budget is set so that not every request fits at once: later
requests wait in the queue and are admitted as earlier ones retire. Each
result must equal the same request run alone in a pool of its own, bit for
bit (a system's arithmetic does not depend on the other systems in the pool,
in FP64 and in mixed precision alike), step counts need not be multiples of
the server's quantum, a request whose state overflows fails alone with
HG_ERR_NUMERIC, and an FP64 request equals the oracle.
"""
import sys
import time

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


def _params(L, precision):
    c, s, h = S.aggregate_params()
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                         precision=precision, contact=c, sinter=s, hamaker=h, facet_contact=1, facet_hamaker=1,
                         surface_contact=c, surface_hamaker=O.hamaker_params(A=1e-20, h0=1e-10))


def _requests(L, count, seed):
    from libgran_amd import workloads as W
    reqs = W.serving_stream(count, 200.0, R, seed=seed, n_range=(60, 400), speed_range=(5.0, 10.0))
    for k, r in enumerate(reqs):
        r["v0"] = np.random.default_rng(500 + k).normal(0.0, 1.0, r["x"].shape)
        r["nsteps"] = 430 if k % 2 == 0 else 575  # not multiples of the quantum
    return reqs


def _alone(L, params, r):
    pool = L.Pool(params, max_particles=r["n"] + 64, max_systems=1, max_facets=4)
    slot = pool.add(r["x"], r["v0"])
    pool.set_facets(slot, *r["facets"])
    pool.step(DT, r["nsteps"], 50)
    st = pool.state(slot)
    pool.close()
    return st


@pytest.mark.parametrize("precision", [0, 1])
def test_server_results_equal_alone(precision):
    L = _L()
    print(f"[gpu-server] alone vs served, precision {precision}", flush=True)
    sys.stdout.flush()
    params = _params(L, precision)
    reqs = _requests(L, 6, seed=3 + precision)
    bpp = L.Pool.bytes_per_particle(params, 4)
    # room for the two largest requests at a time: the rest must queue
    budget = bpp * (sorted(r["n"] for r in reqs)[-1] + sorted(r["n"] for r in reqs)[-2] + 8)
    srv = L.Server(params, budget_bytes=budget, max_requests=8, max_facets=4, quantum=100, rebuild_every=50, dt=DT)
    ids = []
    for k, r in enumerate(reqs):
        ids.append(srv.submit(r["x"], r["nsteps"], v0=r["v0"], facets=r["facets"]))
        if k == 1:
            # later submissions arrive while the first requests run
            t0 = time.time()
            while srv.status(ids[0]) == L._abi.HG_REQ_QUEUED and time.time() - t0 < 30:
                time.sleep(0.001)
    results = [srv.wait(i, timeout=120.0) for i in ids]
    stats = srv.stats()
    srv.close()
    assert stats["done"] == len(reqs) and stats["failed"] == 0
    assert stats["particle_steps"] == sum(r["n"] * r["nsteps"] for r in reqs)
    admitted = np.array([res["times"][1] for res in results])
    done = np.array([res["times"][2] for res in results])
    assert np.all(admitted >= np.array([res["times"][0] for res in results]))
    assert np.all(done > admitted)
    assert admitted.max() >= done.min(), "the budget never made a request wait for a retirement"
    for k, (r, res) in enumerate(zip(reqs, results)):
        ref = _alone(L, params, r)
        for f in ("x", "v", "theta", "omega"):
            np.testing.assert_array_equal(res[f], ref[f], err_msg=f"request {k} field {f}")
    # the plates acted: compressed aggregates were pushed inwards at the ends
    assert any(np.abs(res["v"] - r["v0"]).max() > 1.0 for r, res in zip(reqs, results))


def test_server_rejects_bad_requests():
    L = _L()
    params = _params(L, 1)
    srv = L.Server(params, budget_bytes=1 << 26, max_requests=4, max_facets=4, quantum=100, rebuild_every=50, dt=DT)
    x = L.generate_aggregate(50, R, 1.8, 1.3, 1)
    big = L.generate_aggregate(200, R, 1.8, 1.3, 2)
    small = L.Server(params, budget_bytes=L.Pool.bytes_per_particle(params, 4) * 150, max_requests=4, max_facets=4,
                     quantum=100, rebuild_every=50, dt=DT)
    with pytest.raises(L.HipgranError) as e:
        small.submit(big, 100)  # larger than the whole budget
    assert e.value.status == L._abi.HG_ERR_CAPACITY
    # two requests of 100 particles: the second waits for the first
    a = small.submit(x[:50], 1000)
    b = small.submit(L.generate_aggregate(120, R, 1.8, 1.3, 3), 100)
    with pytest.raises(L.HipgranError):
        small.wait(b, timeout=0.0)  # not done yet: time-out, request kept
    ra, rb = small.wait(a, timeout=60.0), small.wait(b, timeout=60.0)
    assert rb["times"][1] >= ra["times"][2]  # admitted only after a retired
    ok = srv.submit(x, 100)
    res = srv.wait(ok, timeout=60.0)
    assert res["times"][2] >= res["times"][1]
    small.close()
    srv.close()


def _announce(name):
    print(f"[gpu-server] {name}", flush=True)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_server_faulty_request_fails_alone(precision):
    """One request carries a particle at 1e300 m/s (its bond damping overflows
    in the first acceleration pass): hg_server_wait reports HG_ERR_NUMERIC for
    it alone; the requests around it finish and equal their runs alone."""
    _announce(f"faulty request, precision {precision}")
    L = _L()
    params = _params(L, precision)
    reqs = _requests(L, 5, seed=11 + precision)
    bad = 2
    reqs[bad]["v0"] = reqs[bad]["v0"].copy()
    reqs[bad]["v0"][3] = [1.0e300, 0.0, 0.0]
    bpp = L.Pool.bytes_per_particle(params, 4)
    srv = L.Server(params, budget_bytes=bpp * (sum(r["n"] for r in reqs) + 64), max_requests=8, max_facets=4,
                   quantum=50, rebuild_every=50, dt=DT)
    ids = [srv.submit(r["x"], r["nsteps"], v0=r["v0"], facets=r["facets"]) for r in reqs]
    results = {}
    for k, i in enumerate(ids):
        if k == bad:
            with pytest.raises(L.HipgranError) as e:
                srv.wait(i, timeout=120.0)
            assert e.value.status == L._abi.HG_ERR_NUMERIC
            assert "NaN or infinite" in str(e.value)
        else:
            results[k] = srv.wait(i, timeout=120.0)
    stats = srv.stats()
    srv.close()
    assert stats["failed"] == 1 and stats["done"] == len(reqs) - 1
    for k, res in results.items():
        ref = _alone(L, params, reqs[k])
        for f in ("x", "v", "theta", "omega"):
            np.testing.assert_array_equal(res[f], ref[f], err_msg=f"request {k} field {f}")


def test_server_serving_stream_scale():
    """Eight plate requests of the serving workload at full length: N
    log-uniform over 100..20,000 (the largest above 10,000), Df 1.6..2.2,
    plates at 1..10 m/s, 20,000 steps each, mixed precision, submitted at
    once into a budget that holds about half of them. Every result equals
    the request run alone, bit for bit; the server's device-busy time covers
    most of its wall time."""
    _announce("serving stream scale: 8 requests, 20,000 steps")
    L = _L()
    from libgran_amd import workloads as W
    params = _params(L, 1)
    reqs = W.serving_stream(8, 200.0, R, seed=5, n_range=(100, 20000))
    big = int(np.argmax([r["n"] for r in reqs]))
    if reqs[big]["n"] < 10000:  # the stream draws log-uniform sizes: make sure one is large
        x = L.generate_aggregate(12000, R, reqs[big]["df"], 1.3, 4242)
        reqs[big].update(x=x, n=12000, facets=W.plates(x, R, reqs[big]["speed"], reqs[big]["compress"]))
    for k, r in enumerate(reqs):
        r["v0"] = np.random.default_rng(900 + k).normal(0.0, 1.0, r["x"].shape)
        r["nsteps"] = 20000
    total = sum(r["n"] for r in reqs)
    bpp = L.Pool.bytes_per_particle(params, 4)
    srv = L.Server(params, budget_bytes=bpp * max(total // 2, max(r["n"] for r in reqs) + 64), max_requests=8,
                   max_facets=4, quantum=50, rebuild_every=50, dt=DT)
    t0 = time.time()
    ids = [srv.submit(r["x"], r["nsteps"], v0=r["v0"], facets=r["facets"]) for r in reqs]
    results = [srv.wait(i, timeout=600.0) for i in ids]
    wall = time.time() - t0
    stats = srv.stats()
    srv.close()
    print(f"[gpu-server] sizes {[r['n'] for r in reqs]}: {wall:.2f} s, device busy {stats['device_ms'] / 1e3:.2f} s,"
          f" {stats['rounds']} rounds", flush=True)
    assert stats["done"] == len(reqs) and stats["failed"] == 0
    assert stats["particle_steps"] == total * 20000
    assert stats["device_ms"] / 1e3 > 0.5 * wall
    for k, (r, res) in enumerate(zip(reqs, results)):
        ref = _alone(L, params, r)
        for f in ("x", "v", "theta", "omega"):
            np.testing.assert_array_equal(res[f], ref[f], err_msg=f"request {k} (N={r['n']}) field {f}")


def test_server_fp64_request_equals_oracle():
    """An FP64 plate request of 300 particles, 1,234 steps through the server
    (beside two other requests): its final state equals the oracle's bit for
    bit."""
    _announce("FP64 request vs the oracle")
    L = _L()
    from libgran_amd import workloads as W
    params = _params(L, 0)
    reqs = _requests(L, 3, seed=21)
    x = L.generate_aggregate(300, R, 1.8, 1.3, 77)
    v0 = np.random.default_rng(77).normal(0.0, 1.0, x.shape)
    verts, vel = W.plates(x, R, 7.0, True)
    srv = L.Server(params, budget_bytes=1 << 28, max_requests=8, max_facets=4, quantum=50, rebuild_every=50, dt=DT)
    others = [srv.submit(r["x"], r["nsteps"], v0=r["v0"], facets=r["facets"]) for r in reqs]
    rid = srv.submit(x, 1234, v0=v0, facets=(verts, vel))
    res = srv.wait(rid, timeout=120.0)
    for i in others:
        srv.wait(i, timeout=120.0)
    srv.close()
    c, s, h = S.aggregate_params()
    o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R, contact=c,
                                       sinter=s, hamaker=h, facet_contact=1, facet_hamaker=1, surface_contact=c,
                                       surface_hamaker=O.hamaker_params(A=1e-20, h0=1e-10)), x, v0)
    o.set_facets(verts, vel)
    o.step(DT, 1234, rebuild_every=50)
    ref = o.state()
    for f in ("x", "v", "theta", "omega"):
        np.testing.assert_array_equal(res[f], ref[f], err_msg=f)
