"""The continuous-batching server (hg_server) on the device.

Requests are aggregate plate tests (two moving plates of two triangles each).
The arena budget is set so that not every request fits at once: later
requests wait in the queue and are admitted as earlier ones retire. Each
result must equal the same request run alone in a pool of its own, bit for
bit (a system's arithmetic does not depend on the other systems in the pool,
in FP64 and in mixed precision alike).
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
        r["nsteps"] = 400 if k % 2 == 0 else 600
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
    with pytest.raises(L.HipgranError):
        srv.submit(x, 150)  # not a multiple of the quantum
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
