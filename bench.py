"""Throughput benchmark of the DEM per-timestep path on one MI355X.

Headline workload (`value`): the 4,096-aggregate batch. Particle counts
N_k = clip(round(exp(ln 400 + z_k)), 50, 20000) with z_k ~ N(0, 1), fractal
dimensions Df_k ~ U(1.6, 2.2) (numpy default_rng(2024)), kf = 1.3, TD-CCA
aggregates with seeds 0..4095; r = 14 nm, rho = 1700 kg/m^3; alt_sinter bonds
(gap < 1e-3 r, k = 1e4 N/m, damping ratio 0.5) embedding Luding contact
(mu = 1), limited-range Hamaker (A = 1e-20 J, cutoff 3r, h0 = 0.1 nm),
neighbour lists of radius 3.2r rebuilt every 50 steps, dt = 1e-13 s, Gaussian
initial velocities (sigma 1 m/s). All aggregates are stepped together in one
pool, mixed precision (FP32 forces, FP64 state and springs), 2,000 steps. The
pool's launch split (which aggregates take the wide per-step launches) is
tuned online during the warm-up steps.

Beside it on the same line: the same batch in an FP64 pool, the plain FP64
GPU path (one system at a time, per-step launches) on a stratified sample of
the batch, the single N=20,000 aggregate (FP64 and mixed), the energy ledger's
drift on that aggregate with the conservative part of the model (FP64 and
mixed), the serving stream
(Poisson arrivals of plate compression/tension tests through the
continuous-batching server, p50/p99 completion latency), the roofline of the
dominant kernel from HIP events on its own stream, and the CPU oracle on this
host's cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 1500]
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R_PART = 1.4e-8
RHO = 1700.0
MASS = 4.0 / 3.0 * math.pi * R_PART ** 3 * RHO
INERTIA = 2.0 / 5.0 * MASS * R_PART ** 2
DT = 1e-13
REBUILD = 50
HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def log(msg):
    """Progress of the legs, on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)
# The pool's autotuner scores each candidate over two rebuild intervals: two
# chunk shapes, then 21 launch splits on this batch. With fewer warm-up steps
# a fixed choice is set: 512 threads x 1 workgroup per CU (what the tuner
# picks) and wide threshold 4096 with the wide steps run before the chunk
# kernel rather than beside it (the fastest of a sweep of 14 splits, and the
# best chunk-kernel fraction among the fast ones), at which the committed
# counter profiles were taken.
TUNER_STEPS = 46 * REBUILD
PROFILED_SPLIT = (4096, 0)
PROFILED_SHAPE = (512, 1)


def aggregate_params(k=1.0e4, zeta=0.5, mu=1.0, A=1.0e-20, h0=1.0e-10, cutoff=3.0 * R_PART, d_crit=1e-3 * R_PART):
    g = 2.0 * zeta * math.sqrt(k * MASS / 2.0)
    contact = [k, g, k, g, mu, 1.0, k, g, mu, 1.0, k, g, mu, 1.0, R_PART, MASS, INERTIA, DT]
    sinter = [k, g, k, g, k, g, k, g, R_PART, MASS, INERTIA, DT, d_crit]
    ham = [A, h0, cutoff, R_PART, MASS]
    return contact, sinter, ham


def params(L, precision, facets=False):
    c, s, h = aggregate_params()
    extra = {}
    if facets:
        extra = dict(facet_contact=1, facet_hamaker=1, surface_contact=c,
                     surface_hamaker=[1.0e-20, 1.0e-10, 3.0 * R_PART, R_PART, MASS])
    return L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R_PART,
                         contact=c, sinter=s, hamaker=h, precision=precision, **extra)


def batch_workload(L, count=4096):
    from libgran_amd import workloads as W
    sizes, dfs = W.batch_sizes(count)
    xs = L.generate_aggregates(sizes, dfs, 1.3, np.arange(count, dtype=np.uint64), R_PART)
    vs = [np.random.default_rng(1000 + k).normal(0.0, 1.0, x.shape) for k, x in enumerate(xs)]
    return sizes, xs, vs


def state_bytes_per_particle_step(kbar, cbar):
    """Algorithmic HBM bytes per particle-step (FP64 state, FP64 springs):
    x, v, theta, omega read and written (96 + 96), a, alpha read and written
    (48 + 48), neighbour ids (4 k), bond/contact springs read and written
    (144 c), the list rebuild amortised over its interval ((24 + 4 k)/50)."""
    return 288 + 4 * kbar + 144 * cbar + (24 + 4 * kbar) / REBUILD


def run_pool(L, precision, sizes, xs, vs, steps, warmup, split=None, chunk_shape=None, packing=True):
    import torch
    pool = L.Pool(params(L, precision), max_particles=int(sizes.sum()) + 1024, max_systems=len(xs))
    slots = [pool.add(x, v) for x, v in zip(xs, vs)]
    pool.set_packing(packing)
    if chunk_shape:
        pool.set_chunk_shape(*chunk_shape)
    if split is None:
        pool.autotune(True)  # tunes the launch split on the warm-up steps; results do not depend on it
    else:
        pool.set_wide_threshold(split[0])
        pool.set_wide_cus(split[1])
    pool.step(DT, warmup, rebuild_every=REBUILD)
    pool.autotune(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pool.step(DT, steps, rebuild_every=REBUILD)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    prof = pool.last_step_profile()
    tuning = pool.tuning()
    tuning["packing"] = packing
    # measured neighbour and spring counts on a sample of the chunk kernel's systems
    wm = tuning["wide_min"]
    narrow = [s for s, n in zip(slots, sizes) if not wm or n < wm]
    sample = narrow[:: max(1, len(narrow) // 64)]
    npart = sum(pool.info(s)[0] for s in sample)
    kbar = sum(len(pool.neighbors(s)[1]) for s in sample) / npart
    cbar = sum(len(pool.springs(s, 0)[0]) + len(pool.springs(s, 1)[0]) for s in sample) / npart
    narrow_particles = int(sum(int(n) for n in sizes if not wm or n < wm))
    return pool, el, prof, kbar, cbar, tuning, narrow_particles


def plain_fp64_sample(L, sizes, xs, vs, steps, nsample=48):
    """The plain FP64 path: one hg_system at a time, per-step launches, on a
    stratified sample of the batch (every len/nsample-th aggregate by size)."""
    import torch
    order = np.argsort(sizes)
    pick = order[np.linspace(0, len(order) - 1, nsample).astype(int)]
    ps, tt = 0, 0.0
    p = params(L, 0)
    for k in pick:
        s = L.System(p, xs[k], vs[k])
        s.step(DT, REBUILD, rebuild_every=REBUILD)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(DT, steps, rebuild_every=REBUILD)
        torch.cuda.synchronize()
        tt += time.perf_counter() - t0
        ps += int(sizes[k]) * steps
        s.close()
    return ps / tt, f"{nsample} aggregates sampled across the size distribution, {steps} steps each"


def single_aggregate(L, precision, steps, warmup):
    import torch
    x = L.generate_aggregate(20000, R_PART, 1.8, 1.3, 1)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, x.shape)
    s = L.System(params(L, precision), x, v0)
    s.step(DT, warmup, rebuild_every=REBUILD)
    torch.cuda.synchronize()
    s.profile(True)
    t0 = time.perf_counter()
    s.step(DT, steps, rebuild_every=REBUILD)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    launches, ms = s.profile_read()
    return 20000 * steps / el, x, v0, launches, ms


def energy_drift(L, precision, steps=20000, every=1000):
    """Energy ledger drift on the N=20,000 aggregate with the conservative
    part of the benchmark's model: bonds and contacts with their normal
    springs only (no damping, zero tangential, rolling and torsion stiffness)
    and limited-range Hamaker. The reference advances the shear springs after
    taking their force (an explicit step whose energy error is first order in
    dt and, undamped, grows), so they are left out; what remains changes the
    ledger total only by integrator and precision error. Returns the largest
    |E(t) - E(0)| over samples every `every` steps, relative to the ledger's
    scale (the sum of the magnitudes of its components at t = 0)."""
    k = 1.0e4
    c = [k, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, R_PART, MASS, INERTIA, DT]
    s = [k, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, R_PART, MASS, INERTIA, DT, 1e-3 * R_PART]
    h = aggregate_params()[2]
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R_PART,
                      contact=c, sinter=s, hamaker=h, precision=precision)
    x = L.generate_aggregate(20000, R_PART, 1.8, 1.3, 1)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, x.shape)
    sys_ = L.System(p, x, v0)
    e0 = sys_.energy(MASS, INERTIA)
    scale = sum(abs(e0[t]) for t in L.ENERGY_TERMS[:7])
    drift = 0.0
    for _ in range(steps // every):
        sys_.step(DT, every, rebuild_every=REBUILD)
        d = abs(sys_.energy(MASS, INERTIA)["total"] - e0["total"])
        drift = max(drift, d) if math.isfinite(d) else math.inf
    e1 = sys_.energy(MASS, INERTIA)
    sys_.close()
    return {"steps": steps, "max_rel_drift": float(f"{drift / scale:.3e}"), "scale_J": float(f"{scale:.6e}"),
            "start_J": {t: float(f"{e0[t]:.6e}") for t in L.ENERGY_TERMS},
            "end_J": {t: float(f"{e1[t]:.6e}") for t in L.ENERGY_TERMS}}


def legacy_sinter_leg(L, steps=2000):
    """The legacy sinter_functor (five connector springs per bond, moved by
    the step handler) with limited-range Hamaker on the N=20,000 aggregate,
    FP64 neighbour list 3.2r rebuilt every 50 steps: particle-steps/s. The
    connectors are undamped with k_d = 1e4 N/m, the reference's own legacy
    scenario (test/sintered.cpp:66-67): damping at the batch's ratio on all
    five connectors of every bond drives explicit velocity Verlet unstable on
    particles with several bonds (the state leaves the finite range within
    60 steps, oracle and device alike)."""
    import torch
    c, _, h = aggregate_params()
    kd = 1.0e4
    s = [kd, 0.0, kd, 0.0, kd, 0.0, kd, 0.0, R_PART, MASS, INERTIA, DT, 1e-3 * R_PART]
    x = L.generate_aggregate(20000, R_PART, 1.8, 1.3, 1)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, x.shape)
    vecs = np.random.default_rng(2).uniform(-1.0, 1.0, (200000, 3))
    it = iter(vecs / np.linalg.norm(vecs, axis=1, keepdims=True))
    pairs, conn = L.legacy_sinter_frames(x, R_PART, s[12], lambda: next(it))
    p = L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                      r_verlet=3.2 * R_PART, contact=c, sinter=s, hamaker=h)
    sys_ = L.System(p, x, v0)
    sys_.set_legacy_connectors(pairs, conn)
    sys_.step(DT, 200, rebuild_every=REBUILD)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sys_.step(DT, steps, rebuild_every=REBUILD)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    sys_.close()
    return {"value": round(20000 * steps / el, 1), "unit": "particle-steps/s", "bonds": int(len(pairs)),
            "steps": steps, "connectors": "k_d 1e4 N/m, undamped (test/sintered.cpp:66-67)"}


def ledger_and_writer_leg(L, calls=20):
    """Cost of one energy ledger (hg_energy: kernel, copy back, host sum) on
    the N=20,000 aggregate, and of writing its state as a VTK particle file
    (hg_write_vtk_particles, 20,000 particles)."""
    import tempfile
    x = L.generate_aggregate(20000, R_PART, 1.8, 1.3, 1)
    v0 = np.random.default_rng(1).normal(0.0, 1.0, x.shape)
    sys_ = L.System(params(L, L.PRECISION_FP64), x, v0)
    sys_.step(DT, 100, rebuild_every=REBUILD)
    sys_.energy(MASS, INERTIA)
    t0 = time.perf_counter()
    for _ in range(calls):
        sys_.energy(MASS, INERTIA)
    ledger_ms = (time.perf_counter() - t0) * 1e3 / calls
    st = sys_.state()
    sys_.close()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "particles.vtk")
        L.write_vtk_particles(path, st["x"], st["theta"], R_PART)
        t0 = time.perf_counter()
        for _ in range(5):
            L.write_vtk_particles(path, st["x"], st["theta"], R_PART)
        vtk_s = (time.perf_counter() - t0) / 5
        size = os.path.getsize(path)
    return {"energy_ledger_ms_per_call": round(ledger_ms, 3),
            "vtk_particles_per_s": round(20000 / vtk_s, 1), "vtk_mb_per_s": round(size / vtk_s / 1e6, 1),
            "vtk_threads": min(16, os.cpu_count() or 1)}


def small_systems_leg(L, steps=20000):
    """The reference's own 19-particle scenario (test/hamaker.cpp: 18-particle
    lattice and an impactor, contact + Hamaker, dt 1e-13 s) as a single FP64
    system, all pairs and with the 5r neighbour list rebuilt every 20 steps
    (test/hamaker_neighbor.cpp): microseconds per step with the whole call in
    one launch per rebuild interval, with per-step launches (the path of
    every larger system, forced by hg_set_per_step_launches, no event
    timing), and with one queued step per call as a drop-in driver's do_step
    makes them."""
    import torch
    x, v = [], []
    for i in range(3):
        y = -3.0 * R_PART + i * 2.0 * R_PART
        for j in range(3):
            xx = -3.0 * R_PART + j * 2.0 * R_PART
            x += [[xx, y, 0.0], [xx, y, -2.0 * R_PART]]
    x.append([0.0, 0.0, 3.0 * R_PART])
    x = np.array(x)
    v = np.zeros_like(x)
    v[-1] = [1.0, 0.0, -10.0]
    k, gn = 10000.0, 5.0e-9
    contact = [k, gn, k, 0.2 * gn, 1.0, 1.0, k, 0.05 * gn, 0.1, 1.0, k, 0.05 * gn, 0.1, 1.0, R_PART, MASS, INERTIA, DT]
    ham = [1.0e-20, 1.0e-9, 3.0 * R_PART, R_PART, MASS]
    out = {"workload": "19-particle lattice + impactor, contact + Hamaker, FP64 single system", "steps": steps}
    for mode, kw, every in (("all_pairs", {}, 0),
                            ("neighbor_list", dict(pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=5.0 * R_PART), 20)):
        p = L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=contact, hamaker=ham, **kw)
        for name, per_step in (("per_step_launches_us", True), ("one_launch_us", False)):
            s = L.System(p, x, v)
            s.per_step_launches(per_step)
            s.step(DT, 100, rebuild_every=every)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.step(DT, steps, rebuild_every=every)
            torch.cuda.synchronize()
            out[f"{mode}_{name}"] = round((time.perf_counter() - t0) / steps * 1e6, 2)
            s.close()
    s = L.System(L.make_params(L.BINARY_CONTACT, L.HAMAKER_FULL, contact=contact, hamaker=ham), x, v)
    t0 = time.perf_counter()
    for _ in range(steps // 4):
        s.step_async(DT, 1)
    s.state()
    out["all_pairs_queued_step_per_call_us"] = round((time.perf_counter() - t0) / (steps // 4) * 1e6, 2)
    s.close()
    # test/sintered.cpp: the 3x3x3 touching lattice bonded by the legacy
    # sinter_functor (54 bonds, five connectors each, k_d 1e4, undamped); the
    # x_min column moves at +5 m/s in z, the x_max column at +5 m/s in y
    xl, vl = [], []
    for i in range(3):
        for j in range(3):
            xx, yy = (-1.0 + 2.0 * j) * R_PART, (-1.0 + 2.0 * i) * R_PART
            xl += [[xx, yy, 0.0], [xx, yy, -2.0 * R_PART], [xx, yy, -4.0 * R_PART]]
            vl += [[0.0, 0.0, 5.0] if j == 0 else ([0.0, 5.0, 0.0] if j == 2 else [0.0, 0.0, 0.0])] * 3
    xl, vl = np.array(xl), np.array(vl)
    vecs = np.random.default_rng(3).uniform(-1.0, 1.0, (20000, 3))
    it = iter(vecs / np.linalg.norm(vecs, axis=1, keepdims=True))
    pairs, conn = L.legacy_sinter_frames(xl, R_PART, 1.0e-9, lambda: next(it))
    kd = 1.0e4
    sinter = [kd, 0.0, kd, 0.0, kd, 0.0, kd, 0.0, R_PART, MASS, INERTIA, DT, 1.0e-9]
    p = L.make_params(L.BINARY_SINTER_LEGACY, L.HAMAKER_NONE, contact=contact, sinter=sinter)
    for name, per_step in (("per_step_launches_us", True), ("one_launch_us", False)):
        s = L.System(p, xl, vl)
        s.set_legacy_connectors(pairs, conn)
        s.per_step_launches(per_step)
        s.step(DT, 100)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(DT, steps)
        torch.cuda.synchronize()
        out[f"sintered27_legacy_{name}"] = round((time.perf_counter() - t0) / steps * 1e6, 2)
        s.close()
    return out


def serve_stream(L, p, reqs, steps, budget_gib, quantum):
    """Submits `reqs` at their arrival times to a continuous-batching server
    and waits for every result. Returns latencies and the server's figures."""
    srv = L.Server(p, budget_bytes=int(budget_gib * (1 << 30)), max_requests=len(reqs), max_facets=4,
                   quantum=quantum, rebuild_every=REBUILD, dt=DT)
    ids = []
    t0 = time.perf_counter()
    for r in reqs:
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(r["x"], steps, facets=r["facets"]))
    res = [srv.wait(i) for i in ids]
    stats = srv.stats()
    srv.close()
    times = np.array([q["times"] for q in res])
    lat = times[:, 2] - times[:, 0]
    queue = times[:, 1] - times[:, 0]
    span = times[:, 2].max() - times[:, 0].min()
    # device-busy share of the time in which requests were resident
    active = times[:, 2].max() - times[:, 1].min()
    ps = sum(r["n"] for r in reqs) * steps
    rate = (len(reqs) - 1) / max(reqs[-1]["arrival"], 1e-9)
    mean_n = float(np.mean([r["n"] for r in reqs]))
    return {"requests": len(reqs), "arrival_rate_per_s": round(rate, 2),
            "offered_load": round(rate * mean_n * steps, 1),
            "throughput": round(ps / span, 1), "makespan_s": round(float(span), 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
            "queue_wait_p50_s": round(float(np.percentile(queue, 50)), 3),
            "queue_wait_p99_s": round(float(np.percentile(queue, 99)), 3),
            "device_busy_s": round(stats["device_ms"] / 1e3, 3),
            "device_busy_frac": round(stats["device_ms"] / 1e3 / active, 4), "rounds": int(stats["rounds"])}


def serving_leg(L, count, rate=200.0, steps=20000, quantum=REBUILD, budget_gib=4.0, seed=7):
    """The serving stream: plate tests (N log-uniform 100..20000, Df
    U(1.6, 2.2), plates at U(1, 10) m/s, compression or tension) arriving as
    a Poisson stream, each run for `steps` steps through the
    continuous-batching server with an HBM budget of `budget_gib` GiB for
    resident requests; latency = result ready - arrival. Three runs:
    `count` requests at `rate` per second; the same requests as one static
    batch in a pool (all present from the start: the throughput a server
    cannot beat); and `count // 2` further requests at 70% of the measured
    server throughput, so that the queue stays short."""
    from libgran_amd import workloads as W
    p = params(L, L.PRECISION_MIXED, facets=True)
    reqs = W.serving_stream(count, rate, R_PART, seed=seed)
    mean_n = float(np.mean([r["n"] for r in reqs]))
    loaded = serve_stream(L, p, reqs, steps, budget_gib, quantum)
    # the static batch of the same requests
    total = sum(r["n"] for r in reqs)
    pool = L.Pool(p, max_particles=total + 64, max_systems=len(reqs), max_facets=4)
    for r in reqs:
        pool.set_facets(pool.add(r["x"]), *r["facets"])
    pool.autotune(True)
    t0 = time.perf_counter()
    pool.step(DT, steps, REBUILD)
    static = total * steps / (time.perf_counter() - t0)
    pool.close()
    # below capacity
    low_rate = 0.7 * loaded["throughput"] / (mean_n * steps)
    reqs_low = W.serving_stream(max(count // 2, 2), low_rate, R_PART, seed=seed + 1)
    light = serve_stream(L, p, reqs_low, steps, budget_gib, quantum)
    return {"steps_per_request": steps, "budget_gib": budget_gib, "quantum_steps": quantum,
            "unit": "particle-steps/s", "precision": "mixed (FP32 forces, FP64 state/springs)",
            "at_200_per_s": loaded, "static_batch_same_requests": round(static, 1),
            "served_vs_static_batch": round(loaded["throughput"] / static, 4),
            "below_capacity": light}


COUNTERS = os.path.join(ROOT, "profiles", "pool_chunk_mixed_batch4096_hbm_counters.txt")
VALU_COUNTERS = os.path.join(ROOT, "profiles", "pool_chunk_mixed_batch4096_valu_counters.txt")


def same_shape(txt, tuning):
    """Whether a counter profile was taken at this run's chunk-kernel shape and
    packing (a profile that names no packing was taken without)."""
    m = re.search(r"chunk shape\s*(\d+)\s*x\s*(\d+)", txt)
    pk = re.search(r"packing (on|off)", txt)
    packed = bool(pk) and pk.group(1) == "on"
    return (bool(m) and (int(m.group(1)), int(m.group(2))) == (tuning["chunk_threads"], tuning["chunk_wg_per_cu"])
            and packed == tuning.get("packing", True))


def measured_traffic(tuning, count, steps_per_launch):
    """HBM bytes per chunk launch from the committed FETCH_SIZE/WRITE_SIZE
    profile (copy-calibrated), taken at 50 steps per launch and scaled to this
    run's steps per launch. Valid when the run's wide threshold is the
    profile's: the threshold decides which systems the chunk kernel steps, the
    CU room left for the wide launches only its grid."""
    try:
        txt = open(COUNTERS).read()
    except OSError:
        return None, "no counter profile"
    m = re.search(r"traffic per launch\s*=\s*([\d,]+) B", txt)
    split = re.search(r"launch split fixed at\s*(\d+):(\d+)", txt)
    if not m or not split:
        return None, "counter profile unreadable"
    if count != 4096 or tuning["wide_min"] != int(split.group(1)) or not same_shape(txt, tuning):
        return None, f"counter profile was measured at wide threshold {split.group(1)} and another chunk shape"
    scale = steps_per_launch / REBUILD
    return (int(int(m.group(1).replace(",", "")) * scale),
            f"profiles/pool_chunk_mixed_batch4096_hbm_counters.txt (measured at split {split.group(1)}:"
            f"{split.group(2)}, 50 steps per launch; scaled by {scale:.3g} to this run's steps per launch)")


def measured_valu(tuning, count, launch_us, steps_per_launch):
    """FP32 and FP64 VALU throughput of the chunk kernel: FLOPs per launch from
    the committed SQ_INSTS_VALU_* profile (valid at its wide threshold, taken
    at 50 steps per launch and scaled to this run's) over this run's average
    launch duration, against MI355X's vector peaks."""
    try:
        txt = open(VALU_COUNTERS).read()
    except OSError:
        return None
    split = re.search(r"launch split fixed at\s*(\d+):(\d+)", txt)
    f32 = re.search(r"FP32\s+([\d,]+) flop per launch", txt)
    f64 = re.search(r"FP64\s+([\d,]+) flop per launch", txt)
    if not (split and f32 and f64) or count != 4096 or tuning["wide_min"] != int(split.group(1)):
        return None
    if not same_shape(txt, tuning):
        return None
    out = {"source": "profiles/pool_chunk_mixed_batch4096_valu_counters.txt"}
    for name, m, peak in (("fp32", f32, 157.3), ("fp64", f64, 78.6)):
        tf = int(m.group(1).replace(",", "")) * steps_per_launch / REBUILD / (launch_us * 1e-6) / 1e12
        out[name] = {"achieved": round(tf, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(tf / peak, 4)}
    return out


def host_cpu():
    """Model name from lscpu and the CPUs this process may run on."""
    import subprocess
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        model = m.group(1).strip() if m else model
    except (OSError, subprocess.SubprocessError):
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(x, v0, steps, runs=3):
    """The CPU oracle (a restatement of the reference, OpenMP over particles,
    threads bound close) on the N=20,000 aggregate, timed on this host: one
    warm-up run, then the median of `runs` timed runs of `steps` steps."""
    # the OpenMP runtime reads these when the oracle library is loaded
    os.environ.setdefault("OMP_PROC_BIND", "close")
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    c, s, h = aggregate_params()
    p = O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R_PART, contact=c, sinter=s,
                        hamaker=h)
    times = []
    for r in range(runs + 1):
        osys = O.OracleSystem(p, x, v0)
        osys.step(DT, 20, rebuild_every=REBUILD)
        t0 = time.perf_counter()
        osys.step(DT, steps, rebuild_every=REBUILD)
        if r > 0:
            times.append(time.perf_counter() - t0)
    el = float(np.median(times))
    model, cpus = host_cpu()
    return {"value": round(x.shape[0] * steps / el, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "cpus_available": cpus, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "sample": f"N=20,000 aggregate of the batch's parameters (seed 1, Df 1.8), {steps} FP64 steps after 20 "
                      f"warm-up steps, list rebuilt every 50 steps, OpenMP over particles on {threads} threads; one "
                      f"warm-up run, median of {runs} ({', '.join(f'{t:.2f}' for t in times)} s); the batch is "
                      "independent aggregates, so its CPU rate per particle-step is the same"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1500)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=3000)
    ap.add_argument("--serving-requests", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pack", action="store_true", help="one chunk workgroup per system (no packs of small ones)")
    ap.add_argument("--split", default=None, help="WIDE_MIN:WIDE_CUS fixed launch split instead of the autotuner")
    ap.add_argument("--only-batched", action="store_true", help="skip the comparison legs (profiling runs)")
    ap.add_argument("--chunk-shape", default=None, help="THREADS:WG_PER_CU chunk-kernel shape (default: tuned)")
    ap.add_argument("--only-single", action="store_true", help="only the N=20,000 FP64 per-step path (profiling)")
    ap.add_argument("--only-fp64-pool", action="store_true",
                    help="only the batch in an FP64 pool (profiling the FP64 chunk kernel)")
    ap.add_argument("--only-small", action="store_true",
                    help="only the reference's 19-particle lattice scenarios (test/hamaker.cpp, hamaker_neighbor.cpp) "
                         "as single systems: one launch per interval against per-step launches")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run only a 1 GiB device-to-device copy (known traffic for scaling HBM counters)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_read": 1 << 30, "calib_copy_bytes_written": 1 << 30, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the batch shards by aggregate "
              "across GPUs with no communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    # torch first: it bundles its own HIP runtime library with the same soname
    # as ROCm's, and the first one loaded serves the whole process.
    import torch  # noqa: F401  (torch.cuda.synchronize brackets the timed regions)
    import libgran_amd as L

    if args.only_small:
        print(json.dumps(small_systems_leg(L)))
        return

    if args.only_single:
        # the plain FP64 per-step path on the N=20,000 aggregate alone (profiling runs)
        rate, _, _, launches, ms = single_aggregate(L, L.PRECISION_FP64, args.steps, args.warmup)
        print(json.dumps({"workload": "single aggregate N=20000 (seed 1, Df 1.8), FP64, per-step launches",
                          "value": round(rate, 1), "unit": "particle-steps/s", "steps": args.steps,
                          "phase_launches": [int(q) for q in launches],
                          "phase_avg_us": [round(float(m) * 1e3 / max(int(q), 1), 2) for q, m in zip(launches, ms)],
                          "phases": ["force", "kick", "rebuild", "reserved"]}))
        return

    log(f"generating the {args.count}-aggregate batch")
    t_gen = time.perf_counter()
    sizes, xs, vs = batch_workload(L, args.count)
    t_gen = time.perf_counter() - t_gen
    total = int(sizes.sum())
    log(f"batch: {total} particles in {t_gen:.1f} s; mixed pool, {args.warmup} warm-up + {args.steps} timed steps")

    split = tuple(int(q) for q in args.split.split(":")) if args.split else None
    chunk_shape = tuple(int(q) for q in args.chunk_shape.split(":")) if args.chunk_shape else None
    if split is None and args.warmup < TUNER_STEPS:
        # too few warm-up steps for the autotuner: the split it picks on this
        # batch, at which the committed counter profiles were taken
        split = PROFILED_SPLIT
        chunk_shape = chunk_shape or PROFILED_SHAPE
    if args.only_fp64_pool:
        pool, el, prof, kbar, cbar, tuning, narrow = run_pool(L, L.PRECISION_FP64, sizes, xs, vs, args.steps,
                                                              args.warmup, split, chunk_shape, not args.no_pack)
        pool.close()
        c = max(prof["chunks"], 1)
        us = prof["chunk_ms"] * 1e3 / c
        spl = prof["chunk_steps"] / c
        b = state_bytes_per_particle_step(kbar, cbar) * narrow * spl
        print(json.dumps({"kernel": "k_pool_chunk<SINTER,LIMITED,NEIGHBOR_LIST,FP64>", "value": total * args.steps / el,
                          "unit": "particle-steps/s", "avg_launch_us": round(us, 1), "launches": int(prof["chunks"]),
                          "steps_per_launch": spl, "particles_per_launch": narrow, "alg_bytes_per_launch": int(b),
                          "kbar": round(kbar, 3), "cbar": round(cbar, 3),
                          "frac": round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                          "launch_split": f"{tuning['wide_min']}:{tuning['wide_cus']}",
                          "chunk_shape": f"{tuning['chunk_threads']} x {tuning['chunk_wg_per_cu']}"}))
        return
    pool, el_mx, prof_mx, kbar, cbar, tuning, narrow = run_pool(L, L.PRECISION_MIXED, sizes, xs, vs, args.steps,
                                                                 args.warmup, split, chunk_shape,
                                                                 not args.no_pack)
    pool.close()
    value = total * args.steps / el_mx
    log(f"mixed pool: {value:.4g} particle-steps/s")
    chunks = max(prof_mx["chunks"], 1)
    chunk_us = prof_mx["chunk_ms"] * 1e3 / chunks
    steps_per_launch = prof_mx["chunk_steps"] / chunks
    bpps = state_bytes_per_particle_step(kbar, cbar)
    bytes_launch = bpps * narrow * steps_per_launch
    achieved = bytes_launch / (chunk_us * 1e-6) / 1e9
    traffic, traffic_note = measured_traffic(tuning, args.count, steps_per_launch)
    valu = measured_valu(tuning, args.count, chunk_us, steps_per_launch)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_note,
                "kernel": "k_pool_chunk<SINTER,LIMITED,NEIGHBOR_LIST,MIXED> (every aggregate below the wide threshold, "
                          "the steps of one rebuild interval or of the part of one the timed steps cover, per launch)",
                "avg_launch_us": round(chunk_us, 1), "launches": int(prof_mx["chunks"]),
                "chunk_shape": f"{tuning['chunk_threads']} threads x {tuning['chunk_wg_per_cu']} per CU",
                "packing": "on" if tuning["packing"] else "off",
                "steps_per_launch": round(steps_per_launch, 2),
                "particles_per_launch": narrow,
                "alg_bytes_per_particle_step": round(bpps, 1), "alg_bytes_per_launch": int(bytes_launch),
                "kbar": round(kbar, 3), "cbar": round(cbar, 3),
                "rebuilds_timed": sum(1 for k in range(args.warmup, args.warmup + args.steps) if k % REBUILD == 0),
                "rebuild_ms_timed": round(prof_mx["rebuild_ms"], 3),
                "wide_ms_timed": round(prof_mx["wide_ms"], 3),
                "valu": valu}
    extra = {"aggregates": len(xs), "particles": total, "generation_s": round(t_gen, 1),
             "launch_split": {"wide_min": tuning["wide_min"], "wide_cus": tuning["wide_cus"],
                              "chunk_particles": narrow, "wide_particles": total - narrow}}
    cpu = None
    if not args.only_batched:
        log("FP64 pool")
        pool, el_64, prof64, kbar64, cbar64, tuning64, narrow64 = run_pool(L, L.PRECISION_FP64, sizes, xs, vs,
                                                                           args.steps, args.warmup, split, chunk_shape,
                                                                           not args.no_pack)
        pool.close()
        c64 = max(prof64["chunks"], 1)
        us64 = prof64["chunk_ms"] * 1e3 / c64
        b64 = state_bytes_per_particle_step(kbar64, cbar64) * narrow64 * prof64["chunk_steps"] / c64
        extra["roofline_fp64_chunk"] = {
            "kernel": "k_pool_chunk<SINTER,LIMITED,NEIGHBOR_LIST,FP64>", "avg_launch_us": round(us64, 1),
            "steps_per_launch": round(prof64["chunk_steps"] / c64, 2), "particles_per_launch": narrow64,
            "alg_bytes_per_launch": int(b64), "achieved": round(b64 / (us64 * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(b64 / (us64 * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
        log("plain FP64 path on a sample of the batch")
        plain, plain_note = plain_fp64_sample(L, sizes, xs, vs, 200)
        log("single N=20,000 aggregate, FP64 and mixed")
        single64, x20, v20, launches, ms = single_aggregate(L, L.PRECISION_FP64, 5000, 200)
        single_mx = single_aggregate(L, L.PRECISION_MIXED, 5000, 200)[0]
        extra.update({
            "batched_fp64_pool": round(total * args.steps / el_64, 1),
            "plain_fp64_gpu": round(plain, 1), "plain_fp64_gpu_sample": plain_note,
            "batched_mixed_vs_plain_fp64": round(value / plain, 1),
            "batched_fp64_vs_plain_fp64": round(total * args.steps / el_64 / plain, 1),
            "single_n20000_fp64": round(single64, 1), "single_n20000_mixed": round(single_mx, 1),
            "single_n20000_force_kernel_us": round(float(ms[0]) * 1e3 / max(int(launches[0]), 1), 2),
        })
        log("energy ledger drift, FP64 and mixed")
        extra["energy_drift_undamped_n20000"] = {
            "note": "the aggregate above with normal springs (bonds, contacts) and limited-range Hamaker only, no "
                    "damping; largest ledger total drift over 20,000 steps relative to the ledger scale, dt 1e-13 s",
            "fp64": energy_drift(L, L.PRECISION_FP64), "mixed": energy_drift(L, L.PRECISION_MIXED)}
        log("legacy sinter_functor")
        extra["legacy_sinter_n20000_fp64"] = legacy_sinter_leg(L)
        log("energy ledger and VTK writer cost")
        extra["ledger_and_vtk_n20000"] = ledger_and_writer_leg(L)
        if args.serving_requests > 0:
            log(f"serving stream, {args.serving_requests} requests")
            extra["serving"] = serving_leg(L, args.serving_requests)
        if not args.no_cpu_baseline:
            log("CPU oracle baseline")
            cpu = cpu_baseline(x20, v20, args.cpu_steps)
    log("done")
    out = {"metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el_mx * 1e3 / args.steps, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": f"{len(xs)} independent soot aggregates ({total} particles; N log-normal median "
                                  "400 sigma 1 clipped 50..20000, Df U(1.6,2.2), seeds 0..4095), alt_sinter + "
                                  "contact + limited-range Hamaker, neighbour lists 3.2r rebuilt every 50 steps, "
                                  "dt 1e-13 s, mixed precision (FP32 forces, FP64 state/springs), one pool",
                      "rebuild_every": REBUILD, "precision": "mixed"},
           "roofline": roofline, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
