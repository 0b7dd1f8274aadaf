"""Throughput benchmark of the DEM per-timestep path on one MI355X.

Workload (default): one N=20,000 soot aggregate (TD-CCA, seed 1, Df 1.8,
kf 1.3, r = 14 nm, rho = 1700 kg/m^3), alt_sinter bonds (gap < 1e-3 r, k = 1e4
N/m, damping ratio 0.5) embedding Luding contact (mu = 1), limited-range
Hamaker (A = 1e-20 J, cutoff 3r, h0 = 0.1 nm), neighbour list r_list = 3.2r
rebuilt every 50 steps, dt = 1e-13 s, FP64, Gaussian initial velocities
(sigma 1 m/s). A step is one velocity-Verlet step of the whole aggregate.

Prints one JSON line: particle-steps/s, the roofline of the dominant kernel
(the acceleration pass) from HIP events on its own stream over the timed
region, and the CPU oracle timed on this host's cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 5000] [--warmup 200]
"""
import argparse
import json
import math
import os
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
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6


def aggregate_params(k=1.0e4, zeta=0.5, mu=1.0, A=1.0e-20, h0=1.0e-10, cutoff=3.0 * R_PART, d_crit=1e-3 * R_PART):
    g = 2.0 * zeta * math.sqrt(k * MASS / 2.0)
    contact = [k, g, k, g, mu, 1.0, k, g, mu, 1.0, k, g, mu, 1.0, R_PART, MASS, INERTIA, DT]
    sinter = [k, g, k, g, k, g, k, g, R_PART, MASS, INERTIA, DT, d_crit]
    ham = [A, h0, cutoff, R_PART, MASS]
    return contact, sinter, ham


def workload(n=20000, seed=1, df=1.8, kf=1.3):
    import libgran_amd as L
    x = L.generate_aggregate(n, R_PART, df, kf, seed)
    v0 = np.random.default_rng(seed).normal(0.0, 1.0, x.shape)
    return x, v0


def force_bytes_per_particle(kbar, cbar):
    """Algorithmic HBM bytes of one acceleration pass per particle (FP64):
    own x, v, omega (72) + a, alpha written (48) + neighbour count (4) and ids
    (4 k) + pair-state count read/write (8) + pair-state entries read and
    written (c * 2 * (4 + 72)). Partner gathers are not counted: every
    particle's state is counted once, further reads are cache-served."""
    return 72 + 48 + 4 + 4 * kbar + 8 + cbar * 2 * (4 + 72)


def cpu_baseline(x, v0, steps):
    """The CPU oracle (a restatement of the reference, OpenMP over particles)
    on the same workload, timed on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    contact, sinter, ham = aggregate_params()
    p = O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * R_PART, contact=contact,
                        sinter=sinter, hamaker=ham)
    osys = O.OracleSystem(p, x, v0)
    osys.step(DT, 20, rebuild_every=50)  # warm-up
    t0 = time.perf_counter()
    osys.step(DT, steps, rebuild_every=50)
    el = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": x.shape[0] * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"same N={x.shape[0]} aggregate and parameters, {steps} steps after 20 warm-up steps, "
                      f"list rebuilt every 50 steps (cell list), OpenMP over particles, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--cpu-steps", type=int, default=10000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="time without per-launch HIP events")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (one aggregate per step "
              "is a single-GPU workload; see the sharding notes of the design document)", file=sys.stderr)
        sys.exit(2)

    import torch
    import libgran_amd as L

    x, v0 = workload(args.n)
    contact, sinter, ham = aggregate_params()
    params = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST,
                           r_verlet=3.2 * R_PART, contact=contact, sinter=sinter, hamaker=ham)
    sysm = L.System(params, x, v0)
    sysm.step(DT, args.warmup, rebuild_every=50)
    torch.cuda.synchronize()

    if not args.no_events:
        L._abi.check(L._abi.lib().hg_profile_enable(sysm._h, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sysm.step(DT, args.steps, rebuild_every=50)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.n * args.steps / elapsed

    launches = np.zeros(4, dtype=np.uint64)
    dev_ms = np.zeros(4)
    L._abi.check(L._abi.lib().hg_profile_read(sysm._h, launches.ctypes.data_as(L._abi._u64p),
                                              dev_ms.ctypes.data_as(L._abi._dp)))
    off, idx = sysm.neighbors()
    kbar = len(idx) / args.n
    cbar = (len(sysm.springs(0)[0]) + len(sysm.springs(1)[0])) / args.n
    roofline = None
    if launches[0] > 0:
        avg_ms = dev_ms[0] / launches[0]
        bytes_launch = force_bytes_per_particle(kbar, cbar) * args.n
        achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                    "kernel": "k_force_nl<SINTER,LIMITED> (acceleration pass)",
                    "avg_launch_us": round(avg_ms * 1e3, 3), "launches": int(launches[0]),
                    "alg_bytes_per_launch": int(bytes_launch), "kbar": round(kbar, 3), "cbar": round(cbar, 3),
                    "phase_us_per_step": {k: round(float(dev_ms[i]) * 1e3 / args.steps, 3)
                                          for i, k in enumerate(("force", "kick", "rebuild"))}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, v0, args.cpu_steps)

    out = {"metric": "particle-steps/s", "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3 / args.steps, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"single aggregate N={args.n} (seed 1, Df 1.8, kf 1.3), alt_sinter + contact + "
                                  "limited-range Hamaker (3r), neighbour list 3.2r rebuilt every 50 steps, "
                                  "dt 1e-13 s, FP64",
                      "particles": args.n, "rebuild_every": 50, "precision": "fp64"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
