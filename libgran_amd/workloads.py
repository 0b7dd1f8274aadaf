"""Synthetic inputs of the aggregate workloads: the mixed-size batch and the
serving stream of plate compression/tension tests. Geometry comes from the
library's seeded TD-CCA generator (hg_generate_aggregates), so every input is
reproducible from its seeds."""
import math

import numpy as np

from . import generate_aggregates


def batch_sizes(count=4096, median=400, sigma=1.0, lo=50, hi=20000, seed=2024):
    """Log-normal particle counts and uniform fractal dimensions of the batch
    workload: N_k = clip(round(exp(ln median + sigma z)), lo, hi), Df ~ U(1.6, 2.2)."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(count)
    sizes = np.clip(np.round(np.exp(np.log(float(median)) + sigma * z)), lo, hi).astype(np.uint32)
    dfs = rng.uniform(1.6, 2.2, count)
    return sizes, dfs


def plates(x, r_part, speed, compress=True, margin=2.0):
    """Two rigid square plates (two triangles each) normal to z, touching the
    lowest and highest spheres of x, moving towards each other (compress) or
    apart at `speed` m/s each. Returns verts (4, 3, 3) and velocities (4, 3)."""
    lo = x[:, 2].min() - r_part
    hi = x[:, 2].max() + r_part
    cx, cy = x[:, 0].mean(), x[:, 1].mean()
    a = max(np.abs(x[:, 0] - cx).max(), np.abs(x[:, 1] - cy).max()) + margin * r_part
    x0, x1, y0, y1 = cx - a, cx + a, cy - a, cy + a
    verts = np.array([[[x0, y0, lo], [x1, y0, lo], [x0, y1, lo]],
                      [[x1, y1, lo], [x0, y1, lo], [x1, y0, lo]],
                      [[x0, y0, hi], [x0, y1, hi], [x1, y0, hi]],
                      [[x1, y1, hi], [x1, y0, hi], [x0, y1, hi]]])
    s = speed if compress else -speed
    vel = np.array([[0.0, 0.0, s], [0.0, 0.0, s], [0.0, 0.0, -s], [0.0, 0.0, -s]])
    return verts, vel


def serving_stream(count, rate, r_part, seed=7, n_range=(100, 20000), df_range=(1.6, 2.2),
                   speed_range=(1.0, 10.0), kf=1.3, threads=0):
    """A seeded stream of plate tests: Poisson arrivals at `rate` per second,
    N log-uniform over n_range, Df uniform, plate speed uniform, compression
    or tension with equal odds. Returns a list of dicts with arrival (s), x,
    facets (verts, velocities), n, df, speed, compress."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    arrivals -= arrivals[0]
    ns = np.round(np.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]), count))).astype(np.uint32)
    dfs = rng.uniform(df_range[0], df_range[1], count)
    speeds = rng.uniform(speed_range[0], speed_range[1], count)
    compress = rng.random(count) < 0.5
    seeds = np.arange(100000 + seed * count, 100000 + seed * count + count, dtype=np.uint64)
    xs = generate_aggregates(ns, dfs, kf, seeds, r_part, threads)
    out = []
    for k in range(count):
        out.append(dict(arrival=float(arrivals[k]), x=xs[k], facets=plates(xs[k], r_part, speeds[k], compress[k]),
                        n=int(ns[k]), df=float(dfs[k]), speed=float(speeds[k]), compress=bool(compress[k])))
    return out
