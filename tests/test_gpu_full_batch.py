"""The benchmark's own workload at its full size: the 4,096-aggregate batch
(about 2.66 M particles; sizes, fractal dimensions, seeds and parameters as
bench.py builds them) in one pool, 100 steps with lists rebuilt every 50.
Size-independent properties checked on systems sampled across the sizes
(every LDS state tier of the chunk kernel and the wide launches):

- a system stepped inside the full batch equals the same system stepped alone
  in a pool of its own, bit for bit, in mixed precision and in FP64, whatever
  else shares the pool;
- in FP64 it equals the CPU oracle, bit for bit."""
import os
import sys

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS, REBUILD, WIDE_MIN = 100, 50, 4096


def announce(name):
    print(f"[gpu-full-batch] {name}", flush=True)


@pytest.fixture(scope="module")
def batch():
    sys.path.insert(0, ROOT)
    import bench as B
    import libgran_amd as L
    announce("generating the 4,096-aggregate batch")
    sizes, xs, vs = B.batch_workload(L)
    order = np.argsort(sizes)
    # smallest, around each tier boundary, the largest in the chunk kernel and
    # the largest overall (wide launches)
    want = [50, 480, 900, 1100, 2000, 3500]
    picks = sorted({int(order[np.searchsorted(sizes[order], w)]) for w in want} | {int(order[-1])})
    return B, L, sizes, xs, vs, picks


def state_of(pool, slot):
    st = pool.state(slot)
    return {k: st[k] for k in ("x", "v", "theta", "omega", "a", "alpha")}


@pytest.mark.parametrize("precision", [1, 0])
def test_full_batch_systems_equal_alone(batch, precision):
    B, L, sizes, xs, vs, picks = batch
    announce(f"full batch {int(sizes.sum())} particles, precision {precision}, {STEPS} steps")
    p = B.params(L, precision)
    pool = L.Pool(p, max_particles=int(sizes.sum()) + 1024, max_systems=len(xs))
    slots = [pool.add(x, v) for x, v in zip(xs, vs)]
    pool.set_wide_threshold(WIDE_MIN)
    pool.step(B.DT, STEPS, rebuild_every=REBUILD)
    in_batch = {k: state_of(pool, slots[k]) for k in picks}
    pool.close()
    for k in picks:
        announce(f"system {k}: {int(sizes[k])} particles alone")
        alone = L.Pool(p, max_particles=int(sizes[k]) + 64, max_systems=1)
        s = alone.add(xs[k], vs[k])
        alone.set_wide_threshold(WIDE_MIN)
        alone.step(B.DT, STEPS, rebuild_every=REBUILD)
        got = state_of(alone, s)
        alone.close()
        for f, arr in got.items():
            np.testing.assert_array_equal(in_batch[k][f], arr, err_msg=f"system {k} ({sizes[k]}): {f}")
        if precision == 0:
            c, sp, h = (np.asarray(a) for a in B.aggregate_params())
            o = O.OracleSystem(O.system_params(O.BIN_SINTER, O.HAM_LIMITED, neighbor_mode=1, r_verlet=3.2 * O.R_PART,
                                               contact=c, sinter=sp, hamaker=h), xs[k], vs[k])
            o.step(B.DT, STEPS, rebuild_every=REBUILD)
            ref = o.state()
            for f in in_batch[k]:
                np.testing.assert_array_equal(in_batch[k][f], ref[f], err_msg=f"system {k} vs oracle: {f}")
    assert any(sizes[k] >= WIDE_MIN for k in picks)  # one system took the wide launches
