"""hg_pool_energy_noalloc, the entry point that allocates nothing, against
hg_pool_energy, both called through the C ABI.

Two calls in a row on one GPU, the second on a smaller system than the first,
must each equal the allocating entry point bit for bit (the shared device
array keeps the first call's terms beyond the second system's particles). A
system above the array's size is refused, and repeated calls leave the
device's free memory unchanged.
"""
import ctypes as C

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


def energies(pool, slot):
    L = _L()
    lib = L._abi.lib()
    want, got = np.zeros(8), np.zeros(8)
    L._abi.check(lib.hg_pool_energy(pool._h, slot, O.MASS, O.INERTIA, L._abi.dptr(want)))
    L._abi.check(lib.hg_pool_energy_noalloc(pool._h, slot, O.MASS, O.INERTIA, L._abi.dptr(got)))
    return want, got


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_noalloc_energy_equals_allocating_entry_point(precision):
    print(f"[no-alloc] ledger equality, precision {precision}", flush=True)
    L = _L()
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h, precision=precision)
    pool = L.Pool(p, max_particles=4000, max_systems=4)
    sizes = (3000, 150)
    slots = []
    for n, seed in zip(sizes, (71, 72)):
        x = L.generate_aggregate(n, R, 1.8, 1.3, seed)
        slots.append(pool.add(x, np.random.default_rng(seed).normal(0.0, 1.0, x.shape)))
    pool.step(DT, 60, rebuild_every=50)
    for slot, n in zip(slots, sizes):  # the larger system first, then the smaller one
        want, got = energies(pool, slot)
        np.testing.assert_array_equal(got, want, err_msg=f"system of {n} particles")
        assert np.all(np.isfinite(got)) and got[0] > 0.0
    # again after more steps, smaller system first this time
    pool.step(DT, 25, rebuild_every=50)
    for slot in reversed(slots):
        want, got = energies(pool, slot)
        np.testing.assert_array_equal(got, want)
    pool.close()


def test_noalloc_energy_refuses_systems_above_its_array():
    print("[no-alloc] size limit", flush=True)
    L = _L()
    lib = L._abi.lib()
    m = 51  # 51^3 = 132,651 particles, spaced 4 r apart (no pairs)
    g = np.arange(m) * 4.0 * R
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    assert x.shape[0] > L._abi.HG_NOALLOC_MAX_PARTICLES
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h)
    pool = L.Pool(p, max_particles=x.shape[0] + 16, max_systems=1)
    slot = pool.add(x)
    out = np.zeros(8)
    st = lib.hg_pool_energy_noalloc(pool._h, slot, O.MASS, O.INERTIA, L._abi.dptr(out))
    assert st == L._abi.HG_ERR_CAPACITY
    assert "131072" in lib.hg_last_error().decode()
    pool.close()


def test_noalloc_energy_leaves_device_memory_unchanged():
    print("[no-alloc] free device memory", flush=True)
    L = _L()
    lib = L._abi.lib()
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    c, s, h = S.aggregate_params()
    p = L.make_params(L.BINARY_SINTER, L.HAMAKER_LIMITED, pair_mode=L.PAIRS_NEIGHBOR_LIST, r_verlet=3.2 * R,
                      contact=c, sinter=s, hamaker=h)
    pool = L.Pool(p, max_particles=2000, max_systems=2)
    x = L.generate_aggregate(1500, R, 1.8, 1.3, 73)
    slot = pool.add(x, np.random.default_rng(73).normal(0.0, 1.0, x.shape))
    pool.step(DT, 10, rebuild_every=50)
    out = np.zeros(8)
    L._abi.check(lib.hg_pool_energy_noalloc(pool._h, slot, O.MASS, O.INERTIA, L._abi.dptr(out)))
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    before = free.value
    for _ in range(5):
        L._abi.check(lib.hg_pool_energy_noalloc(pool._h, slot, O.MASS, O.INERTIA, L._abi.dptr(out)))
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    assert free.value == before
    pool.close()
