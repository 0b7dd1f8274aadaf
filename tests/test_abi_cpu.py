"""Host-side checks of the hipgran library that need no GPU: the shared
object loads, exports every function include/hipgran.h declares, and its host
logic (bond topology, aggregate generator) agrees with the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import _oracle as O

libgran_amd = pytest.importorskip("libgran_amd")
from libgran_amd import _abi  # noqa: E402

AGG150 = np.load(os.path.join(os.path.dirname(__file__), "golden", "aggregate150.npz"))["x_in_r"] * O.R_PART


def test_library_exports_every_declared_symbol():
    declared = _abi.header_symbols()
    assert len(declared) >= 18
    L = C.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(_abi.SIGNATURES), "binding and header disagree"


def test_params_struct_layout_matches_header():
    # 10 int32 then a double: r_verlet at byte 40; 59 doubles of functor parameters.
    assert _abi.Params.r_verlet.offset == 40
    assert C.sizeof(_abi.Params) == 48 + 8 * (18 + 13 + 5 + 18 + 5)


def test_topology_matches_reference_geometries():
    """The host bond search on the reference's shipped aggregate
    (resources/aggregate.h, a tree: 149 bonds) and the 3x3x3 lattice of
    test/alt_sintered.cpp (26 bonds, 28 cycle-closing pairs skipped), against
    the oracle's, which test_oracle_pins.py holds to an independent union-find
    scan."""
    import _scenarios as S
    pairs, sk = libgran_amd.sinter_topology(AGG150, O.R_PART, 1e-3 * O.R_PART)
    np.testing.assert_array_equal(pairs, O.oracle_topology(AGG150, O.R_PART, 1e-3 * O.R_PART)[0])
    assert len(pairs) == 149 and sk == 0
    xl, _ = S.sinter_lattice()
    pairs, sk = libgran_amd.sinter_topology(xl, O.R_PART, 1.0e-9)
    np.testing.assert_array_equal(pairs, O.oracle_topology(xl, O.R_PART, 1.0e-9)[0])
    assert len(pairs) == 26 and sk == 28


@pytest.mark.parametrize("n,seed,df", [(200, 0, 1.8), (1000, 3, 2.2), (3000, 1, 1.6)])
def test_topology_matches_oracle_on_generated_aggregates(n, seed, df):
    x = libgran_amd.generate_aggregate(n, O.R_PART, df, 1.3, seed)
    for d_crit in (1e-3 * O.R_PART, 1e-9, 0.2 * O.R_PART):
        p1, s1 = libgran_amd.sinter_topology(x, O.R_PART, d_crit)
        p2, s2 = O.oracle_topology(x, O.R_PART, d_crit)
        np.testing.assert_array_equal(p1, p2)
        assert s1 == s2


def test_generator_is_deterministic_touching_and_connected():
    n = 400
    x = libgran_amd.generate_aggregate(n, O.R_PART, 1.8, 1.3, 7)
    x2 = libgran_amd.generate_aggregate(n, O.R_PART, 1.8, 1.3, 7)
    np.testing.assert_array_equal(x, x2)
    assert not np.array_equal(x, libgran_amd.generate_aggregate(n, O.R_PART, 1.8, 1.3, 8))
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    assert d.min() > 2 * O.R_PART * (1 - 1e-9), "spheres overlap"
    pairs, sk = libgran_amd.sinter_topology(x, O.R_PART, 1e-3 * O.R_PART)
    assert len(pairs) == n - 1 and sk >= 0  # the touching graph spans the aggregate
    # centred on its centre of mass
    assert np.linalg.norm(x.mean(axis=0)) < 1e-9 * O.R_PART * n


def test_generator_fractal_scaling():
    """N = kf (Rg/r)^Df holds approximately for the generated aggregates."""
    for df in (1.6, 1.8, 2.2):
        n = 2000
        x = libgran_amd.generate_aggregate(n, O.R_PART, df, 1.3, 11)
        rg = np.sqrt(((x - x.mean(axis=0)) ** 2).sum(axis=1).mean())
        n_pred = 1.3 * (rg / O.R_PART) ** df
        assert 0.5 < n_pred / n < 2.0, (df, n_pred)


LEGACY_DRIVER = r"""
#include <functional>
#include <libgran/granular_system/granular_system.h>
#include <libgran/sinter_bridge/sinter_bridge.h>
#include <libgran/vec3.h>
using V = libgran_amd::vec3;
template <typename C, typename F> using handler_double = sinter_step_handler<C, F, double>;
int main() {
    std::vector<V> x0{V(0, 0, 0), V(2.8e-8, 0, 0)}, z(2);
    std::function<V()> gen = [] { return V(0, 1, 0); };
    sinter_functor<V, double> s(2, 1e4, 0.0, 1e4, 5e-9, 1e4, 1e-9, 1.0, 1.0, 1e4, 2.5e-10, 0.1, 1.0, 1e4, 2.5e-10,
                                0.1, 1.0, 1.4e-8, 1e-20, 1e-36, 1e-13, V(), 0.0, x0.cbegin(), gen, 1e-9);
    binary_force_functor_container<V, double, sinter_functor<V, double>> b{s};
    unary_force_functor_container<V, double> u;
    auto h = s.get_step_handler<std::vector<V>>();
    granular_system<V, double, rotational_velocity_verlet_half, HANDLER, decltype(b), decltype(u)>
        sys(x0, z, z, z, 0.0, V(), 0.0, h, b, u);
    return 0;
}
"""


@pytest.mark.parametrize("handler,ok", [("handler_double", True), ("rotational_step_handler", False)])
def test_dropin_legacy_sinter_needs_its_step_handler(tmp_path, handler, ok):
    """sinter_functor compiles with the step handler it hands out
    (sinter_bridge.h:228-232, as test/sintered.cpp:117-127 uses it) and is a
    compile-time error with the default handler: the device moves the
    connectors, which the default handler would leave behind."""
    import subprocess
    src = tmp_path / "d.cpp"
    src.write_text(LEGACY_DRIVER.replace("HANDLER", handler))
    if handler == "rotational_step_handler":  # the default handler would also not bind to h's type
        src.write_text(src.read_text().replace("auto h = s.get_step_handler<std::vector<V>>();",
                                               "rotational_step_handler<std::vector<V>, V> h;"))
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libgran_amd", "include")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, str(src)], capture_output=True, text=True)
    if ok:
        assert r.returncode == 0, r.stderr
    else:
        assert r.returncode != 0 and "sinter_functor's own step handler" in r.stderr


def test_snapshot_loader_refuses_foreign_and_damaged_bytes():
    """hg_load checks the magic, checksum and sizes before it touches a
    device, so these run on the host."""
    lib = _abi.lib()
    out = C.c_void_p()
    for junk in (b"", b"HGSNAP01", b"not a snapshot at all, just some bytes" * 10, b"HGSNAP01" + bytes(64)):
        st = lib.hg_load(junk, len(junk), C.byref(out))
        assert st == _abi.HG_ERR_INVALID, junk[:16]
        assert not out.value
    assert "checksum" in (lib.hg_load(b"HGSNAP01" + bytes(64), 72, C.byref(out)) and lib.hg_last_error().decode())
