"""The C++ drop-in layer (libgran_amd/include/libgran) against the CPU oracle.

tests/cpp/dropin_scenarios.cpp is written against the libgran template
interface (granular_system, granular_system_neighbor_list, the functor
containers, triangular_facet, libgran_amd::vec3) exactly as a program written
for the reference would be, compiled with the host compiler against the
drop-in headers and linked with libhipgran.so. It runs in a child process and
writes its final state; the FP64 device path is bit-exact with the oracle, so
the comparison is exact equality.
"""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import _scenarios as S

pytestmark = pytest.mark.gpu

R = O.R_PART
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "build", "dropin_scenarios")


def run_driver(tmp_path, scenario, steps):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C tests/cpp, or __graft_entry__.build())")
    out = tmp_path / f"{scenario}.bin"
    print(f"[gpu-dropin] {scenario} {steps} steps", flush=True)
    r = subprocess.run([DRIVER, scenario, str(steps), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"driver exit {r.returncode}: {r.stderr}"
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    vals = np.frombuffer(raw[8:], dtype=np.float64).reshape(4, n, 3)
    return dict(zip(("x", "v", "theta", "omega"), vals))


def vtk_oracle():
    import importlib.util
    spec = importlib.util.spec_from_file_location("vtk_oracle", os.path.join(ROOT, "oracle", "vtk_writer.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def compare(osys, got, what):
    ref = osys.state()
    for k in ("x", "v", "theta", "omega"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{what}: field {k}")


def oracle(binary_model, hamaker_model, x, v, **kw):
    return O.OracleSystem(O.system_params(binary_model, hamaker_model, **kw), x, v)


def test_dropin_contact(tmp_path):
    """test/contact.cpp set-up through granular_system<contact_force_functor>."""
    steps = 32000
    got = run_driver(tmp_path, "contact", steps)
    x, v = S.lattice_with_impactor()
    o = oracle(O.BIN_CONTACT, O.HAM_NONE, x, v)
    o.step(1e-13, steps)
    compare(o, got, "contact")
    assert np.abs(got["v"][:-1]).max() > 0.0  # the impactor reached the lattice


def test_dropin_hamaker_all_pairs(tmp_path):
    """test/hamaker.cpp set-up: contact + Hamaker in one binary container."""
    steps = 6000
    got = run_driver(tmp_path, "hamaker", steps)
    x, v = S.lattice_with_impactor()
    o = oracle(O.BIN_CONTACT, O.HAM_FULL, x, v, hamaker=O.hamaker_params(A=1e-20, h0=1e-9))
    o.step(1e-13, steps)
    compare(o, got, "hamaker")


def test_dropin_hamaker_neighbor_list(tmp_path):
    """test/hamaker_neighbor.cpp set-up: granular_system_neighbor_list with
    r_verlet = 5 r and update_neighbor_list() every 20 steps."""
    steps = 16000
    got = run_driver(tmp_path, "hamaker_neighbor", steps)
    x, v = S.lattice_with_impactor()
    o = oracle(O.BIN_CONTACT, O.HAM_FULL, x, v, neighbor_mode=1, r_verlet=5.0 * R,
               hamaker=O.hamaker_params(A=1e-20, h0=1e-9))
    o.step(1e-13, steps, rebuild_every=20)
    compare(o, got, "hamaker_neighbor")


def test_dropin_simplified_hamaker(tmp_path):
    """test/hamaker.cpp set-up with simplified_hamaker_functor as the second
    functor of the binary container."""
    steps = 4000
    got = run_driver(tmp_path, "simplified", steps)
    x, v = S.lattice_with_impactor()
    o = oracle(O.BIN_CONTACT, O.HAM_SIMPLIFIED, x, v, hamaker=O.hamaker_params(A=1e-20, h0=1e-9))
    o.step(1e-13, steps)
    compare(o, got, "simplified")


def test_dropin_alt_sintered(tmp_path):
    """test/alt_sintered.cpp set-up: alt_sinter_functor built from x0 (the
    driver also checks the dense bonded_contacts view and that
    get_contact_springs() reads the device springs)."""
    steps = 5000
    got = run_driver(tmp_path, "alt_sintered", steps)
    x, v = S.sinter_lattice()
    o = oracle(O.BIN_SINTER, O.HAM_NONE, x, v, contact=O.contact_params(dt=5e-14),
               sinter=O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14))
    o.step(5e-14, steps)
    compare(o, got, "alt_sintered")
    # the driver wrote the final state and the bonds' necks through the
    # drop-in writer.h (writer.cpp:51-172): the oracle's text for that state
    V = vtk_oracle()
    xs, ths = got["x"].tolist(), got["theta"].tolist()
    assert (tmp_path / "particles_1.vtk").read_text() == V.particles_text(xs, ths, R)
    bonds = [tuple(map(int, p)) for p in o.bonds()[0]]
    assert len(bonds) == 26
    assert (tmp_path / "necks_1.vtk").read_text() == V.necks_text(xs, bonds, R)


def test_dropin_sintered(tmp_path):
    """test/sintered.cpp set-up through the legacy sinter_functor and the step
    handler it hands out, unit vectors from the reference's own
    std::mt19937_64(0) generator: the oracle, given the vectors the driver
    drew, reproduces the run bit for bit."""
    steps = 5000
    got = run_driver(tmp_path, "sintered", steps)
    raw = (tmp_path / "sintered.bin.vecs").read_bytes()
    nv = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    vecs = np.frombuffer(raw[8:], dtype=np.float64).reshape(nv, 3)
    assert nv == 54  # one draw per bond of the lattice
    x, v = S.sinter_lattice()
    c = O.contact_params(dt=1e-13)
    s13 = O.sinter_params(k_bond=1.0e4, d_crit=1.0e-9, dt=1e-13)
    s13[1] = 0.0
    pairs, conn, used = O.legacy_frames(x, R, 1.0e-9, vecs)
    assert used == nv
    o = oracle(O.BIN_SINTER_LEGACY, O.HAM_NONE, x, v, contact=c, sinter=s13)
    o.set_legacy(pairs, conn)
    o.step(1e-13, steps)
    compare(o, got, "sintered")


def test_dropin_alt_breaking_sintered(tmp_path):
    """The same lattice through alt_breaking_sinter_functor (every touching
    pair bonded, neck strengths 1e-8 N * (1 + b % 7) in draw order); the driver
    also checks that bonded_contacts follows the breaks on the device."""
    steps = 6000
    got = run_driver(tmp_path, "breaking", steps)
    x, v = S.sinter_lattice()
    o = oracle(O.BIN_SINTER, O.HAM_NONE, x, v, contact=O.contact_params(dt=5e-14),
               sinter=O.sinter_params(k_bond=1000.0, d_crit=1.0e-9, dt=5e-14), sinter_breaking=1)
    nb = len(o.bonds()[0])
    o.set_neck_strengths(1.0e-8 * (1.0 + np.arange(nb) % 7))
    o.step(5e-14, steps)
    compare(o, got, "breaking")


@pytest.mark.parametrize("scenario", ["surface", "plate"])
def test_dropin_triangular_facet(tmp_path, scenario):
    """test/surface.cpp set-up: a triangular_facet in the unary container; the
    'plate' variant moves the facet (update_positions after every step) and
    switches on surface Hamaker."""
    steps = 3000
    got = run_driver(tmp_path, scenario, steps)
    x, v, w, verts = S.surface_contacts()
    moving = scenario == "plate"
    cp = O.contact_params(mu=0.5)
    o = O.OracleSystem(O.system_params(O.BIN_CONTACT, O.HAM_NONE, contact=cp, facet_contact=1, facet_hamaker=1,
                                       surface_contact=cp,
                                       surface_hamaker=O.hamaker_params(A=1e-20 if moving else 0.0, h0=1e-9)),
                       x, v, w0=w)
    o.set_facets(verts, np.array([[0.5, -0.2, 1.0]]) if moving else np.zeros((1, 3)))
    o.step(1e-13, steps)
    compare(o, got, scenario)
    assert got["v"][0, 2] > -0.5 and got["v"][1, 2] > -0.5  # both particles met the facet
