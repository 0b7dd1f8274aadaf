"""The legacy sinter_functor's host set-up (hg_legacy_sinter_frames,
sinter_bridge.h:138-211) against the oracle's restatement, whose frame
geometry test_oracle_pins.py holds to its closed form: the same bonds in the
same order, the same unit vectors drawn (rejections included) and
bit-identical connectors. Host code: no GPU needed."""
import numpy as np
import pytest

import _oracle as O

R = O.R_PART


def _L():
    import libgran_amd
    return libgran_amd


def unit_vectors(seed, count):
    v = np.random.default_rng(seed).uniform(-1.0, 1.0, (count, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def drawing(vecs):
    """A next_unit_vector() over vecs that counts its draws."""
    state = {"k": 0}

    def nxt():
        k = state["k"]
        state["k"] += 1
        return vecs[k]
    return nxt, state


@pytest.mark.parametrize("case", ["lattice", "aggregate", "gapped"])
def test_frames_equal_oracle(case):
    L = _L()
    if case == "lattice":  # test/sintered.cpp:77-98
        import _scenarios as S
        x, _ = S.sinter_lattice()
        d_crit = 1.0e-9
    elif case == "aggregate":
        x = L.generate_aggregate(600, R, 1.8, 1.3, 5)
        d_crit = 1e-3 * R
    else:  # gaps of +-(d_crit less one ulp-scale rounding) and d_crit + 1e-16 m
        x = np.array([[0.0, 0.0, 0.0], [2.0 * R + 1e-9, 0.0, 0.0], [2.0 * R + 1e-9, 2.0 * R - 1e-9, 0.0],
                      [2.0 * R + 1e-9, 4.0 * R - 1e-9 + 1.0000001e-9, 0.0]])
        d_crit = 1.0e-9
    vecs = unit_vectors(11, 5000)
    vecs[0] = [0.0, 0.0, 0.0]  # a zero draw is collinear with every normal: drawn again
    vecs[1] = [1.0, 0.0, 0.0] if case == "gapped" else vecs[1]
    nxt, state = drawing(vecs)
    pairs, conn = L.legacy_sinter_frames(x, R, d_crit, nxt)
    opairs, oconn, used = O.legacy_frames(x, R, d_crit, vecs)
    np.testing.assert_array_equal(pairs, opairs)
    np.testing.assert_array_equal(conn, oconn)
    assert state["k"] == used
    assert len(pairs) > 0
    if case == "gapped":
        # separation and overlap at d_crit (within rounding) bond, 1e-16 m more does not
        assert pairs.tolist() == [[0, 1], [1, 2]]


def test_frames_count_only_draws_nothing():
    L = _L()
    import ctypes as C
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R, 0.0, 0.0]])
    nb = C.c_size_t(0)
    L._abi.check(L._abi.lib().hg_legacy_sinter_frames(2, L._abi.dptr(x), R, 1e-9, None, None, None, None, 0,
                                                      C.byref(nb)))
    assert nb.value == 1


def test_frames_source_errors_propagate():
    L = _L()
    x = np.array([[0.0, 0.0, 0.0], [2.0 * R, 0.0, 0.0]])

    def broken():
        raise RuntimeError("no more vectors")
    with pytest.raises(RuntimeError, match="no more vectors"):
        L.legacy_sinter_frames(x, R, 1e-9, broken)
    # a source of vectors collinear with the normal only never ends the draw
    # on its own; the library gives up with an error
    with pytest.raises(L.HipgranError):
        L.legacy_sinter_frames(x, R, 1e-9, lambda: (1.0, 0.0, 0.0))
