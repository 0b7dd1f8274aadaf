"""Writes tests/golden/functors.npz: inputs and outputs of the REFERENCE functor
headers (compiled from the reference tree into oracle/_ref/libref.so by
`make -C oracle ref`), for the pair/facet sequences of tests/_seq.py, plus the
bond topology of the reference's shipped aggregate (resources/aggregate.h) and
of the 3x3x3 lattice of test/alt_sintered.cpp:69-90, and sequences of the
legacy sinter_functor with its step handler (sinter_bridge.h).

Run from the repo root:  python tests/golden/make_golden.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402
import _seq  # noqa: E402

REF_ROOT = "/root/reference"


def read_aggregate_h():
    """resources/aggregate.h:8-159, coordinates in units of r_part."""
    txt = open(os.path.join(REF_ROOT, "resources", "aggregate.h")).read()
    body = txt[txt.index("{") + 1: txt.rindex("}")]
    nums = [float(t) for t in re.findall(r"-?\d+\.\d+", body)]
    return np.array(nums).reshape(-1, 3)


def lattice_333(r):
    """test/alt_sintered.cpp:69-77 particle positions."""
    x = []
    for i in range(3):
        y = -1.0 * r + i * 2.0 * r
        for j in range(3):
            xx = -1.0 * r + j * 2.0 * r
            x += [[xx, y, 0.0], [xx, y, -2.0 * r], [xx, y, -4.0 * r]]
    return np.array(x)


def legacy_four(r):
    """Four spheres: 0-1 touching along x, 1-2 along y with a 0.5 nm gap
    (within the 1 nm critical separation), 3 overlapping 0 by 3 nm (a contact,
    not a bond)."""
    return np.array([[0.0, 0.0, 0.0], [2.0 * r, 0.0, 0.0], [2.0 * r, 2.0 * r + 5e-10, 0.0],
                     [-(2.0 * r - 3e-9) * 0.6, (2.0 * r - 3e-9) * 0.8, 0.0]])


def main():
    if O.ref() is None:
        raise SystemExit("build the reference harness first: make -C oracle ref")
    out = {}
    contact_sets = {
        "std": O.contact_params(),
        "slip": O.contact_params(mu=0.05, mu_o=0.01, k=2.0e4),
        "dyn": O.contact_params(mu=0.3, phi=0.7, gamma_n=5e-8),
    }
    sinter = O.sinter_params()
    ham = O.hamaker_params(h0=1.0e-10)
    for cname, cp in contact_sets.items():
        out[f"contact_{cname}_params"] = cp
        for seed in range(3):
            s = _seq.pair_sequence(100 + seed, v_sigma=2.0 if cname != "slip" else 20.0)
            out[f"contact_{cname}_{seed}_states"] = s
            out[f"contact_{cname}_{seed}_out"] = O.ref_contact_seq(cp, s)
    out["sinter_params"] = sinter
    for seed in range(3):
        s = _seq.bonded_sequence(200 + seed)
        o, sp, b = O.ref_sinter_seq(sinter, contact_sets["std"], s)
        assert b == 1
        out[f"sinter_{seed}_states"] = s
        out[f"sinter_{seed}_out"] = o
        out[f"sinter_{seed}_springs"] = sp
    # alt_breaking_sinter: the ordered pair (0, 1) with a neck strength that
    # the sequence's force exceeds part-way (0.9 of the unbroken run's peak)
    for seed in range(3):
        s = _seq.bonded_sequence(250 + seed)
        o_inf, _ = O.ref_breaking_seq(sinter, contact_sets["std"], float("inf"), s)
        strength = 0.9 * float(np.max(np.linalg.norm(o_inf[:, :3], axis=1))) * O.MASS
        o, broken = O.ref_breaking_seq(sinter, contact_sets["std"], strength, s)
        assert 0 < broken.sum() < len(broken)
        out[f"breaking_{seed}_states"] = s
        out[f"breaking_{seed}_strength"] = np.array(strength)
        out[f"breaking_{seed}_out"] = o
        out[f"breaking_{seed}_broken"] = broken
    out["ham_params"] = ham
    for seed in range(2):
        s = _seq.far_sequence(300 + seed)
        out[f"ham_{seed}_states"] = s
        out[f"ham_full_{seed}_out"] = O.ref_hamaker_seq(O.HAM_FULL, ham, s)
        out[f"ham_lr_{seed}_out"] = O.ref_hamaker_seq(O.HAM_LIMITED, ham, s)
        out[f"ham_simplified_{seed}_out"] = O.ref_hamaker_seq(O.HAM_SIMPLIFIED, ham, s)
    for seed in range(2):
        s = _seq.bonded_sequence(400 + seed, amp=0.08)
        out[f"fold_{seed}_states"] = s
        out[f"fold_{seed}_out"] = O.ref_fold_seq(sinter, contact_sets["std"], ham, s)
    verts = np.array([[-3.0, -3.0, 0.0], [-3.0, 3.0, 0.0], [3.0, -3.0, 0.0]]) * O.R_PART
    out["facet_verts"] = verts
    for seed, vf in enumerate(([0.0, 0.0, 0.0], [0.3, -0.2, 0.5])):
        s = _seq.facet_sequence(500 + seed)
        out[f"facet_{seed}_vf"] = np.array(vf)
        out[f"facet_{seed}_states"] = s
        out[f"facet_{seed}_out"] = O.ref_facet_seq(verts, np.array(vf), contact_sets["std"],
                                                   O.hamaker_params(h0=1.0e-10), 1.0e-11, s)
    # Topologies: the shipped aggregate with the configurations' d_crit = 1e-3 r
    agg = read_aggregate_h() * O.R_PART
    out["agg150_x"] = agg
    out["agg150_bonds"] = O.ref_topology(agg, O.sinter_params(d_crit=1e-3 * O.R_PART), contact_sets["std"])
    lat = lattice_333(O.R_PART)
    out["lattice_x"] = lat
    out["lattice_bonds"] = O.ref_topology(lat, O.sinter_params(d_crit=1.0e-9), contact_sets["std"])
    # sinter_functor + sinter_step_handler (sinter_bridge.h): frames from
    # the given unit vectors, per-step increments through the step handler,
    # every ordered pair evaluated
    for name, x0, steps in (("legacy4", legacy_four(O.R_PART), 40), ("legacy_lattice", lat, 3)):
        n = len(x0)
        rng = np.random.default_rng(77 + n)
        vecs = rng.normal(size=(400, 3))
        vecs /= np.linalg.norm(vecs, axis=1, keepdims=True)
        vecs[0] = [1.0, 0.0, 0.0]  # collinear with the first bond's normal: drawn again
        incs = np.concatenate([rng.normal(0.0, 2e-11, (steps, n, 3)), rng.normal(0.0, 2e-3, (steps, n, 3)),
                               rng.normal(0.0, 1.0, (steps, n, 3)), rng.normal(0.0, 1e7, (steps, n, 3))], axis=2)
        th0 = rng.normal(0.0, 0.5, (n, 3))
        s13 = O.sinter_params(k_bond=1000.0, d_crit=1.0e-9)
        s13[1] = 3e-9  # bond damping gamma_d
        ref, used = O.ref_legacy_seq(n, x0, th0, s13, contact_sets["std"], vecs, incs)
        out[f"{name}_x0"], out[f"{name}_th0"], out[f"{name}_vecs"] = x0, th0, vecs
        out[f"{name}_incs"], out[f"{name}_sinter"], out[f"{name}_out"] = incs, s13, ref
        out[f"{name}_used"] = np.array([used])
    np.savez_compressed(os.path.join(HERE, "functors.npz"), **out)
    print("wrote", os.path.join(HERE, "functors.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
