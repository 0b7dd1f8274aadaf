"""Writes tests/golden/aggregate150.npz: the coordinates of the reference's
shipped soot aggregate (resources/aggregate.h:8-159, 150 touching spheres, in
units of the particle radius), the only geometry fixture the reference holds.
Everything else the tests compare against is computed (closed forms, brute
force, the reference's known answers in test/*.cpp).

Run from the repo root, where the reference tree is present:
    python tests/golden/make_golden.py
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference"


def read_aggregate_h():
    """resources/aggregate.h:8-159, coordinates in units of r_part."""
    txt = open(os.path.join(REF_ROOT, "resources", "aggregate.h")).read()
    body = txt[txt.index("{") + 1: txt.rindex("}")]
    nums = [float(t) for t in re.findall(r"-?\d+\.\d+", body)]
    return np.array(nums).reshape(-1, 3)


def main():
    x = read_aggregate_h()
    assert x.shape == (150, 3)
    np.savez_compressed(os.path.join(HERE, "aggregate150.npz"), x_in_r=x)
    print(f"wrote aggregate150.npz ({x.shape[0]} particles)")


if __name__ == "__main__":
    main()
