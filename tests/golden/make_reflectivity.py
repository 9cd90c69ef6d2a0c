"""Writes tests/golden/reflectivity.npz: the reference's own mesh-parameter
sampler, ReflectivityAndSeeds of src/interferometer.py:9-54, run on a few
(n_modes, seed) cases.  Its reflectivities come from scipy's generic
rv_continuous sampler (a numeric inverse CDF driven by numpy's global RNG),
which no closed form pins bit for bit, so the vectors are taken from the
reference itself: this script imports that one file (numpy and scipy only)
from /root/reference in the build container.  Run from the repository root:
`python tests/golden/make_reflectivity.py`.
"""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/interferometer.py"
CASES = [(4, 0), (6, 1), (8, 7), (12, 3), (16, 11)]   # even n_modes: the reference
                                                      # raises IndexError for odd ones


def used_slots(M):
    """(layer, index) slots that hold a beam splitter (mode l % 2 + 2 i <= M - 2);
    the reference leaves the others as np.empty's contents."""
    mask = np.zeros((M, M // 2), bool)
    for lay in range(M):
        for i, m in enumerate(range(lay % 2, M - 1, 2)):
            mask[lay, i] = True
    return mask


def main():
    spec = importlib.util.spec_from_file_location("reference_interferometer", SRC)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    out = {}
    for M, seed in CASES:
        r, s = ref.ReflectivityAndSeeds(M, seed)
        mask = used_slots(M)
        out[f"r_{M}_{seed}"] = np.where(mask, r, 0.0)
        out[f"s_{M}_{seed}"] = np.where(mask, s, 0)
    out["cases"] = np.array(CASES)
    np.savez_compressed(os.path.join(HERE, "reflectivity.npz"), **out)
    print("wrote reflectivity.npz for", CASES)


if __name__ == "__main__":
    main()
