"""Writes pop_mpqp1.npz: the inputs of the 100 problems of the reference's
POP_mpQP1.mat (the struct fields main.py:73-80 reads: A, b, c, F, Q, Ht, CRA,
CRb), keyed "<id>_<field>", float64.  The .mat holds inputs only; no expected
outputs exist in the reference.  Run in a checkout where the reference data is
readable:  python tests/golden/make_pop.py /path/to/POP_mpQP1.mat
"""
import os
import sys

import numpy as np
import scipy.io

FIELDS = ("A", "b", "c", "F", "Q", "Ht", "CRA", "CRb")


def main(path):
    ds = scipy.io.loadmat(path)["problem"][0]
    out = {}
    for pid in range(len(ds)):
        for f in FIELDS:
            out[f"{pid}_{f}"] = np.array(ds[pid][f], dtype=np.float64)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pop_mpqp1.npz")
    np.savez_compressed(dst, **out)
    print(dst, len(ds), "problems")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/POP_mpQP1.mat")
