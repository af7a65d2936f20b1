"""Writes tests/golden/mt_ref.npz: the MT19937 stream of the reference's own
mtwist.c (compiled from its source into oracle/_ref by oracle/Makefile) for the
seeds the engine uses, so the oracle's restatement can be checked where the
reference tree is absent.

Run from the repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def main():
    oracle.build()
    out = {}
    for seed in (2013, 4357, 1, 123456789):
        s = oracle.ref_mt_stream(seed, 2000)
        if s is None:
            raise SystemExit("oracle/_ref/libmtwist_ref.so is missing (reference tree absent?)")
        out[f"seed_{seed}"] = s
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mt_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
