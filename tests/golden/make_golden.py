"""Write tests/golden/oracle_vectors.npz: seeded inputs with the oracle's outputs.

The inputs are the reference tests' LCG generators (tests/inputs.py); the outputs come
from the CPU oracle (oracle/lbvh_oracle.c), itself pinned by tests/test_oracle.py. The
fixture freezes those outputs so a later change to the oracle that alters any of them
fails tests/test_oracle.py::test_oracle_matches_committed_vectors, and GPU tests can
compare against data without rebuilding it.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import inputs as I  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_vectors.npz")


def main():
    d = {}
    for seed in (1, 7, 42):
        pos = I.cloud(seed, 2000, 3.0)
        b, lanes, codes = O.reference_morton(pos)
        order = O.reference_sort(codes)
        left, right, parent = O.reference_karras(codes[order])
        d[f"cloud{seed}_pos"] = pos
        d[f"cloud{seed}_bounds"] = np.concatenate([b["bmin"], [b["size"], b["scale"]]])
        d[f"cloud{seed}_codes"] = codes
        d[f"cloud{seed}_order"] = order
        d[f"cloud{seed}_left"] = left
        d[f"cloud{seed}_right"] = right
        d[f"cloud{seed}_parent"] = parent
    codes = I.random_codes(2, 4000)
    d["random2_codes"] = codes
    d["random2_order"] = O.reference_sort(codes)
    for seed in (0, 1):
        pos, _, mass = I.cluster(seed, 120)
        acc, cnt = O.lbvh_accelerations(pos, mass, 1.0, 0.05, 0.5)
        d[f"cluster{seed}_pos"] = pos
        d[f"cluster{seed}_mass"] = mass
        d[f"cluster{seed}_lbvh_theta05"] = acc
        d[f"cluster{seed}_lbvh_theta05_count"] = np.array([cnt], np.uint64)
        d[f"cluster{seed}_direct"] = O.direct_sum(pos, mass, 1.0, 0.05)
    np.savez_compressed(OUT, **d)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
