"""Writes the golden fixtures of the test suite from the CPU oracle.

    python tests/golden/gen_golden.py [--big]

Fixtures (data only: inputs as parameter dicts, outputs as integer arrays):
  v1_100_seed0.npz     rows 0..4000 of the 100x100 v1 workload, seed 0
  v2s_1600_seed0.npz   rows 0..4000 of the 1600x1600 v2-style workload, seed 0,
                       capacity 2,457,600 (--big: about five minutes of CPU)
  state_digests.json   SHA-256 of the full state (agents in slot order and
                       grass) after each of 100 iterations on small grids

The reference has no tests, golden outputs or fixtures for this path (its
RNG streams and sort come from the unvendored cl_ops library), so these pin
the oracle against regressions; the oracle itself is pinned by the checks in
tests/test_oracle_*.py.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import v1, v2s  # noqa: E402
from oracle.pp_oracle import Oracle, simulate  # noqa: E402

DIGEST_CASES = {
    "v1_8x8_seed3": (v1(gx=8, gy=8), 3),
    "v2s_16x12_seed5": (v2s(gx=16, gy=12), 5),
    "v2s_32x32_seed0": (v2s(32), 0),
}


def state_digest(state):
    h = hashlib.sha256()
    for a in state:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def digests(d, seed, iters=100):
    o = Oracle(d, seed=seed)
    out = [state_digest(o.state())]
    for _ in range(iters):
        assert o.step() == 0
        out.append(state_digest(o.state()))
    o.close()
    return out


def main():
    big = "--big" in sys.argv
    rows = simulate(v1(100), seed=0)
    np.savez_compressed(os.path.join(HERE, "v1_100_seed0.npz"), rows=rows,
                        params=json.dumps(v1(100)), seed=0, capacity=16777216)
    print("v1_100_seed0.npz", rows.shape)
    dig = {name: {"params": d, "seed": seed, "digests": digests(d, seed)}
           for name, (d, seed) in DIGEST_CASES.items()}
    with open(os.path.join(HERE, "state_digests.json"), "w") as f:
        json.dump(dig, f, indent=1)
    print("state_digests.json", len(dig))
    if big:
        cap = 4 * (409600 + 204800)
        rows = simulate(v2s(1600), seed=0, capacity=cap)
        np.savez_compressed(os.path.join(HERE, "v2s_1600_seed0.npz"),
                            rows=rows, params=json.dumps(v2s(1600)), seed=0,
                            capacity=cap)
        print("v2s_1600_seed0.npz", rows.shape)


if __name__ == "__main__":
    main()
