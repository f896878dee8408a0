"""Writes the reference-side focal measures for the statistical tests.

    python tests/golden/gen_statref.py [--replicas 30] [--procs 8]

Runs the independent NetLogo-rules simulator (oracle/pp_statref.c, MT19937)
on the dense v2-style 1600x1600 workload for 4000 iterations, seeds
1001..1000+replicas, and stores each replica's 36 focal measures
(pphpc_amd/replicas.py) in statref_v2s1600.json. Data only: parameters,
seeds and the resulting measures.
"""
import argparse
import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def one(seed):
    from helpers import v2s
    from oracle.pp_oracle import statref_run
    from pphpc_amd.replicas import focal_measures
    d = v2s(1600)
    rows = statref_run(d, seed)
    print(f"seed {seed} done", flush=True)
    return focal_measures(rows, d["grid_x"] * d["grid_y"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=30)
    ap.add_argument("--procs", type=int, default=8)
    args = ap.parse_args()
    from helpers import v2s
    seeds = list(range(1001, 1001 + args.replicas))
    with Pool(args.procs) as pool:
        focal = pool.map(one, seeds)
    out = {"params": v2s(1600), "seeds": seeds, "iters": 4000,
           "steady_from": 1000, "simulator": "NetLogo rules, MT19937",
           "focal": focal}
    with open(os.path.join(HERE, "statref_v2s1600.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
