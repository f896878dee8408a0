"""Writes the reference-side focal measures for the statistical tests.

    python tests/golden/gen_statref.py [--replicas 30] [--procs 8]
                                       [--rules cells|netlogo]

Runs the independent simulator (oracle/pp_statref.c, MT19937) on the dense
v2-style 1600x1600 workload for 4000 iterations, seeds 1001..1000+replicas,
and stores each replica's 36 focal measures (pphpc_amd/replicas.py):
  --rules cells    the Java single-thread semantics (per-cell shuffled acting
                   order, a wolf takes the first live sheep of the list) ->
                   statref_v2s1600.json
  --rules netlogo  the NetLogo rules (global random order, a wolf takes a
                   uniformly random sheep) -> statref_netlogo_v2s1600.json
Data only: parameters, rule set, seeds and the resulting measures.
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

SIMULATOR = {
    "cells": "Java single-thread rules (per-cell shuffled acting order, wolf "
             "takes the first live sheep of the list), MT19937",
    "netlogo": "NetLogo rules (global random acting order, wolf takes a "
               "uniformly random sheep), MT19937",
}


def one(args):
    seed, rules = args
    from helpers import v2s
    from oracle.pp_oracle import statref_run
    from pphpc_amd.replicas import focal_measures
    d = v2s(1600)
    rows = statref_run(d, seed, rules=rules)
    print(f"seed {seed} done", flush=True)
    return focal_measures(rows, d["grid_x"] * d["grid_y"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=30)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--rules", choices=list(SIMULATOR), default="cells")
    args = ap.parse_args()
    from helpers import v2s
    seeds = list(range(1001, 1001 + args.replicas))
    with Pool(args.procs) as pool:
        focal = pool.map(one, [(s, args.rules) for s in seeds])
    out = {"params": v2s(1600), "seeds": seeds, "iters": 4000,
           "steady_from": 1000, "rules": args.rules,
           "simulator": SIMULATOR[args.rules], "focal": focal}
    name = ("statref_v2s1600.json" if args.rules == "cells"
            else "statref_netlogo_v2s1600.json")
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
