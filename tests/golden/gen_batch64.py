"""Writes the fixture of the 64-replica batch workload from the CPU oracle.

    python tests/golden/gen_batch64.py [--threads T] [--only I,J,...]

The workload: 64 replicas whose grid side is drawn uniformly from
{100, 200, 400, 800, 1600, 3200}, v1 or v2s with probability 1/2, seeds
1..64, 4000 iterations each (pphpc_amd.workloads.batch_replicas(64)), each
with the agent reservation the batch gives it (workloads.reserve_for).

Every replica is run through oracle/build/pp_omp, the multi-threaded
restatement whose results equal the single-threaded oracle's bit for bit at
any thread count (tests/test_omp_baseline.py), because the single-threaded
oracle would take hours over the 3200 x 3200 replicas.

Outputs (data only):
  batch64_digests.json  per replica: kind, side, seed, reservation, rows
                        produced, error code, the SHA-256 of all its rows
                        (uint64 little-endian, seven columns, rows 0..4000)
                        and of each chunk of 500 rows (to name where a
                        mismatch starts)
  batch64_small.npz     the full rows of every replica of side <= 400
                        (array "r<index>")

About 20 minutes on 8 cores.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from pphpc_amd.workloads import batch_replicas, reserve_for  # noqa: E402

KEYS = ("INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
        "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
        "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
        "GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS")
CHUNK = 500
SMALL_SIDE = 400


def write_params(path, d):
    with open(path, "w") as f:
        for k in KEYS:
            f.write(f"{k} = {d[k.lower()]}\n")


def digest(rows):
    return hashlib.sha256(np.ascontiguousarray(rows, dtype="<u8").tobytes()).hexdigest()


def chunk_digests(rows):
    return [digest(rows[i:i + CHUNK]) for i in range(0, len(rows), CHUNK)]


def run_one(omp, d, seed, reserve, threads, tmp):
    prm = os.path.join(tmp, "p.txt")
    raw = os.path.join(tmp, "rows.bin")
    write_params(prm, d)
    r = subprocess.run([omp, "-p", prm, "-r", str(seed), "-m", str(reserve),
                        "-t", str(threads), "-q", "-b", raw],
                       capture_output=True, text=True)
    rows = np.fromfile(raw, dtype="<u8").reshape(-1, 7)
    return r.returncode, rows, r.stdout.strip()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    omp = os.path.join(ROOT, "oracle", "build", "pp_omp")
    if not os.path.exists(omp):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    work = batch_replicas(64)
    only = {int(x) for x in a.only.split(",") if x}
    out_json = os.path.join(HERE, "batch64_digests.json")
    out_npz = os.path.join(HERE, "batch64_small.npz")
    have = {}
    if only and os.path.exists(out_json):
        with open(out_json) as f:
            have = {r["index"]: r for r in json.load(f)["replicas"]}
    small = dict(np.load(out_npz)) if only and os.path.exists(out_npz) else {}
    with tempfile.TemporaryDirectory() as tmp:
        for i, (kind, L, d, seed) in enumerate(work):
            if only and i not in only:
                continue
            res = reserve_for(d)
            t0 = time.time()
            rc, rows, msg = run_one(omp, d, seed, res, a.threads, tmp)
            have[i] = {"index": i, "kind": kind, "side": L, "seed": seed,
                       "reserve": res, "rows": int(len(rows)), "error": rc,
                       "sha256": digest(rows), "chunks": chunk_digests(rows)}
            if L <= SMALL_SIDE:
                small[f"r{i}"] = rows
            print(f"{i:2d} {kind:3s} {L:5d} seed {seed:2d} rc {rc} rows "
                  f"{len(rows)} {time.time() - t0:6.1f} s  {msg}", flush=True)
    with open(out_json, "w") as f:
        json.dump({"workload": "batch_replicas(64), 4000 iterations",
                   "chunk_rows": CHUNK,
                   "replicas": [have[i] for i in sorted(have)]}, f, indent=1)
    np.savez_compressed(out_npz, **small)


if __name__ == "__main__":
    main()
