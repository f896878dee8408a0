"""The batched and serving workloads at their full size, bit for bit:

* batched: the 64 replicas of the batched-replica workload (grid sides drawn
  from {100, ..., 3200}, v1 or v2s, seeds 1..64, 4000 iterations) as one
  batch with the autotuner's shape; every replica's 4001 statistics rows
  equal the CPU oracle's, recorded in tests/golden/batch64_digests.json by
  tests/golden/gen_batch64.py (full rows for the grids up to 400 x 400);
* serving: a burst of requests including 3200 x 3200 and 6400 x 6400 grids
  and 1000..4000 iterations, admitted into a batch smaller than their total
  footprint: each result equals the same simulation alone, and the requests
  up to 400 x 400 equal the oracle.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _sha(rows):
    return hashlib.sha256(np.ascontiguousarray(rows, dtype="<u8").tobytes()
                          ).hexdigest()


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN,
                                                     "batch64_digests.json")),
                    reason="fixture not written (tests/golden/gen_batch64.py)")
def test_batch64_workload_equals_oracle_fixture():
    from pphpc_amd import Batch, Params
    from pphpc_amd.workloads import batch_replicas, reserve_for
    with open(os.path.join(GOLDEN, "batch64_digests.json")) as f:
        fix = json.load(f)
    small = np.load(os.path.join(GOLDEN, "batch64_small.npz"))
    chunk = fix["chunk_rows"]
    reps = batch_replicas(64)
    assert len(fix["replicas"]) == len(reps)
    cells = sum(L * L for _, L, _, _ in reps)
    agents = sum(reserve_for(d) for _, _, d, _ in reps)
    _case(f"64-replica batch, {cells} cells, {agents} reserved agents, tuned")
    with Batch(agents, cells, len(reps), 10, device=0) as b:
        ids = []
        for (kind, L, d, seed), fx in zip(reps, fix["replicas"]):
            assert (fx["kind"], fx["side"], fx["seed"], fx["reserve"]) == \
                (kind, L, seed, reserve_for(d))
            ids.append(b.add(Params.from_values(**d), seed, reserve_for(d)))
        tuned = b.tune(16)
        _case(f"tuned shape {tuned}")
        while b.resources()["active"]:
            b.run(512)
        for i, (rid, fx) in enumerate(zip(ids, fix["replicas"])):
            st = b.replica(rid)
            rows = b.rows_array(rid)
            what = f"replica {i} ({fx['kind']} {fx['side']} seed {fx['seed']})"
            assert st["error"] == fx["error"], (what, st)
            assert len(rows) == fx["rows"], (what, len(rows))
            if f"r{i}" in small.files:
                ref = small[f"r{i}"]
                bad = np.nonzero((rows != ref).any(axis=1))[0]
                assert len(bad) == 0, (what, int(bad[0]), rows[bad[0]],
                                       ref[bad[0]])
            if _sha(rows) != fx["sha256"]:
                first = next(k for k, h in enumerate(fx["chunks"])
                             if _sha(rows[k * chunk:(k + 1) * chunk]) != h)
                raise AssertionError(f"{what}: rows differ from the oracle "
                                     f"in rows {first * chunk}.."
                                     f"{(first + 1) * chunk - 1}")


def test_serving_large_requests_same_alone_and_as_oracle():
    from pphpc_amd import Params, Simulation, stats_rows_to_array
    from pphpc_amd.serve import Server
    from pphpc_amd.workloads import reserve_for, serve_requests
    from oracle.pp_oracle import simulate
    stream = serve_requests(rate=1000.0, duration=0.012, draw_key=25,
                            sizes=(100, 200, 400, 3200, 6400),
                            iters_range=(1000, 4000))
    sides = [s["L"] for s in stream]
    assert 6400 in sides and 3200 in sides, sides
    _case(f"serving burst of {len(stream)} requests, sides {sides}, 3 GiB")
    with Server(3 << 30, chunk=32, keep_rows=True) as srv:
        reqs, summary = srv.run(stream)
    assert summary["completed"] == len(stream), summary
    # the batch cannot hold every request at once: some waited for room
    assert any(r.admitted > min(x.done for x in reqs) for r in reqs)
    for r in reqs:
        d = r.params
        _case(f"request {r.rid}: {r.kind} {r.L} x {r.L}, {d['iters']} "
              "iterations, alone")
        with Simulation(Params.from_values(**d), seed=r.seed,
                        capacity=reserve_for(d), device=0) as sim:
            alone = stats_rows_to_array(sim.simulate())
        assert r.rows.shape == alone.shape, r.rid
        bad = np.nonzero((r.rows != alone).any(axis=1))[0]
        assert len(bad) == 0, (r.rid, bad[:3])
        if r.L <= 400:
            ref = simulate(d, seed=r.seed, capacity=reserve_for(d))
            assert np.array_equal(r.rows, ref), r.rid
