"""Serving loop: requests arriving while others run are admitted as room
frees up, run concurrently in one batch, and each returns exactly the
statistics of the same simulation run alone."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_each_result_same_alone_and_in_the_mix():
    from pphpc_amd import Params, Simulation, stats_rows_to_array
    from pphpc_amd.serve import Server
    from pphpc_amd.workloads import reserve_for, serve_requests
    _case("serving a burst of requests, sides 50..400, then each alone")
    stream = serve_requests(rate=2000.0, duration=0.01, draw_key=11,
                            sizes=(50, 100, 200, 400),
                            iters_range=(100, 300))
    assert len(stream) >= 12
    # four replica slots: the burst runs four at a time, the rest wait for
    # slots to free up
    with Server(24 << 20, max_replicas=4, chunk=16, keep_rows=True) as srv:
        reqs, summary = srv.run(stream)
    assert summary["completed"] == len(stream), summary
    assert summary["max_concurrent"] == 4
    first_done = min(r.done for r in reqs)
    assert sum(r.admitted >= first_done for r in reqs) >= len(reqs) - 4
    for r in reqs:
        d = r.params
        with Simulation(Params.from_values(**d), seed=r.seed,
                        capacity=reserve_for(d), device=0) as sim:
            alone = stats_rows_to_array(sim.simulate())
        assert r.rows.shape == alone.shape
        bad = np.nonzero((r.rows != alone).any(axis=1))[0]
        assert len(bad) == 0, (r.rid, bad[:3])
