"""Limits of the step, each checked bit-exactly against the oracle:

* the energy limit: an agent above 2^24 - 1 after its action fails the
  simulation (in a batch, only its replica) with PP_ENERGY_OVERFLOW at the
  iteration the oracle fails, and every row before it is the oracle's;
* agent capacities above 33.5 M: one replica (the emit sums the block sums
  before its tiles) and a batch of two, where the end-of-tick scan of the
  action tiles' block sums takes its 1024-thread form;
* the full 12800 x 12800 grid (163.84 M cells, 39.3 M agents): the state
  and the statistics after two iterations.
"""
import sys

import numpy as np
import pytest

from helpers import assert_state_equal, gpu_params, v1, v2s

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _energy_case(side=16, iters=600):
    """Wolves gaining 2^20 per sheep, never reproducing, among sheep that
    breed fast: some wolf passes 2^24 - 1 within a few hundred iterations."""
    return dict(v2s(side, iters=iters), wolves_gain_from_food=1 << 20,
                wolves_reproduce_prob=0, sheep_reproduce_prob=60)


def _oracle_failure(d, seed, capacity):
    from oracle.pp_oracle import Oracle
    o = Oracle(d, seed=seed, capacity=capacity)
    row0 = np.array([o.stats()], np.uint64)
    rc, rows = o.run(d["iters"])
    step = o.step_count
    o.close()
    return rc, step, np.vstack([row0, rows[:step]])


def test_energy_overflow_plain_matches_oracle():
    from pphpc_amd import PP_ENERGY_OVERFLOW, PPError, Simulation
    d = _energy_case()
    _case("energy limit, plain 16x16, seed 3")
    cap = 4 * (d["init_sheep"] + d["init_wolves"])
    rc, fail, ref = _oracle_failure(d, 3, cap)
    assert rc == PP_ENERGY_OVERFLOW and 10 < fail < d["iters"]
    with Simulation(gpu_params(d), seed=3, capacity=cap, device=0) as sim:
        row0 = sim.init().as_tuple()
        with pytest.raises(PPError) as e:
            sim.run(d["iters"])
        assert e.value.code == PP_ENERGY_OVERFLOW
        assert sim.error_step == fail and sim.step_count == fail
    with Simulation(gpu_params(d), seed=3, capacity=cap, device=0) as sim:
        got = np.array([sim.init().as_tuple()]
                       + [r.as_tuple() for r in sim.run(fail)], np.uint64)
    assert row0 == tuple(int(x) for x in ref[0])
    assert np.array_equal(got, ref)


def test_energy_overflow_is_isolated_in_a_batch():
    """The overflowing replica fails alone; its neighbours, including one
    sharing action tiles with it, finish with the oracle's rows."""
    from pphpc_amd import PP_ENERGY_OVERFLOW, Batch
    from oracle.pp_oracle import simulate
    bad = _energy_case(16)
    good = [(v2s(24, iters=500), 21), (v1(40, iters=500), 22)]
    _case("energy limit inside a batch of three")
    res = lambda d: max(4 * (d["init_sheep"] + d["init_wolves"]),
                        3 * d["grid_x"] * d["grid_y"])
    cells = 16 * 16 + sum(d["grid_x"] * d["grid_y"] for d, _ in good)
    agents = res(bad) + sum(res(d) for d, _ in good)
    rc, fail, ref = _oracle_failure(bad, 3, res(bad))
    assert rc == PP_ENERGY_OVERFLOW
    with Batch(agents, cells, 4, 10, device=0) as b:
        ids = [b.add(gpu_params(good[0][0]), good[0][1], res(good[0][0])),
               b.add(gpu_params(bad), 3, res(bad)),
               b.add(gpu_params(good[1][0]), good[1][1], res(good[1][0]))]
        b.run(600)
        st = b.replica(ids[1])
        assert st["state"] == 3 and st["error"] == PP_ENERGY_OVERFLOW, st
        assert st["steps"] == fail
        assert np.array_equal(b.rows_array(ids[1]), ref)
        for rid, (d, s) in zip((ids[0], ids[2]), good):
            assert b.replica(rid)["state"] == 2
            assert np.array_equal(b.rows_array(rid),
                                  simulate(d, seed=s, capacity=res(d)))


@pytest.mark.parametrize("capacity", [34_000_000])
def test_wide_tile_scan_matches_oracle(capacity):
    """A small grid in buffers of more than 32768 blocks of action tiles
    (33.5 M agents) runs the oracle's rows and state. (One replica: the
    emit's groups sum the block sums before their tiles; the scan kernel
    is not launched.)"""
    from pphpc_amd import Simulation
    from oracle.pp_oracle import Oracle
    d = v2s(100, iters=300)
    _case(f"v2s 100x100 with capacity {capacity}")
    with Simulation(gpu_params(d), seed=12, capacity=capacity,
                    device=0) as sim:
        rows = np.array([r.as_tuple() for r in sim.simulate()], np.uint64)
        state = sim.state()
    o = Oracle(d, seed=12, capacity=capacity)
    ref = [o.stats()]
    rc, more = o.run(300)
    assert rc == 0
    assert np.array_equal(rows, np.vstack([np.array(ref, np.uint64), more]))
    assert_state_equal(state, o.state(), "final state")
    o.close()


def test_wide_tile_scan_of_a_batch_matches_oracle():
    """Two replicas in a batch of 34 M agents (above 32768 blocks of action
    tiles): the end-of-tick kernel scans the block sums with 1024 threads
    instead of one wave; both replicas run the oracle's rows."""
    from pphpc_amd import Batch
    from pphpc_amd.workloads import reserve_for as res
    from oracle.pp_oracle import simulate
    reps = [(v2s(100, iters=300), 12), (v1(64, iters=300), 5)]
    cells = sum(d["grid_x"] * d["grid_y"] for d, _ in reps)
    _case("two replicas in a batch of 34 M agents")
    with Batch(34_000_000, cells, 2, 10, device=0) as b:
        ids = [b.add(gpu_params(d), s, res(d)) for d, s in reps]
        b.run(300)
        for rid, (d, s) in zip(ids, reps):
            st = b.replica(rid)
            assert st["state"] == 2 and st["steps"] == 300, st
            assert np.array_equal(b.rows_array(rid),
                                  simulate(d, seed=s, capacity=res(d)))


def test_full_size_12800_equals_oracle():
    """The 12800 x 12800 workload (26.2 M sheep, 13.1 M wolves, capacity
    4 x N0 on the GPU; the oracle needs no more than 2 x N0) bit for bit:
    rows 0..2 and the whole state after two iterations."""
    from pphpc_amd import Simulation
    from oracle.pp_oracle import Oracle
    d = v2s(12800, iters=2)
    n0 = d["init_sheep"] + d["init_wolves"]
    _case("12800^2 seed 0, two iterations, against the oracle")
    with Simulation(gpu_params(d), seed=0, capacity=4 * n0, device=0) as sim:
        rows = np.array([sim.init().as_tuple()]
                        + [r.as_tuple() for r in sim.run(2)], np.uint64)
        state = sim.state()
    _case("12800^2 oracle")
    o = Oracle(d, seed=0, capacity=2 * n0)
    ref = [o.stats()]
    rc, more = o.run(2)
    assert rc == 0
    assert np.array_equal(rows, np.vstack([np.array(ref, np.uint64), more]))
    assert_state_equal(state, o.state(), "12800^2 after 2 iterations")
    o.close()
