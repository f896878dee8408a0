"""Batches of replicas (pp_batch_*): every replica's statistics rows and state
equal those of the same simulation alone (the CPU oracle, and pp_sim), bit
for bit, whatever else shares the batch, whenever it was admitted, and
whatever sort shape the batch uses.
"""
import sys

import numpy as np
import pytest

from helpers import assert_state_equal, gpu_params, v1, v2ref, v2s

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_rows(d, seed, capacity):
    from oracle.pp_oracle import simulate
    return simulate(d, seed=seed, capacity=capacity)


def _reserve(d):
    """Agent reservation: 4x the initial agents, at least three per cell (the
    v1 and reference-v2 sets outgrow 4x; v2 peaks near 2.2 per cell)."""
    return max(4 * (d["init_sheep"] + d["init_wolves"]),
               3 * d["grid_x"] * d["grid_y"])


def _check_rows(b, rid, d, seed, what):
    rows = b.rows_array(rid)
    ref = _oracle_rows(d, seed, _reserve(d))
    assert rows.shape == ref.shape, f"{what}: {rows.shape} != {ref.shape}"
    bad = np.nonzero((rows != ref).any(axis=1))[0]
    assert len(bad) == 0, (f"{what}: first differing row {bad[0]}: "
                           f"{rows[bad[0]]} != {ref[bad[0]]}")


def _batch(agents, cells, reps, R=15):
    from pphpc_amd import Batch
    return Batch(agents, cells, reps, R, device=0)


def test_batch_of_one_equals_plain():
    from pphpc_amd import Simulation
    _case("batch of one v2s 100x100 seed 3, 600 iterations")
    d = v2s(100, iters=600)
    with Simulation(gpu_params(d), seed=3, capacity=_reserve(d),
                    device=0) as sim:
        plain = np.array([r.as_tuple() for r in sim.simulate()], np.uint64)
    with _batch(_reserve(d), 100 * 100, 1) as b:
        rid = b.add(gpu_params(d), 3, _reserve(d))
        b.run(600)
        assert b.replica(rid)["state"] == 2
        assert np.array_equal(b.rows_array(rid), plain)


HETERO = [
    # (params, seed, admitted after tick)
    (v2s(64, iters=150), 11, 0),
    (v1(100, iters=200), 5, 0),
    (v2ref(40, iters=90), 2, 0),
    (dict(v2s(gx=50, gy=30), iters=170), 7, 0),
    (v1(8, iters=120), 1, 0),
    (v2s(gx=1, gy=37, iters=80), 9, 0),
    (v2s(128, iters=60), 4, 37),
    (v1(30, iters=140), 8, 37),
    (v2ref(24, iters=100), 6, 95),
]


@pytest.mark.parametrize("prefix,narrow", [(False, False), (True, False), (False, True)],
                         ids=["levels", "prefix", "narrow"])
def test_heterogeneous_batch_equals_alone(prefix, narrow, monkeypatch):
    """Replicas of different grids, parameter sets, seeds and iteration
    counts, admitted at ticks 0, 37 and 95, each equal to its run alone;
    the bucket sort reading its range bounds from the count levels, and from
    the bucket records (used by default from 6144 buckets, here for any
    layout); the act and emit in groups of 512 threads (their default for
    this few agents) and of 256."""
    if prefix:
        monkeypatch.setenv("PPHPC_BPREFIX_MIN", "0")
    if narrow:
        monkeypatch.setenv("PPHPC_WIDE_GROUPS", "0")
    _case(f"heterogeneous batch of 9 replicas, staggered admission, prefix={prefix}, "
          f"narrow groups={narrow}")
    cells = sum(d["grid_x"] * d["grid_y"] for d, _, _ in HETERO)
    agents = sum(_reserve(d) for d, _, _ in HETERO)
    with _batch(agents, cells, 16) as b:
        ids = {}
        tick = 0
        for at in sorted({a for _, _, a in HETERO}):
            if at > tick:
                b.run(at - tick)
                tick = at
            for i, (d, seed, a) in enumerate(HETERO):
                if a == at:
                    rid = b.add(gpu_params(d), seed, _reserve(d))
                    assert rid is not None
                    ids[i] = rid
        b.run(1000)
        for i, (d, seed, a) in enumerate(HETERO):
            st = b.replica(ids[i])
            assert st["state"] == 2 and st["steps"] == d["iters"], (i, st)
            _check_rows(b, ids[i], d, seed, f"replica {i}")
        assert b.resources()["active"] == 0


def test_many_tiny_replicas_share_tiles():
    """40 replicas of 8x8 cells: one action tile holds the agents of more
    replicas than its LDS slots (the global-atomic path)."""
    _case("40 tiny replicas in one action tile")
    reps = [(dict(v2s(8), iters=100), 100 + i) for i in range(40)]
    with _batch(sum(_reserve(d) for d, _ in reps), 40 * 64, 64) as b:
        ids = [b.add(gpu_params(d), s, _reserve(d)) for d, s in reps]
        b.run(100)
        for rid, (d, s) in zip(ids, reps):
            _check_rows(b, rid, d, s, f"tiny replica seed {s}")


def test_running_replica_state_equals_oracle():
    from oracle.pp_oracle import Oracle
    _case("state of a running replica at tick 25 of a batch")
    ds = [(v2s(48, iters=500), 21), (v1(64, iters=500), 22),
          (v2s(gx=20, gy=70, iters=500), 23)]
    cells = sum(d["grid_x"] * d["grid_y"] for d, _ in ds)
    with _batch(sum(_reserve(d) for d, _ in ds), cells, 4) as b:
        ids = [b.add(gpu_params(d), s, _reserve(d)) for d, s in ds]
        b.run(25)
        for rid, (d, s) in zip(ids, ds):
            o = Oracle(d, seed=s, capacity=_reserve(d))
            assert o.run(25)[0] == 0
            assert_state_equal(b.state(rid), o.state(), f"seed {s} tick 25")
            o.close()


def test_capacity_failure_is_isolated():
    """A replica above its reservation fails alone, at the iteration the
    oracle with that capacity fails; its neighbours finish normally."""
    from oracle.pp_oracle import Oracle
    _case("replica capacity failure inside a batch")
    bad = dict(v2s(32, iters=300), sheep_reproduce_prob=60,
               wolves_reproduce_prob=40)
    n0 = bad["init_sheep"] + bad["init_wolves"]
    small = n0 + n0 // 4
    good = [(v2s(40, iters=120), 31), (v1(50, iters=120), 32)]
    cells = 32 * 32 + sum(d["grid_x"] * d["grid_y"] for d, _ in good)
    with _batch(small + sum(_reserve(d) for d, _ in good), cells, 4) as b:
        rb = b.add(gpu_params(bad), 30, small)
        ids = [b.add(gpu_params(d), s, _reserve(d)) for d, s in good]
        b.run(400)
        o = Oracle(bad, seed=30, capacity=small)
        rc, orows = o.run(300)
        assert rc == -8
        st = b.replica(rb)
        assert st["state"] == 3 and st["error"] == -8, st
        assert st["steps"] == o.step_count, (st, o.step_count)
        rows = b.rows_array(rb)
        assert np.array_equal(rows[1:], orows[:o.step_count])
        o.close()
        for rid, (d, s) in zip(ids, good):
            _check_rows(b, rid, d, s, f"neighbour seed {s}")


def test_shapes_and_tuner_do_not_change_results():
    _case("sort shapes and tuner")
    ds = [(v2s(150, iters=260), 41), (v1(90, iters=260), 42),
          (v2ref(70, iters=260), 43)]
    cells = sum(d["grid_x"] * d["grid_y"] for d, _ in ds)
    agents = sum(_reserve(d) for d, _ in ds)
    with _batch(agents, cells, 4) as b:
        ids = [b.add(gpu_params(d), s, _reserve(d)) for d, s in ds]
        b.set_shape(8, 8)
        b.run(40)
        b.set_shape(16, 11)
        b.run(30)
        # launch shapes: one round of action and emit groups, half a round
        # of bucket-sort groups; then eight rounds
        b.set_launch(1, 4)
        b.run(30)
        b.set_launch(8, 8)
        b.run(30)
        t = b.tune(16)
        assert t["keys_per_thread"] in (8, 16)
        assert t["max_digit_bits"] in (8, 11)
        assert t["tile_rounds"] in (1, 2, 4, 8)
        assert t["bucket_eighths"] in (4, 8)
        b.run(1000)
        for rid, (d, s) in zip(ids, ds):
            _check_rows(b, rid, d, s, f"seed {s} after tuning")
        with pytest.raises(Exception):
            b.set_launch(3, 8)


def test_admission_waits_for_room():
    """Slots, cells and reservations come back when replicas finish."""
    _case("admission by resources")
    d = v2s(40, iters=30)
    with _batch(2 * _reserve(d), 2 * 1600, 2) as b:
        a = b.add(gpu_params(d), 1, _reserve(d))
        c = b.add(gpu_params(d), 2, _reserve(d))
        assert a is not None and c is not None
        assert b.add(gpu_params(d), 3, _reserve(d)) is None
        b.run(30)
        r = b.resources()
        assert r["slots_free"] == 2 and r["active"] == 0
        e = b.add(gpu_params(d), 3, _reserve(d))
        assert e is not None
        b.run(30)
        for rid, s in ((a, 1), (c, 2), (e, 3)):
            _check_rows(b, rid, d, s, f"seed {s}")
        b.release(a)
        with pytest.raises(Exception):
            b.replica(a)
