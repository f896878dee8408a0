"""Randomised parity: seeded random small configurations (grid shapes from
a single row to 2100 wide, sparse to dense, wolves or none, grass restart 1
to 300, every sort mode), 150 iterations each, every statistics row equal to
the oracle's; where the population outgrows the capacity, the failure at the
oracle's iteration with equal rows before it (pp_gpu.c:847-852)."""
import random
import sys

import numpy as np
import pytest

from helpers import gpu_params

pytestmark = pytest.mark.gpu

ITERS = 150


def _configs(n, seed=20261203):
    rng = random.Random(seed)
    for case in range(n):
        gx = rng.choice([7, 33, 64, 100, 257, 300, 513, 700, 2100])
        gy = rng.choice([1, 2, 9, 50, 100, 260])
        cells = gx * gy
        ds = rng.choice([0.02, 0.1, 0.25, 0.6])
        dw = rng.choice([0.0, 0.02, 0.12, 0.3])
        d = dict(init_sheep=int(cells * ds),
                 sheep_gain_from_food=rng.choice([1, 4, 9]),
                 sheep_reproduce_threshold=rng.choice([2, 5]),
                 sheep_reproduce_prob=rng.choice([4, 10, 40]),
                 init_wolves=int(cells * dw),
                 wolves_gain_from_food=rng.choice([5, 20]),
                 wolves_reproduce_threshold=rng.choice([2, 7]),
                 wolves_reproduce_prob=rng.choice([5, 20]),
                 grass_restart=rng.choice([1, 5, 10, 30, 300]),
                 grid_x=gx, grid_y=gy, grid_xy=cells, iters=ITERS)
        cap = max(16, 6 * (d["init_sheep"] + d["init_wolves"]))
        yield case, d, cap, rng.randrange(1 << 30), rng.choice([1, 1, 0, 8, 11])


@pytest.mark.parametrize("case,d,cap,seed,sort", list(_configs(24)),
                         ids=lambda v: None)
def test_random_configuration_matches_oracle(case, d, cap, seed, sort):
    from oracle.pp_oracle import Oracle
    from pphpc_amd import PPError, Simulation
    print(f"[case] random {case}: {d['grid_x']}x{d['grid_y']} sheep "
          f"{d['init_sheep']} wolves {d['init_wolves']} R {d['grass_restart']} "
          f"sort {sort} seed {seed}", flush=True)
    sys.stdout.flush()
    o = Oracle(d, seed=seed, capacity=cap)
    ref = [o.stats()]
    rc, rows = o.run(ITERS)
    fail = o.step_count
    o.close()
    ref = np.vstack([np.array(ref, np.uint64), rows]) if len(rows) else \
        np.array(ref, np.uint64)

    sim = Simulation(gpu_params(d), seed=seed, capacity=cap, device=0)
    sim.set_sort(sort)
    got = [sim.init().as_tuple()]
    if rc == 0:
        got += [r.as_tuple() for r in sim.run(ITERS)]
        sim.close()
        assert np.array_equal(np.array(got, np.uint64), ref)
        return
    with pytest.raises(PPError):
        sim.run(ITERS)
    assert sim.error_step == fail
    sim.close()
    sim = Simulation(gpu_params(d), seed=seed, capacity=cap, device=0)
    sim.set_sort(sort)
    got = [sim.init().as_tuple()] + (
        [r.as_tuple() for r in sim.run(fail)] if fail else [])
    sim.close()
    assert np.array_equal(np.array(got, np.uint64), ref[:fail + 1])
