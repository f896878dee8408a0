"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Bit-exact: every agent (cell, energy, type) in slot order, every grass
countdown, and every per-iteration statistic.
"""
import sys

import numpy as np
import pytest

from helpers import assert_state_equal, gpu_params, v1, v2ref, v2s

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(d, seed=0, capacity=0):
    from oracle.pp_oracle import Oracle
    from pphpc_amd import Simulation
    sim = Simulation(gpu_params(d), seed=seed, capacity=capacity, device=0)
    orc = Oracle(d, seed=seed, capacity=sim.capacity)
    return sim, orc


STATE_CASES = [
    ("v1 8x8", v1(gx=8, gy=8), 3),
    ("v2s 16x8", v2s(gx=16, gy=8), 11),
    ("v2s 1x37", v2s(gx=1, gy=37), 5),
    ("v2s 41x1", v2s(gx=41, gy=1), 7),
    ("v1 dense 5x3", dict(v1(gx=5, gy=3), init_sheep=40, init_wolves=12), 2),
    ("v2ref 33x29", dict(v2ref(33), grid_x=33, grid_y=29, init_sheep=200,
                         init_wolves=90), 9),
    ("v2s 64x64", v2s(64), 0),
]


@pytest.mark.parametrize("name,d,seed", STATE_CASES, ids=[c[0] for c in
                                                         STATE_CASES])
def test_state_parity_per_step(name, d, seed):
    """Full state after init and after each of 60 iterations."""
    _case(f"state parity {name} seed={seed}")
    sim, orc = _pair(d, seed)
    s0 = sim.init()
    assert s0.as_tuple() == orc.stats(), "row 0"
    assert_state_equal(sim.state(), orc.state(), f"{name} init")
    for t in range(60):
        g = sim.step()
        assert orc.step() == 0
        assert g.as_tuple() == orc.stats(), f"{name} stats row {t + 1}"
        assert_state_equal(sim.state(), orc.state(), f"{name} step {t + 1}")
    sim.close()


@pytest.mark.parametrize("name,d,seed,iters", [
    ("config1 v1 100x100", v1(100), 0, 4000),
    ("v2s 100x100", v2s(100), 1, 4000),
    ("v2ref 200x200", v2ref(200), 7, 1500),
    ("v2s 400x400", v2s(400), 0, 400),
])
def test_stats_parity_long(name, d, seed, iters):
    """Every statistics row of a long run (graph-captured chunks)."""
    from oracle.pp_oracle import simulate
    _case(f"stats parity {name} seed={seed} iters={iters}")
    sim, orc = _pair(d, seed)
    orc.close()
    rows = np.array([r.as_tuple() for r in [sim.init()] + sim.run(iters)],
                    dtype=np.uint64)
    ref = simulate(d, seed=seed, capacity=sim.capacity, iters=iters)
    bad = np.nonzero((rows != ref).any(axis=1))[0]
    assert len(bad) == 0, (f"{name}: first differing row {bad[0]}: "
                           f"{rows[bad[0]]} != {ref[bad[0]]}")
    sim.close()


def test_event_counters_match_oracle():
    """The run's exact event sums (the counts the roofline's algorithmic
    bytes are computed from) against the oracle's per-step events and
    state: agents at the start of each iteration, survivors of the move,
    agents after actions, grass eaten, births."""
    from oracle.pp_oracle import Oracle
    from pphpc_amd import Simulation
    d = v2s(200)
    _case("event counters v2s 200x200, 150 iterations")
    sim = Simulation(gpu_params(d), seed=4, capacity=0, device=0)
    orc = Oracle(d, seed=4, capacity=sim.capacity)
    sim.init()
    sim.run(150, collect=False)
    c = sim.counters()
    sim.close()
    want = dict(agent_steps=0, live=0, out=0, eaten=0, births=0)
    for _ in range(150):
        n = orc.num_agents
        assert orc.step() == 0
        ev = orc.last_events()
        want["agent_steps"] += n
        want["live"] += n - ev["starved"]
        want["out"] += orc.num_agents
        want["eaten"] += ev["grass_eaten"]
        want["births"] += ev["sheep_births"] + ev["wolf_births"]
    orc.close()
    assert {k: c[k] for k in want} == want
    assert 0 < c["occupied"] <= c["live"]
