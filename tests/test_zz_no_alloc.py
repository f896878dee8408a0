"""The allocation-free entry point (pp_sim_simulate_noalloc) gives exactly the
rows of pp_sim_create + pp_sim_simulate, call after call on the same device
scratch: a larger run, then smaller ones (whose buffers sit over what the
first left behind), then the first again. Collected last (the name), so
that it also runs after every other GPU test in one process."""
import sys

import numpy as np
import pytest

from helpers import gpu_params, v1, v2s

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _plain(d, seed, capacity):
    from pphpc_amd import Simulation, stats_rows_to_array
    with Simulation(gpu_params(d), seed=seed, capacity=capacity,
                    device=0) as sim:
        return stats_rows_to_array(sim.simulate())


CALLS = [
    # (params, seed, capacity): large, then smaller, then large again
    (v2s(256, iters=400), 5, 4 * (10485 + 5242)),
    (v1(100, iters=300), 0, 30000),   # v1 peaks near 4.5x its start
    (v2s(64, iters=200), 3, 0),       # capacity 0: 4x the initial agents
    (dict(v2s(gx=90, gy=40), iters=250), 9, 4000),
    (v2s(256, iters=400), 5, 4 * (10485 + 5242)),
]


def test_noalloc_equals_simulate_call_after_call():
    from pphpc_amd import simulate_noalloc
    for i, (d, seed, cap) in enumerate(CALLS):
        _case(f"noalloc call {i}: {d['grid_x']}x{d['grid_y']} seed {seed} "
              f"capacity {cap}")
        got = simulate_noalloc(gpu_params(d), seed, cap, device=0)
        ref = _plain(d, seed, cap or 4 * (d["init_sheep"] + d["init_wolves"]))
        assert got.shape == ref.shape
        bad = np.nonzero((got != ref).any(axis=1))[0]
        assert len(bad) == 0, (i, int(bad[0]), got[bad[0]], ref[bad[0]])


def test_noalloc_at_its_limits():
    """A run at every limit at once (262,144 cells, 262,144 agents of capacity,
    grass restart 4096) fits the scratch and equals the plain run."""
    from pphpc_amd import simulate_noalloc
    d = dict(v2s(512, iters=40), grass_restart=4096)
    _case("noalloc at its limits: 512x512, capacity 262144, grass restart 4096")
    got = simulate_noalloc(gpu_params(d), 11, 1 << 18, device=0)
    ref = _plain(d, 11, 1 << 18)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from pphpc_amd import PP_INVALID_ARGS, PPError, simulate_noalloc
    _case("noalloc limits")
    with pytest.raises(PPError) as e:
        simulate_noalloc(gpu_params(v2s(600, iters=10)), 1, 0, device=0)
    assert e.value.code == PP_INVALID_ARGS
    with pytest.raises(PPError) as e:
        simulate_noalloc(gpu_params(v1(100, iters=10)), 1, 1 << 20, device=0)
    assert e.value.code == PP_INVALID_ARGS
