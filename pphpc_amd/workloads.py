"""Workloads of the benchmark and the tests: the parameter sets, the draw of a
batch of heterogeneous replicas and the request stream of the serving loop.

Parameter sets (per 100 x 100 cells, scaled with the grid):
  v1   the reference's configs/config100v1.txt: 400 sheep, 200 wolves, sheep
       gain 4 / threshold 2 / probability 4, wolf gain 20 / 2 / 5, grass
       restart 10
  v2s  the dense v2-style set of the 1600 x 1600 benchmark: 1600 sheep, 800
       wolves, sheep gain 4 / 2 / 10, wolf gain 20 / 2 / 5, grass restart 5

Random draws of workloads (grid sizes, parameter sets, seeds, arrival times)
use numpy's Philox4x32-10 generator under a fixed key, the generator family
of the step itself.
"""
from __future__ import annotations

import numpy as np

BATCH_SIZES = (100, 200, 400, 800, 1600, 3200)
SERVE_SIZES = (100, 200, 400, 800, 1600, 3200, 6400)
KINDS = ("v1", "v2s")


def v1(L: int = 100, iters: int = 4000, gx: int | None = None,
       gy: int | None = None) -> dict:
    gx = gx or L
    gy = gy or L
    cells = gx * gy
    return dict(init_sheep=cells * 400 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
                init_wolves=cells * 200 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=10, grid_x=gx, grid_y=gy, grid_xy=cells,
                iters=iters)


def v2s(L: int = 1600, iters: int = 4000, gx: int | None = None,
        gy: int | None = None) -> dict:
    gx = gx or L
    gy = gy or L
    cells = gx * gy
    return dict(init_sheep=cells * 1600 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=10,
                init_wolves=cells * 800 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=5, grid_x=gx, grid_y=gy, grid_xy=cells,
                iters=iters)


def params_of(kind: str, L: int, iters: int) -> dict:
    return v1(L, iters) if kind == "v1" else v2s(L, iters)


def reserve_for(d: dict) -> int:
    """Agent reservation of a replica: 4x its initial agents and at least one
    agent per cell (v1 populations outgrow 4x, peaking near 0.3 agents per
    cell; v2s stays under 0.5)."""
    return max(4 * (d["init_sheep"] + d["init_wolves"]),
               d["grid_x"] * d["grid_y"])


def footprint_bytes(d: dict) -> int:
    """Device bytes a replica holds in a batch: 33 per reserved agent (state
    9, sort pairs 24), 1 per cell of grass (restart <= 191), its grass ring
    and statistics ring."""
    return (33 * reserve_for(d) + d["grid_x"] * d["grid_y"]
            + 8 * (d["grass_restart"] + 1) + 56 * 1024)


def _rng(key: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=key))


def batch_replicas(n: int = 64, draw_key: int = 4, iters: int = 4000):
    """The batched-replica workload: n replicas, grid side uniform over
    BATCH_SIZES, v1 or v2s with probability 1/2 each, seeds 1..n. Returns
    [(kind, L, params dict, seed)]."""
    g = _rng(draw_key)
    sides = g.choice(BATCH_SIZES, size=n)
    kinds = g.integers(0, 2, size=n)
    return [(KINDS[k], int(L), params_of(KINDS[k], int(L), iters), i + 1)
            for i, (L, k) in enumerate(zip(sides, kinds))]


def serve_requests(rate: float, duration: float, draw_key: int = 5,
                   sizes=SERVE_SIZES, iters_range=(1000, 4000)):
    """A Poisson stream of simulation requests over [0, duration) seconds at
    `rate` per second: grid side uniform over `sizes` (log-uniform over
    powers of two), v1 or v2s with probability 1/2, iterations uniform over
    iters_range (inclusive), a random 64-bit seed. Returns a list of dicts
    with arrival, kind, L, params and seed, in arrival order."""
    g = _rng(draw_key)
    out = []
    t = 0.0
    while True:
        t += g.exponential(1.0 / rate)
        if t >= duration:
            break
        L = int(g.choice(sizes))
        kind = KINDS[int(g.integers(0, 2))]
        iters = int(g.integers(iters_range[0], iters_range[1] + 1))
        seed = int(g.integers(0, 2 ** 63))
        out.append({"arrival": t, "kind": kind, "L": L,
                    "params": params_of(kind, L, iters), "seed": seed})
    return out
