"""Shared parameter sets and comparison helpers for the tests."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIGS = os.path.join(ROOT, "configs")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def v1(L=100, iters=4000, gx=None, gy=None):
    """Reference v1 parameters (configs/config100v1.txt values) scaled to the
    reference densities: 400 sheep, 200 wolves per 100x100 cells."""
    gx = gx or L
    gy = gy or L
    cells = gx * gy
    return dict(init_sheep=cells * 400 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
                init_wolves=cells * 200 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=10, grid_x=gx, grid_y=gy, grid_xy=cells,
                iters=iters)


def v2s(L=1600, iters=4000, gx=None, gy=None):
    """The dense v2-style set of the 1600x1600 benchmark: 1600 sheep and 800
    wolves per 100x100 cells, sheep gain 4 / prob 10, wolf gain 20 / prob 5,
    thresholds 2, grass restart 5."""
    gx = gx or L
    gy = gy or L
    cells = gx * gy
    return dict(init_sheep=cells * 1600 // 10000, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=10,
                init_wolves=cells * 800 // 10000, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=5, grid_x=gx, grid_y=gy, grid_xy=cells,
                iters=iters)


def v2ref(L=100, iters=4000):
    """Reference v2 parameters (configs/config100v2.txt values)."""
    cells = L * L
    return dict(init_sheep=cells * 400 // 10000, sheep_gain_from_food=30,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=10,
                init_wolves=cells * 200 // 10000, wolves_gain_from_food=10,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
                grass_restart=15, grid_x=L, grid_y=L, grid_xy=cells,
                iters=iters)


def gpu_params(d):
    from pphpc_amd import Params
    return Params.from_values(**d)


def assert_state_equal(a, b, what=""):
    names = ("cell", "energy", "type", "grass")
    for name, x, y in zip(names, a, b):
        assert x.shape == y.shape, f"{what} {name}: shape {x.shape} != {y.shape}"
        if not np.array_equal(x, y):
            bad = np.nonzero(x != y)[0]
            raise AssertionError(
                f"{what} {name}: {len(bad)} mismatches, first at {bad[0]}: "
                f"{x[bad[0]]} != {y[bad[0]]}")
