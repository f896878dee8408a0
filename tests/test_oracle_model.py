"""Pins the oracle's step semantics on the CPU.

* an independent Python statement of the model (tests/spec_model.py) agrees
  with the oracle state after every step on small grids;
* hand-worked single-cell cases with outcomes fixed by the rules;
* distributions of init, move and shuffle (chi-square);
* closed-form invariants of counts and countdown sums from event counts;
* planted mistakes (oracle fault modes) are caught;
* golden fixtures (regression pins) are reproduced.
"""
import hashlib
import json
import os

import numpy as np
import pytest
from scipy import stats as sps

from helpers import GOLDEN, v1, v2ref, v2s
from oracle.pp_oracle import Oracle, simulate
from spec_model import SpecModel


def _state_lists(o):
    c, e, t, g = o.state()
    return list(map(int, c)), list(map(int, e)), list(map(int, t)), \
        list(map(int, g))


SPEC_CASES = [
    ("v1 8x8", v1(gx=8, gy=8), 3, 40),
    ("v2s 12x7", v2s(gx=12, gy=7), 5, 40),
    ("v2ref 9x9", dict(v2ref(9), init_sheep=30, init_wolves=12), 11, 40),
    ("dense 4x3", dict(v1(gx=4, gy=3), init_sheep=30, init_wolves=9), 2, 30),
    ("1x9 torus", dict(v2s(gx=1, gy=9), init_sheep=6, init_wolves=3), 4, 30),
    ("seed high bits", v2s(gx=10, gy=10), (7 << 32) | 5, 20),
]


@pytest.mark.parametrize("name,d,seed,iters", SPEC_CASES,
                         ids=[c[0] for c in SPEC_CASES])
def test_oracle_matches_independent_spec(name, d, seed, iters):
    o = Oracle(d, seed=seed)
    m = SpecModel(d, seed)
    assert _state_lists(o) == m.state(), f"{name}: init state"
    assert o.stats() == m.stats()
    for t in range(iters):
        assert o.step() == 0
        m.step()
        assert _state_lists(o) == m.state(), f"{name}: state after {t + 1}"
        assert o.stats() == m.stats(), f"{name}: stats after {t + 1}"
    o.close()


# ------------------------------------------------------- hand-worked cases

def _one_cell(**kw):
    base = dict(init_sheep=0, sheep_gain_from_food=4,
                sheep_reproduce_threshold=2, sheep_reproduce_prob=0,
                init_wolves=0, wolves_gain_from_food=20,
                wolves_reproduce_threshold=2, wolves_reproduce_prob=0,
                grass_restart=10, grid_x=1, grid_y=1, iters=10)
    base.update(kw)
    return base


def _run_one(d, agents, grass, seed=0):
    """agents: list of (type, energy) in slot order, all in cell 0."""
    o = Oracle(d, seed=seed, capacity=64)
    n = len(agents)
    o.set_state(np.zeros(n), [e for _, e in agents], [t for t, _ in agents],
                [grass])
    assert o.step() == 0
    c, e, t, g = o.state()
    ev = o.last_events()
    o.close()
    return list(zip(map(int, t), map(int, e))), int(g[0]), ev


def test_sheep_eats_live_grass():
    out, g, ev = _run_one(_one_cell(), [(0, 5)], grass=0)
    assert out == [(0, 5 - 1 + 4)] and g == 10 and ev["grass_eaten"] == 1


def test_sheep_finds_dead_grass():
    out, g, ev = _run_one(_one_cell(), [(0, 5)], grass=3)
    assert out == [(0, 4)] and g == 2 and ev["grass_eaten"] == 0


def test_grass_regrows_at_countdown_one():
    """countdown 1 reaches 0 in the regrowth phase, before actions."""
    out, g, _ = _run_one(_one_cell(), [(0, 5)], grass=1)
    assert out == [(0, 8)] and g == 10


def test_starving_agent_dies_before_acting():
    out, g, ev = _run_one(_one_cell(), [(0, 1)], grass=0)
    assert out == [] and g == 0 and ev["starved"] == 1


def test_reproduction_splits_energy():
    d = _one_cell(sheep_reproduce_prob=100)
    out, g, ev = _run_one(d, [(0, 6)], grass=0)
    # 6 - 1 + 4 = 9 > 2: child floor(9/2) = 4 right after its parent (5)
    assert out == [(0, 5), (0, 4)] and ev["sheep_births"] == 1


def test_reproduction_needs_energy_above_threshold():
    d = _one_cell(sheep_reproduce_prob=100, sheep_reproduce_threshold=7)
    out, _, _ = _run_one(d, [(0, 5)], grass=3)     # 4, not above 7
    assert out == [(0, 4)]


def test_wolf_eats_the_only_sheep():
    out, g, ev = _run_one(_one_cell(), [(0, 5), (1, 10)], grass=5)
    assert out == [(1, 10 - 1 + 20)] and ev["sheep_eaten"] == 1


def test_wolf_eats_one_of_two_sheep():
    for seed in range(20):
        out, _, ev = _run_one(_one_cell(), [(0, 5), (0, 7), (1, 10)], 5, seed)
        sheep = [e for t, e in out if t == 0]
        wolves = [e for t, e in out if t == 1]
        assert wolves == [29] and len(sheep) == 1 and sheep[0] in (4, 6)
        assert ev["sheep_eaten"] == 1


def test_two_wolves_one_sheep():
    for seed in range(20):
        out, _, _ = _run_one(_one_cell(), [(1, 10), (0, 5), (1, 30)], 5, seed)
        wolves = sorted(e for t, e in out if t == 1)
        assert [t for t, _ in out].count(0) == 0
        assert wolves in ([9, 49], [29, 29])


def test_newborn_is_never_eaten():
    """Sheep reproduces only if it acts before the wolf; the wolf then eats
    the parent, never the newborn. Outcomes: {child only} or {nothing}."""
    d = _one_cell(sheep_reproduce_prob=100)
    seen = set()
    for seed in range(40):
        out, _, _ = _run_one(d, [(0, 9), (1, 10)], grass=5, seed=seed)
        sheep = [e for t, e in out if t == 0]
        assert sheep in ([], [4]), sheep   # 9-1=8: child 4, parent 4 eaten
        seen.add(len(sheep))
    assert seen == {0, 1}, "both acting orders should occur over 40 seeds"


def test_eaten_sheep_that_already_acted_keeps_its_grass_effect():
    """Sheep acting first eats the grass; then the wolf eats it: grass is
    reset to R even though the sheep is gone."""
    outs = set()
    for seed in range(40):
        out, g, _ = _run_one(_one_cell(), [(0, 5), (1, 10)], grass=0,
                             seed=seed)
        assert out == [(1, 29)]
        outs.add(g)
    assert outs == {0, 10}


# ----------------------------------------------------------- distributions

def test_init_distributions():
    d = v1(200)          # 40000 cells, 1600 sheep, 800 wolves
    d["init_sheep"], d["init_wolves"] = 20000, 20000
    o = Oracle(d, seed=9)
    cell, energy, typ, grass = o.state()
    o.close()
    alive = int((grass == 0).sum())
    assert abs(alive - 20000) < 5 * np.sqrt(10000)
    cd = np.bincount(grass[grass > 0], minlength=11)[1:]
    assert sps.chisquare(cd).pvalue > 1e-4                 # U{1..10}
    es = np.bincount(energy[typ == 0], minlength=9)[1:]
    assert sps.chisquare(es).pvalue > 1e-4                 # U{1..8}
    ew = np.bincount(energy[typ == 1], minlength=41)[1:]
    assert sps.chisquare(ew).pvalue > 1e-4                 # U{1..40}
    assert sps.chisquare(np.bincount(cell % 200, minlength=200)).pvalue > 1e-4
    assert sps.chisquare(np.bincount(cell // 200, minlength=200)).pvalue > 1e-4
    assert list(typ[:20000]) == [0] * 20000 and set(typ[20000:]) == {1}


def test_moves_stay_within_one_step():
    """Every agent ends within one torus step of where some agent started,
    and a population that cannot die or breed is conserved."""
    L = 50
    d = dict(v1(L), init_sheep=0, init_wolves=0, wolves_reproduce_prob=0)
    n = 20000
    cells = np.random.default_rng(0).integers(0, L * L, n)
    o = Oracle(d, seed=3, capacity=n)
    o.set_state(cells, np.full(n, 50), np.ones(n, np.uint8),
                np.full(L * L, 5))
    assert o.step() == 0
    c, e, _, _ = o.state()
    o.close()
    assert len(c) == n and (e == 49).all()
    occ = (np.bincount(cells, minlength=L * L) > 0).reshape(L, L)
    reach = occ | np.roll(occ, 1, 0) | np.roll(occ, -1, 0) | \
        np.roll(occ, 1, 1) | np.roll(occ, -1, 1)
    assert reach[c // L, c % L].all()


def test_move_direction_table_single_agents():
    """One agent per far-apart cell: the direction drawn by word 0 of block
    (id, step, MOVE, 0) maps 1:+x 2:-x 3:+y 4:-y with wrap."""
    from oracle.pp_oracle import philox
    L = 7
    d = dict(v1(L), init_sheep=0, init_wolves=0, wolves_reproduce_prob=0)
    counts = np.zeros(5, int)
    for seed in range(200):
        o = Oracle(d, seed=seed, capacity=4)
        start = [0, 6, 42, 48]                     # the four corners
        o.set_state(start, [9, 9, 9, 9], [1, 1, 1, 1], np.full(L * L, 5))
        assert o.step() == 0
        c, _, _, _ = o.state()
        o.close()
        got = sorted(int(x) for x in c)
        want = []
        for i, s in enumerate(start):
            dirn = (philox((i, 0, 2, 0), (seed, 0))[0] * 5) >> 32
            counts[dirn] += 1
            x, y = s % L, s // L
            dx, dy = {0: (0, 0), 1: (1, 0), 2: (-1, 0), 3: (0, 1),
                      4: (0, -1)}[dirn]
            want.append(((y + dy) % L) * L + (x + dx) % L)
        assert got == sorted(want), seed
    assert sps.chisquare(counts).pvalue > 1e-4


def test_shuffle_is_uniform():
    """Three sheep on a 1x1 torus with dead grass and no reproduction: the
    state lists them in acting order, so each step's permutation of the
    previous order is observed (energies 10000/20000/30000 tell them apart).
    Every step uses a fresh Philox block (cell 0, step)."""
    d = dict(v1(gx=1, gy=1), init_sheep=0, init_wolves=0,
             sheep_reproduce_prob=0)
    o = Oracle(d, seed=5, capacity=3)
    o.set_state([0, 0, 0], [10000, 20000, 30000], [0, 0, 0], [10 ** 6])
    counts = {}
    prev = [10000, 20000, 30000]
    for _ in range(3000):
        assert o.step() == 0
        _, e, _, _ = o.state()
        cur = [int(x) + 0 for x in e]
        # rank of each current agent in the previous order (energy - 1 each)
        perm = tuple(prev.index(x + 1) for x in cur)
        counts[perm] = counts.get(perm, 0) + 1
        prev = cur
    o.close()
    assert len(counts) == 6
    assert sps.chisquare(list(counts.values())).pvalue > 1e-3


# -------------------------------------------------------------- invariants

@pytest.mark.parametrize("d,seed", [(v1(100), 0), (v2s(100), 1),
                                    (v2ref(80), 2)])
def test_closed_form_invariants(d, seed):
    """N' = N - starved - eaten sheep + births;
    countdown sum' = countdown sum - dead cells + R * grass eaten."""
    o = Oracle(d, seed=seed)
    C = d["grid_x"] * d["grid_y"]
    R = d["grass_restart"]
    prev = o.stats()
    for _ in range(300):
        assert o.step() == 0
        cur = o.stats()
        ev = o.last_events()
        n0, n1 = prev[0] + prev[1], cur[0] + cur[1]
        assert n1 == n0 - ev["starved"] - ev["sheep_eaten"] + \
            ev["sheep_births"] + ev["wolf_births"]
        assert cur[5] == prev[5] - (C - prev[2]) + R * ev["grass_eaten"]
        assert cur[1] - prev[1] <= ev["wolf_births"]
        prev = cur
    o.close()


def test_capacity_overflow_is_reported():
    d = v1(100)
    o = Oracle(d, seed=0, capacity=700)
    rc, rows = o.run(400)
    assert rc == -8
    assert 0 < o.step_count < 400
    o.close()


# ------------------------------------------------------- energy limit

ENERGY_MAX = (1 << 24) - 1


def _step_rc(d, agents, grass, seed=0):
    o = Oracle(d, seed=seed, capacity=64)
    n = len(agents)
    o.set_state(np.zeros(n), [e for _, e in agents], [t for t, _ in agents],
                [grass])
    rc = o.step()
    o.close()
    return rc


def test_energy_limit_boundary():
    """Energy may reach 2^24 - 1 after an action; one more fails with -9."""
    g = 1 << 20
    d = _one_cell(wolves_gain_from_food=g, sheep_gain_from_food=g)
    # wolf: e - 1 + gain after eating the sheep
    assert _step_rc(d, [(0, 5), (1, ENERGY_MAX - g + 1)], 5) == 0
    assert _step_rc(d, [(0, 5), (1, ENERGY_MAX - g + 2)], 5) == -9
    # sheep on live grass
    assert _step_rc(d, [(0, ENERGY_MAX - g + 1)], 0) == 0
    assert _step_rc(d, [(0, ENERGY_MAX - g + 2)], 0) == -9
    # the move alone never exceeds the limit it started under
    assert _step_rc(d, [(0, ENERGY_MAX)], 3) == 0


def test_energy_limit_applies_after_the_split():
    """The parent keeps ceil(e/2): above the limit only if that half is."""
    d = _one_cell(sheep_gain_from_food=1 << 20, sheep_reproduce_prob=100)
    e_max_ok = 2 * ENERGY_MAX - (1 << 20) + 1   # e - 1 + gain = 2 ENERGY_MAX
    assert _step_rc(d, [(0, e_max_ok)], 0) == 0
    assert _step_rc(d, [(0, e_max_ok + 1)], 0) == -9
    out, _, _ = _run_one(d, [(0, e_max_ok)], grass=0)
    assert out == [(0, ENERGY_MAX), (0, ENERGY_MAX)]


def test_energy_limit_run_matches_spec():
    """Wolves with a large gain and no reproduction grow until one passes
    the limit: the oracle fails with -9 at the iteration the independent
    statement does, and every row before it agrees."""
    d = dict(v2s(gx=8, gy=8), init_sheep=60, init_wolves=6,
             wolves_gain_from_food=1 << 20, wolves_reproduce_prob=0,
             sheep_reproduce_prob=60)
    o = Oracle(d, seed=7)
    m = SpecModel(d, 7)
    fail = None
    for t in range(400):
        rc_o, rc_m = o.step(), m.step()
        assert rc_o == rc_m, f"step {t}: oracle {rc_o}, spec {rc_m}"
        if rc_o:
            fail = t
            break
        assert o.stats() == m.stats(), f"stats after {t + 1}"
    o.close()
    assert fail is not None and fail > 10 and rc_o == -9


# ------------------------------------------------------- planted mistakes

@pytest.mark.parametrize("mode", [1, 2, 3])
def test_planted_mistakes_are_caught(mode):
    """Each fault mode changes some state within 40 steps of a small run, so
    a state/stat comparison against the oracle would fail."""
    d = v2s(gx=16, gy=16)
    a, b = Oracle(d, seed=1), Oracle(d, seed=1)
    b.set_fault(mode)
    differs = False
    for _ in range(40):
        a.step()
        b.step()
        if any(not np.array_equal(x, y) for x, y in zip(a.state(), b.state())):
            differs = True
            break
    assert differs


# ---------------------------------------------------------- golden pins

def test_golden_config1_rows():
    g = np.load(os.path.join(GOLDEN, "v1_100_seed0.npz"))
    rows = simulate(json.loads(str(g["params"])), seed=int(g["seed"]),
                    capacity=int(g["capacity"]))
    assert np.array_equal(rows, g["rows"])


def test_golden_state_digests():
    with open(os.path.join(GOLDEN, "state_digests.json")) as f:
        cases = json.load(f)
    for name, c in cases.items():
        o = Oracle(c["params"], seed=c["seed"])
        for t, want in enumerate(c["digests"]):
            if t:
                assert o.step() == 0
            h = hashlib.sha256()
            for a in o.state():
                h.update(np.ascontiguousarray(a).tobytes())
            assert h.hexdigest() == want, f"{name} step {t}"
        o.close()
