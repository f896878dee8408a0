"""GPU: golden runs, the `pp` command line, edge cases and full-size
properties. Every comparison is bit-exact against the oracle or a fixture."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import (GOLDEN, ROOT, assert_state_equal, gpu_params, v1, v2ref,
                     v2s)

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _sim(d, seed=0, capacity=0):
    from pphpc_amd import Simulation
    return Simulation(gpu_params(d), seed=seed, capacity=capacity, device=0)


def _rows(sim, iters):
    return np.array([sim.init().as_tuple()] +
                    [r.as_tuple() for r in sim.run(iters)], dtype=np.uint64)


@pytest.mark.parametrize("fixture", ["v1_100_seed0.npz", "v2s_1600_seed0.npz"])
def test_golden_full_runs(fixture):
    """Config 1 (100x100 v1) and config 2 (1600x1600 v2-style, capacity 4x
    the initial agents): all 4001 rows equal the oracle's fixture."""
    _case(f"golden {fixture}")
    g = np.load(os.path.join(GOLDEN, fixture))
    d = json.loads(str(g["params"]))
    sim = _sim(d, seed=int(g["seed"]), capacity=int(g["capacity"]))
    rows = _rows(sim, d["iters"])
    sim.close()
    bad = np.nonzero((rows != g["rows"]).any(axis=1))[0]
    assert len(bad) == 0, f"first bad row {bad[0]}"


def test_cli_writes_the_oracle_stats_file(tmp_path):
    """`pp -p FILE -s FILE -r SEED` writes the same bytes as the oracle CLI."""
    _case("cli config1")
    params = os.path.join(ROOT, "configs", "v1_100.txt")
    ours, ref = str(tmp_path / "gpu.txt"), str(tmp_path / "oracle.txt")
    prof = str(tmp_path / "prof.tsv")
    r = subprocess.run([os.path.join(ROOT, "pphpc_amd", "bin", "pp"), "-p",
                        params, "-s", ours, "-r", "0", "-d", "0", "-i", prof],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    subprocess.run([os.path.join(ROOT, "oracle", "build", "pp_oracle"), "-p",
                    params, "-s", ref, "-r", "0"], check=True,
                   capture_output=True, timeout=120)
    assert open(ours, "rb").read() == open(ref, "rb").read()
    lines = open(prof).read().splitlines()
    names = [ln.split("\t")[0].strip('"') for ln in lines]
    # one replica: the emit ends each tick (no act_scan launches)
    assert {"move0", "sort_scan", "sort_down", "act",
            "emit_move"} <= set(names)
    # %lg keeps 6 significant digits per line
    assert abs(sum(float(ln.split("\t")[2]) for ln in lines) - 1.0) < 1e-5


def test_cli_f32_stats_equal_the_reference_format_of_the_oracle(tmp_path):
    """`pp -F` writes the statistics as the reference prints them (32-bit
    counts, single-precision means): the same bytes as pp_stats_save_f32 of
    the oracle's rows."""
    import pphpc_amd as pp
    from oracle.pp_oracle import simulate
    _case("cli config1 -F")
    params = os.path.join(ROOT, "configs", "v1_100.txt")
    ours, ref = str(tmp_path / "gpu.txt"), str(tmp_path / "oracle.txt")
    r = subprocess.run([os.path.join(ROOT, "pphpc_amd", "bin", "pp"), "-p",
                        params, "-s", ours, "-r", "0", "-d", "0", "-F"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p = pp.load_params(params)
    d = {f: getattr(p, f) for f in pp.sim.PARAM_FIELDS}
    rows = simulate(d, seed=0)
    pp.stats_save(ref, [pp.Stats(*map(int, row)) for row in rows], p,
                  f32=True)
    assert open(ours, "rb").read() == open(ref, "rb").read()


@pytest.mark.parametrize("dump_type", [0x01, 0x10])
def test_cli_dump_equals_the_oracle_dump(tmp_path, dump_type):
    """`pp -D PREFIX` dumps every state (after init and each iteration) with
    the same bytes as the oracle CLI's dump."""
    _case(f"cli dump type {dump_type:#x}")
    d = v2s(24, iters=15)
    path = tmp_path / "p.txt"
    keys = ("INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
            "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
            "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
            "GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS")
    path.write_text("".join(f"{k} = {d[k.lower()]}\n" for k in keys))
    ours, ref = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    r = subprocess.run([os.path.join(ROOT, "pphpc_amd", "bin", "pp"), "-p",
                        str(path), "-s", str(tmp_path / "s1.txt"), "-r", "7",
                        "-d", "0", "-m", "4096", "-D", ours, "-T",
                        str(dump_type)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    subprocess.run([os.path.join(ROOT, "oracle", "build", "pp_oracle"), "-p",
                    str(path), "-s", str(tmp_path / "s2.txt"), "-r", "7",
                    "-m", "4096", "-D", ref, "-T", str(dump_type)],
                   check=True, capture_output=True, timeout=120)
    for part in ("_agents.txt", "_cells.txt"):
        a, b = open(ours + part).read(), open(ref + part).read()
        assert a.count("\nIter") == 16
        assert a == b, part
    assert (open(tmp_path / "s1.txt", "rb").read()
            == open(tmp_path / "s2.txt", "rb").read())


def test_extinction_and_empty_population():
    """Wolves without sheep starve: population reaches zero and stays; an
    empty initial population is valid."""
    _case("extinction")
    from oracle.pp_oracle import simulate
    d = dict(v1(30), init_sheep=0, init_wolves=50, iters=80)
    sim = _sim(d, seed=2, capacity=1000)
    rows = _rows(sim, 80)
    assert rows[-1][0] == 0 and rows[-1][1] == 0
    assert np.array_equal(rows, simulate(d, seed=2, capacity=1000, iters=80))
    sim.close()
    d = dict(v1(10), init_sheep=0, init_wolves=0, iters=5)
    sim = _sim(d, capacity=16)
    rows = _rows(sim, 5)
    assert np.array_equal(rows, simulate(d, capacity=16, iters=5))
    sim.close()


def test_capacity_overflow_matches_oracle():
    """PP_OUT_OF_RESOURCES at the same iteration as the oracle, and rows
    before it identical (pp_gpu.c:847-852 semantics)."""
    _case("capacity overflow")
    from oracle.pp_oracle import Oracle
    from pphpc_amd import PPError, PP_OUT_OF_RESOURCES
    d = v1(100)
    cap = 700
    o = Oracle(d, seed=0, capacity=cap)
    rc, ref = o.run(500)
    assert rc == -8
    fail_step = o.step_count          # steps completed before the failure
    o.close()
    sim = _sim(d, capacity=cap)
    sim.init()
    with pytest.raises(PPError) as e:
        sim.run(500)
    assert e.value.code == PP_OUT_OF_RESOURCES
    assert sim.error_step == fail_step
    assert sim.step_count == fail_step
    sim.close()
    # rows up to the failure agree
    sim = _sim(d, capacity=cap)
    sim.init()
    got = np.array([r.as_tuple() for r in sim.run(fail_step)], np.uint64)
    assert np.array_equal(got, ref[:fail_step])
    sim.close()


def test_large_grass_restart_uses_global_ring():
    """R + 1 > 4096 takes the global-atomic histogram path at init."""
    _case("large R")
    from oracle.pp_oracle import Oracle
    d = dict(v2s(60), grass_restart=9000)
    sim = _sim(d, seed=4)
    o = Oracle(d, seed=4, capacity=sim.capacity)
    assert sim.init().as_tuple() == o.stats()
    assert_state_equal(sim.state(), o.state(), "init R=9000")
    for t in range(30):
        assert sim.step().as_tuple() == (o.step(), o.stats())[1]
    assert_state_equal(sim.state(), o.state(), "R=9000 step 30")
    sim.close()
    o.close()


def test_reinit_graph_and_profiling_modes_agree():
    """Re-init reproduces the run; graph replay, per-launch and profiled
    modes give identical rows."""
    _case("modes")
    d = v2s(200)
    sim = _sim(d, seed=3)
    a = _rows(sim, 300)
    b = _rows(sim, 300)                       # re-init on the same handle
    os.environ["PPHPC_NO_GRAPH"] = "1"
    try:
        c = _rows(sim, 300)
    finally:
        del os.environ["PPHPC_NO_GRAPH"]
    sim.set_profiling(True)
    e = _rows(sim, 300)
    prof = sim.profile()
    sim.close()
    assert np.array_equal(a, b) and np.array_equal(a, c) and \
        np.array_equal(a, e)
    assert prof["act"][0] == 300 and prof["emit_move"][0] == 300
    # a plain simulation ends each tick in the emit's work-group 0
    assert prof.get("act_scan", (0, 0))[0] == 0 and prof["move0"][0] == 1
    # iteration 1 sorts the freshly placed agents with the radix sort
    # (40000 cells: 16 bits = 2 passes of 8) after clearing the bucket
    # counts, the 299 others by row buckets; the 200 buckets fit one scan
    # chunk, which the bucket sort scans itself
    assert prof["sort_down"][0] == 2 and prof["sort_scan"][0] == 2
    assert prof["sort_up"][0] == 2
    assert prof["bucket_sort"][0] == 299 and prof["bucket_scan"][0] == 1


@pytest.mark.parametrize("shift", [0, 8, 9, 11])
def test_sort_variants_agree(shift):
    """The radix sort every iteration (0) and row buckets of 2^8..2^11 cells
    give the oracle's rows and state, on grids narrower and wider than a
    block, one and two rows high, and with a dense corner."""
    from oracle.pp_oracle import Oracle
    _case(f"sort variant {shift}")
    cases = [(v2s(600), 1), (v2s(gx=3000, gy=20), 2), (v2s(gx=257, gy=1), 3),
             (v2s(gx=700, gy=2), 4),
             (dict(v2ref(60), grid_x=60, grid_y=60, init_sheep=6000,
                   init_wolves=300), 5),
             # rows of ~5000 agents: more than a bucket sorts in LDS at 2^11
             (dict(v2ref(60), grid_x=2048, grid_y=4, grid_xy=8192,
                   init_sheep=16000, init_wolves=4000), 6),
             # more buckets than one scan chunk holds: the separate scans
             (v2s(gx=300, gy=2500), 7)]
    for d, seed in cases:
        sim = _sim(d, seed=seed)
        sim.set_sort(shift)
        rows = _rows(sim, 120)
        o = Oracle(d, seed=seed, capacity=sim.capacity)
        rc, orows = o.run(120)
        assert rc == 0
        assert np.array_equal(rows[1:], orows), (d["grid_x"], d["grid_y"])
        assert_state_equal(sim.state(), o.state(), f"{shift} {d['grid_x']}")
        o.close()
        sim.close()


@pytest.mark.parametrize("shift", [8, 11])
def test_bucket_prefix_tables_agree(shift, monkeypatch):
    """The bucket sort reading its bucket records, which k_bprefix and k_brec
    write (used by default from 6144 buckets, here forced for
    every layout) gives the oracle's rows and state: grids narrower and wider
    than a block, a dense corner, a bucket above the LDS capacity and a
    layout with more than 4096 buckets (all three count levels)."""
    from oracle.pp_oracle import Oracle
    monkeypatch.setenv("PPHPC_BPREFIX_MIN", "0")
    _case(f"bucket prefix tables, shift {shift}")
    cases = [(v2s(600), 1), (v2s(gx=3000, gy=20), 2), (v2s(gx=257, gy=1), 3),
             (dict(v2ref(60), grid_x=60, grid_y=60, init_sheep=6000,
                   init_wolves=300), 5),
             (dict(v2ref(60), grid_x=2048, grid_y=4, grid_xy=8192,
                   init_sheep=16000, init_wolves=4000), 6),
             (v2s(gx=300, gy=2500), 7)]
    for d, seed in cases:
        sim = _sim(d, seed=seed)
        sim.set_sort(shift)
        rows = _rows(sim, 120)
        o = Oracle(d, seed=seed, capacity=sim.capacity)
        rc, orows = o.run(120)
        assert rc == 0
        assert np.array_equal(rows[1:], orows), (d["grid_x"], d["grid_y"])
        assert_state_equal(sim.state(), o.state(), f"prefix {shift} {d['grid_x']}")
        o.close()
        sim.close()


@pytest.mark.parametrize("wide", ["0", "1"], ids=["wg256", "wg512"])
def test_group_width_agrees(wide, monkeypatch):
    """One replica with the act and emit in groups of 256 and of 512 threads
    (chosen by the agents' tile count; here forced): rows and state equal
    the oracle's. Among the cases, a grid whose tick outputs end mid-round, a
    one-row grid and a dense corner whose tiles emit many children."""
    from oracle.pp_oracle import Oracle
    monkeypatch.setenv("PPHPC_WIDE_GROUPS", wide)
    _case(f"group width {wide}")
    cases = [(v2s(600), 11), (v2s(gx=3000, gy=20), 12), (v2s(gx=257, gy=1), 13),
             (dict(v2ref(60), grid_x=60, grid_y=60, init_sheep=6000,
                   init_wolves=300), 14), (v2s(gx=1000, gy=1000), 15)]
    for d, seed in cases:
        sim = _sim(d, seed=seed)
        rows = _rows(sim, 60)
        o = Oracle(d, seed=seed, capacity=sim.capacity)
        rc, orows = o.run(60)
        assert rc == 0
        assert np.array_equal(rows[1:], orows), (d["grid_x"], d["grid_y"])
        assert_state_equal(sim.state(), o.state(), f"width {wide} {d['grid_x']}")
        o.close()
        sim.close()


def test_step_by_step_equals_run():
    """Rows and final agent state agree whether 50 iterations run as one
    call (graphs of 8 and single ticks; only the last tick writes the
    per-agent state) or one at a time (every tick writes it)."""
    _case("step vs run")
    d = v2s(128)
    sim = _sim(d, seed=8)
    a = _rows(sim, 50)
    sa = sim.state()
    sim.init()
    b = [r.as_tuple() for r in [sim.step() for _ in range(50)]]
    assert np.array_equal(a[1:], np.array(b, np.uint64))
    assert_state_equal(sim.state(), sa, "step vs run 50")
    sim.run(13)
    from oracle.pp_oracle import Oracle
    o = Oracle(d, seed=8, capacity=sim.capacity)
    rc, _ = o.run(63)
    assert rc == 0
    assert_state_equal(sim.state(), o.state(), "run 13 after 50 steps")
    o.close()
    sim.close()


def test_two_simulations_share_the_gpu():
    """Independent handles on one device do not interfere."""
    _case("two handles")
    d1, d2 = v2s(100), v1(150)
    s1, s2 = _sim(d1, seed=1), _sim(d2, seed=2)
    s1.init()
    s2.init()
    r1, r2 = [], []
    for _ in range(20):
        r1 += [r.as_tuple() for r in s1.run(25)]
        r2 += [r.as_tuple() for r in s2.run(25)]
    from oracle.pp_oracle import simulate
    assert np.array_equal(np.array(r1, np.uint64),
                          simulate(d1, 1, s1.capacity, 500)[1:])
    assert np.array_equal(np.array(r2, np.uint64),
                          simulate(d2, 2, s2.capacity, 500)[1:])
    s1.close()
    s2.close()


def test_full_size_12800_properties():
    """12800x12800 (163.84 M cells, 39.3 M agents), capacity 4x: the run is
    deterministic (two runs, same rows), the statistics agree with the state
    (counts, energy sums, grass from the materialised countdowns), and the
    state is ordered by cell."""
    _case("12800^2 properties")
    d = v2s(12800)
    n0 = d["init_sheep"] + d["init_wolves"]
    sim = _sim(d, seed=0, capacity=4 * n0)
    a = _rows(sim, 4)
    cell, energy, typ, grass = sim.state()
    b = _rows(sim, 4)
    sim.close()
    assert np.array_equal(a, b)
    last = a[-1]
    assert last[0] == (typ == 0).sum() and last[1] == (typ == 1).sum()
    assert last[3] == energy[typ == 0].astype(np.int64).sum()
    assert last[4] == energy[typ == 1].astype(np.int64).sum()
    assert last[2] == (grass == 0).sum()
    assert last[5] == grass.astype(np.uint64).sum()
    assert (np.diff(cell.astype(np.int64)) >= 0).all()
    assert (energy >= 0).all() and cell.max() < 12800 * 12800


@pytest.mark.parametrize("width", ["1", "2", "4"])
def test_grass_storage_widths_agree(width):
    """The regrow step stored modulo 2^8, 2^16 or 2^32 (with the periodic
    re-stamp of alive cells) gives the oracle's rows over 1200 iterations
    (many re-stamp periods) and its final grass countdowns."""
    _case(f"grass width {width}")
    from oracle.pp_oracle import Oracle
    d = v2s(150)
    os.environ["PPHPC_GRASS_BYTES"] = width
    try:
        sim = _sim(d, seed=6)
    finally:
        del os.environ["PPHPC_GRASS_BYTES"]
    o = Oracle(d, seed=6, capacity=sim.capacity)
    rows = _rows(sim, 1200)
    ref = [o.stats()]
    _, more = o.run(1200)
    assert np.array_equal(rows, np.vstack([np.array(ref, np.uint64), more]))
    assert_state_equal(sim.state(), o.state(), f"width {width} final")
    sim.close()
    o.close()
