"""State dumps (pp_sim_dump / pp_dump_state and the oracle's ppo_dump), the
counterpart of the reference's PPG_DUMP output (opencl/pp/pp_gpu.c:438-555):
hand-worked known answers for both writers, the dump-type filter, and the
oracle CLI's dumps against its own statistics rows. No GPU calls."""
import ctypes
import os
import subprocess

import numpy as np

import pphpc_amd as pp
from helpers import ROOT, v1
from oracle.pp_oracle import Oracle

# 3 x 2 torus: agents in cells 0, 0, 4 (x, y = (0,0), (0,0), (1,1)).
PARAMS = dict(init_sheep=2, sheep_gain_from_food=4,
              sheep_reproduce_threshold=2, sheep_reproduce_prob=4,
              init_wolves=1, wolves_gain_from_food=20,
              wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
              grass_restart=10, grid_x=3, grid_y=2, grid_xy=6, iters=5)
CELL = np.array([0, 0, 4], np.uint32)
ENERGY = np.array([5, 7, 2], np.int32)
TYPE = np.array([0, 1, 0], np.uint8)
GRASS = np.array([0, 3, 0, 1, 0, 2], np.uint32)

AGENTS_TXT = ("\nIter 0, agents=3\n"
              "[   0] (   0,    0) type=0 energy=5\n"
              "[   1] (   0,    0) type=1 energy=7\n"
              "[   2] (   1,    1) type=0 energy=2\n")
CELLS_TXT = ("\nIteration 0\n"
             "(0, 0) -> (0, 1) More than 1 agent present [Grass: 0]\n"
             "(1, 0) -> (-, -)  [Grass: 3]\n"
             "(2, 0) -> (-, -)  [Grass: 0]\n"
             "(0, 1) -> (-, -)  [Grass: 1]\n"
             "(1, 1) -> (2, 2)  [Grass: 0]\n"
             "(2, 1) -> (-, -)  [Grass: 2]\n")
CELLS_AGENTS_ONLY = ("\nIteration 0\n"
                     "(0, 0) -> (0, 1) More than 1 agent present [Grass: 0]\n"
                     "\n(1, 1) -> (2, 2)  [Grass: 0]\n")


def _product_dump(tmp_path, dump_type, tag):
    a, c = str(tmp_path / f"{tag}_a.txt"), str(tmp_path / f"{tag}_c.txt")
    vp = ctypes.c_void_p
    rc = pp.sim._lib().pp_dump_state(
        a.encode(), c.encode(), 0, 0, dump_type,
        ctypes.byref(pp.Params.from_values(**PARAMS)), len(CELL),
        CELL.ctypes.data_as(vp), ENERGY.ctypes.data_as(vp),
        TYPE.ctypes.data_as(vp), GRASS.ctypes.data_as(vp))
    assert rc == 0
    return open(a).read(), open(c).read()


def _oracle_dump(tmp_path, dump_type, tag):
    a, c = str(tmp_path / f"{tag}_oa.txt"), str(tmp_path / f"{tag}_oc.txt")
    o = Oracle(PARAMS, seed=0, capacity=16)
    o.set_state(CELL, ENERGY, TYPE, GRASS)
    assert o.dump(a, c, dump_type) == 0
    o.close()
    return open(a).read(), open(c).read()


def test_dump_known_answer():
    import tempfile, pathlib
    with tempfile.TemporaryDirectory() as d:
        t = pathlib.Path(d)
        for writer in (_product_dump, _oracle_dump):
            a, c = writer(t, 0x01, writer.__name__)
            assert a == AGENTS_TXT, writer.__name__
            assert c == CELLS_TXT, writer.__name__
            a, c = writer(t, 0x11, writer.__name__ + "f")
            assert c == CELLS_AGENTS_ONLY, writer.__name__


def test_dump_appends_and_rejects_bad_cells(tmp_path):
    a = str(tmp_path / "a.txt")
    vp = ctypes.c_void_p
    p = pp.Params.from_values(**PARAMS)
    for append in (0, 1):
        assert pp.sim._lib().pp_dump_state(
            a.encode(), None, append, 0, 1, ctypes.byref(p), len(CELL),
            CELL.ctypes.data_as(vp), ENERGY.ctypes.data_as(vp),
            TYPE.ctypes.data_as(vp), GRASS.ctypes.data_as(vp)) == 0
    assert open(a).read() == AGENTS_TXT * 2
    bad = np.array([0, 0, 6], np.uint32)     # cell 6 is off the 3 x 2 grid
    assert pp.sim._lib().pp_dump_state(
        None, str(tmp_path / "c.txt").encode(), 0, 0, 1, ctypes.byref(p), 3,
        bad.ctypes.data_as(vp), ENERGY.ctypes.data_as(vp),
        TYPE.ctypes.data_as(vp), GRASS.ctypes.data_as(vp)) == pp.PP_INVALID_ARGS


def _blocks(text, head):
    out = []
    for block in text.split("\n" + head)[1:]:
        lines = block.split("\n")
        out.append((lines[0], [ln for ln in lines[1:] if ln]))
    return out


def test_oracle_cli_dump_matches_its_rows(tmp_path):
    """Every dumped state agrees with the statistics row of its iteration:
    sheep, wolves, energy sums, alive grass and countdown sum."""
    d = v1(20, iters=12)
    path = tmp_path / "p.txt"
    keys = ("INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
            "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
            "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
            "GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS")
    path.write_text("".join(f"{k} = {d[k.lower()]}\n" for k in keys))
    exe = os.path.join(ROOT, "oracle", "build", "pp_oracle")
    prefix = str(tmp_path / "dump")
    subprocess.run([exe, "-p", str(path), "-s", str(tmp_path / "s.txt"),
                    "-r", "3", "-D", prefix], check=True, capture_output=True)
    stats = [ln.split("\t") for ln in
             open(tmp_path / "s.txt").read().splitlines()]
    agents = _blocks(open(prefix + "_agents.txt").read(), "Iter ")
    cells = _blocks(open(prefix + "_cells.txt").read(), "Iteration ")
    assert len(agents) == len(cells) == d["iters"] + 1 == len(stats)
    for t, ((ha, la), (hc, lc), row) in enumerate(zip(agents, cells, stats)):
        assert ha.startswith(f"{t}, agents=") and hc == f"{t}"
        typ = np.array([int(ln.split("type=")[1].split()[0]) for ln in la])
        en = np.array([int(ln.split("energy=")[1]) for ln in la])
        grass = np.array([int(ln.split("[Grass: ")[1][:-1]) for ln in lc])
        assert len(lc) == 400
        sheep, wolves = int((typ == 0).sum()), int((typ == 1).sum())
        assert (sheep, wolves, int((grass == 0).sum())) == tuple(
            int(x) for x in row[:3])
        for col, s, n in ((3, en[typ == 0].sum(), sheep),
                          (4, en[typ == 1].sum(), wolves),
                          (5, grass.sum(), 400)):
            assert row[col] == f"{(s / n if n else 0.0):f}"
