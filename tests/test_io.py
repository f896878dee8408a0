"""Parameter-file parser and statistics writer (CPU; no GPU calls).

Known answers: the values of the reference's 20 parameter files
(configs/config{25..12800}v{1,2}.txt: 400 sheep and 200 wolves per 100x100
cells, v1 = gains 4/20, probabilities 4/5, restart 10; v2 = gains 30/10,
probabilities 10/5, restart 15; thresholds 2; 4000 iterations), regenerated
here as files in a temporary directory; and the error codes of
pp_load_params (pp_common.c:50-52, :60-62, :163-174).
"""
import os
import random

import numpy as np
import pytest

import pphpc_amd as pp
from oracle import pp_oracle

SIZES = (25, 50, 100, 200, 400, 800, 1600, 3200, 6400, 12800)


def ref_values(L, v):
    cells = L * L
    if v == 1:
        gains, probs, restart = (4, 20), (4, 5), 10
    else:
        gains, probs, restart = (30, 10), (10, 5), 15
    return {"INIT_SHEEP": cells * 400 // 10000, "SHEEP_GAIN_FROM_FOOD": gains[0],
            "SHEEP_REPRODUCE_THRESHOLD": 2, "SHEEP_REPRODUCE_PROB": probs[0],
            "INIT_WOLVES": cells * 200 // 10000,
            "WOLVES_GAIN_FROM_FOOD": gains[1],
            "WOLVES_REPRODUCE_THRESHOLD": 2,
            "WOLVES_REPRODUCE_PROB": probs[1], "GRASS_RESTART": restart,
            "GRID_X": L, "GRID_Y": L, "ITERS": 4000}


def write(path, items):
    with open(path, "w") as f:
        for k, v in items:
            f.write(f"{k} = {v}\n")


@pytest.mark.parametrize("L", SIZES)
@pytest.mark.parametrize("v", (1, 2))
def test_reference_parameter_files(tmp_path, L, v):
    vals = ref_values(L, v)
    path = str(tmp_path / f"config{L}v{v}.txt")
    write(path, vals.items())
    p = pp.load_params(path)
    for key, field in zip(pp.PARAM_KEYS, [f for f in pp.sim.PARAM_FIELDS
                                          if f != "grid_xy"]):
        assert getattr(p, field) == vals[key]
    assert p.grid_xy == L * L
    rc, q = pp_oracle.load_params(path)
    assert rc == 0 and all(getattr(q, f) == getattr(p, f)
                           for f in pp.sim.PARAM_FIELDS)


def test_25_cell_known_answer(tmp_path):
    # config25v1.txt: INIT_SHEEP = 25, INIT_WOLVES = 12 (floor of 12.5)
    vals = ref_values(25, 1)
    assert vals["INIT_SHEEP"] == 25 and vals["INIT_WOLVES"] == 12


def test_key_order_is_free(tmp_path):
    vals = list(ref_values(100, 1).items())
    random.Random(3).shuffle(vals)
    path = str(tmp_path / "c.txt")
    write(path, vals)
    assert pp.load_params(path).grid_xy == 10000


@pytest.mark.parametrize("mutate,code", [
    (lambda it: it[:-1], pp.PP_INVALID_PARAMS_FILE),                  # missing
    (lambda it: it + [it[0]], pp.PP_INVALID_PARAMS_FILE),             # repeat
    (lambda it: it + [("NOT_A_KEY", 3)], pp.PP_INVALID_PARAMS_FILE),  # unknown
    (lambda it: [("init_sheep", 3)] + it[1:], pp.PP_INVALID_PARAMS_FILE),
])
def test_invalid_files(tmp_path, mutate, code):
    path = str(tmp_path / "bad.txt")
    write(path, mutate(list(ref_values(100, 1).items())))
    with pytest.raises(pp.PPError) as e:
        pp.load_params(path)
    assert e.value.code == code
    assert pp_oracle.load_params(path)[0] == code


def test_missing_file():
    with pytest.raises(pp.PPError) as e:
        pp.load_params("/nonexistent/params.txt")
    assert e.value.code == pp.PP_UNABLE_TO_OPEN_PARAMS_FILE


def test_parsers_agree_on_fuzzed_files(tmp_path):
    """Product parser and oracle parser (two restatements) agree on random
    mixes of valid, repeated, missing and unknown keys and spacing."""
    rng = random.Random(11)
    keys = list(ref_values(100, 1).items())
    for i in range(200):
        items = [kv for kv in keys if rng.random() > 0.05]
        if rng.random() < 0.2:
            items.append(rng.choice(keys))
        if rng.random() < 0.1:
            items.append(("BOGUS", 1))
        rng.shuffle(items)
        path = str(tmp_path / f"f{i}.txt")
        with open(path, "w") as f:
            for k, v in items:
                sep = rng.choice([" = ", "  =  ", " =\t", " = "])
                f.write(f"{k}{sep}{v}\n")
        try:
            p = pp.load_params(path)
            rc = 0
        except pp.PPError as e:
            rc, p = e.code, None
        orc, q = pp_oracle.load_params(path)
        assert rc == orc, path
        if rc == 0:
            assert all(getattr(q, f) == getattr(p, f)
                       for f in pp.sim.PARAM_FIELDS)


def test_stats_writer_closed_form(tmp_path):
    p = pp.Params.from_values(**{f: 1 for f in pp.sim.PARAM_FIELDS
                                 if f != "grid_xy"} | {"grid_x": 4,
                                                       "grid_y": 5})
    rows = [pp.Stats(10, 3, 7, 25, 31, 9, 0), pp.Stats(0, 0, 20, 0, 0, 0, 0),
            pp.Stats(2**40, 1, 0, 3 * 2**40, 7, 2**33, 0)]
    path = str(tmp_path / "s.txt")
    pp.stats_save(path, rows, p)
    want = ("10\t3\t7\t2.500000\t10.333333\t0.450000\n"
            "0\t0\t20\t0.000000\t0.000000\t0.000000\n"
            f"{2**40}\t1\t0\t3.000000\t7.000000\t{2**33 / 20:f}\n")
    assert open(path).read() == want


def test_stats_writer_f32_known_answers(tmp_path):
    """pp_stats_save_f32 prints what the reference's pp_stats_save
    (pp_common.c:216-224) prints for the same rows: 32-bit cl_uint fields,
    "%d" counts, single-precision quotients. Hand-worked answers: 2^24 + 1
    becomes 2^24 as a float, so 16777217/3 is 5592405.5 (the float nearest
    16777216/3), not 5592405.666667; 1234567/1000 is the float nearest
    1234.567, 1234.5670166015625; 9/20 rounds back to 0.450000; sums above
    2^32 - 1 wrap (2^32 + 6 -> 6, 2^33 -> 0) and a count of 2^31 + 5 prints
    as -2147483643."""
    p = pp.Params.from_values(**{f: 1 for f in pp.sim.PARAM_FIELDS
                                 if f != "grid_xy"} | {"grid_x": 4,
                                                       "grid_y": 5})
    rows = [pp.Stats(1000, 3, 7, 1234567, 16777217, 9, 0),
            pp.Stats(0, 0, 20, 0, 0, 0, 0),
            pp.Stats(2**31 + 5, 1, 0, 2**32 + 6, 7, 2**33, 0)]
    path = str(tmp_path / "s.txt")
    pp.stats_save(path, rows, p, f32=True)
    want = ("1000\t3\t7\t1234.567017\t5592405.500000\t0.450000\n"
            "0\t0\t20\t0.000000\t0.000000\t0.000000\n"
            "-2147483643\t1\t0\t0.000000\t7.000000\t0.000000\n")
    assert open(path).read() == want


def test_stats_writers_agree(tmp_path):
    """Product writer and oracle writer give identical bytes."""
    rng = np.random.default_rng(5)
    arr = rng.integers(0, 2**40, size=(50, 7), dtype=np.uint64)
    arr[::7, 0] = 0
    arr[::5, 1] = 0
    p = pp.Params.from_values(init_sheep=1, sheep_gain_from_food=1,
                              sheep_reproduce_threshold=1,
                              sheep_reproduce_prob=1, init_wolves=1,
                              wolves_gain_from_food=1,
                              wolves_reproduce_threshold=1,
                              wolves_reproduce_prob=1, grass_restart=1,
                              grid_x=37, grid_y=11, iters=49)
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    pp.stats_save(a, [pp.Stats(*map(int, r)) for r in arr], p)
    orows = (pp_oracle.OStats * len(arr))(*[pp_oracle.OStats(*map(int, r))
                                            for r in arr])
    assert pp_oracle.lib().ppo_stats_save(b.encode(), orows, len(arr),
                                          37 * 11) == 0
    assert open(a, "rb").read() == open(b, "rb").read()


def test_stats_save_unwritable():
    p = pp.Params.from_values(**{f: 1 for f in pp.sim.PARAM_FIELDS
                                 if f != "grid_xy"})
    with pytest.raises(pp.PPError) as e:
        pp.stats_save("/nonexistent/dir/s.txt", [pp.Stats()], p)
    assert e.value.code == pp.PP_UNABLE_SAVE_STATS


def test_strerror_covers_codes():
    fallback = pp.strerror(-7)
    assert fallback == "unknown error code"
    msgs = [pp.strerror(code) for code in (0, -1, -2, -3, -4, -5, -6, -8)]
    assert fallback not in msgs and len(set(msgs)) == len(msgs)


def test_repository_parameter_files():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = pp.load_params(os.path.join(root, "configs", "v2s_1600.txt"))
    assert (p.init_sheep, p.init_wolves, p.grid_x, p.grass_restart,
            p.sheep_reproduce_prob, p.wolves_gain_from_food) == \
        (409600, 204800, 1600, 5, 10, 20)
    p = pp.load_params(os.path.join(root, "configs", "v1_100.txt"))
    assert p.as_dict() == {**{f: v for f, v in zip(
        [f for f in pp.sim.PARAM_FIELDS if f != "grid_xy"],
        ref_values(100, 1).values())}, "grid_xy": 10000}
