"""The multi-threaded CPU restatement (oracle/pp_omp.c, the benchmark's CPU
baseline) writes the same statistics file as the single-threaded oracle, for
any thread count: same semantics, same counter-based random streams."""
import os
import subprocess

import pytest

from helpers import ROOT, v1, v2ref, v2s

BUILD = os.path.join(ROOT, "oracle", "build")
KEYS = ("INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
        "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
        "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
        "GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS")


def _write(path, d):
    path.write_text("".join(f"{k} = {d[k.lower()]}\n" for k in KEYS))


def _run(exe, args):
    return subprocess.run([os.path.join(BUILD, exe), *args],
                          capture_output=True, text=True, timeout=300)


CASES = [
    ("v1_100", v1(100, iters=300), 0, 0),
    ("v2s_64", v2s(64, iters=200), 5, 0),
    ("v2ref_50", v2ref(50, iters=150), 9, 0),
    ("v2s_40x17", dict(v2s(gx=40, gy=17), iters=120), 2, 0),
    ("v1_1x37", dict(v1(gx=1, gy=37), init_sheep=20, init_wolves=6,
                     iters=80), 4, 0),
    ("seed64", v2s(30, iters=60), 0x123456789ABCDEF, 0),
]


@pytest.mark.parametrize("name,d,seed,_", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_omp_equals_oracle(tmp_path, name, d, seed, _, threads):
    _write(tmp_path / "p.txt", d)
    cap = str(max(4 * (d["init_sheep"] + d["init_wolves"]),
                  3 * d["grid_x"] * d["grid_y"]))
    r1 = _run("pp_oracle", ["-p", str(tmp_path / "p.txt"), "-s",
                            str(tmp_path / "o.txt"), "-r", str(seed), "-m",
                            cap])
    r2 = _run("pp_omp", ["-p", str(tmp_path / "p.txt"), "-s",
                         str(tmp_path / "m.txt"), "-r", str(seed), "-m", cap,
                         "-t", str(threads)])
    assert r1.returncode == 0 and r2.returncode == 0, (r1.stderr, r2.stderr)
    assert (open(tmp_path / "o.txt", "rb").read()
            == open(tmp_path / "m.txt", "rb").read())


def test_omp_capacity_failure_like_oracle(tmp_path):
    d = dict(v2s(32, iters=300), sheep_reproduce_prob=60,
             wolves_reproduce_prob=40)
    _write(tmp_path / "p.txt", d)
    cap = str(d["init_sheep"] + d["init_wolves"] + 50)
    r1 = _run("pp_oracle", ["-p", str(tmp_path / "p.txt"), "-q", "-m", cap])
    r2 = _run("pp_omp", ["-p", str(tmp_path / "p.txt"), "-q", "-m", cap,
                         "-t", "4"])
    assert r1.returncode == r2.returncode != 0
    assert r1.stderr.strip() == r2.stderr.strip()


def test_omp_energy_limit_like_oracle(tmp_path):
    """Both stop with -9 at the same iteration when a wolf's energy passes
    2^24 - 1."""
    d = dict(v2s(16, iters=600), wolves_gain_from_food=1 << 20,
             wolves_reproduce_prob=0, sheep_reproduce_prob=60)
    _write(tmp_path / "p.txt", d)
    r1 = _run("pp_oracle", ["-p", str(tmp_path / "p.txt"), "-q", "-r", "3"])
    r2 = _run("pp_omp", ["-p", str(tmp_path / "p.txt"), "-q", "-r", "3",
                         "-t", "4"])
    assert r1.returncode == r2.returncode == 256 - 9, (r1.stderr, r2.stderr)
    assert "energy above its limit" in r1.stderr
    assert r1.stderr.strip() == r2.stderr.strip()
