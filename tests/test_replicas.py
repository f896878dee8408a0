"""Replica statistics (focal measures, Welch t, Mann-Whitney, Holm) and the
statistical equivalence of the model under test with an independent
simulator of the Java single-thread rules (oracle/pp_statref.c, MT19937),
plus the NetLogo rules at the size where the two rule sets are not yet
distinguishable."""
import math
from multiprocessing import Pool

import numpy as np
import pytest
from scipy import special, stats

from helpers import v1, v2ref
from pphpc_amd import replicas as rp


def test_betainc_matches_scipy():
    rng = np.random.default_rng(0)
    for _ in range(300):
        a, b = rng.uniform(0.2, 40, 2)
        x = rng.uniform(0, 1)
        assert abs(rp.betainc_reg(a, b, x) - special.betainc(a, b, x)) < 1e-10


def test_welch_matches_scipy():
    rng = np.random.default_rng(1)
    for _ in range(100):
        a = rng.normal(0, rng.uniform(0.5, 3), 30)
        b = rng.normal(rng.uniform(-1, 1), rng.uniform(0.5, 3), 30)
        t, df, p = rp.welch_t(a, b)
        ref = stats.ttest_ind(a, b, equal_var=False)
        assert abs(t - ref.statistic) < 1e-9 and abs(p - ref.pvalue) < 1e-9


def test_mann_whitney_matches_scipy():
    rng = np.random.default_rng(2)
    for _ in range(100):
        a = rng.integers(0, 40, 30).astype(float)   # with ties
        b = rng.integers(0, 45, 30).astype(float)
        u, p = rp.mann_whitney(a, b)
        ref = stats.mannwhitneyu(a, b, alternative="two-sided",
                                 method="asymptotic", use_continuity=True)
        assert abs(u - ref.statistic) < 1e-9 and abs(p - ref.pvalue) < 1e-9


def test_focal_measures_closed_form():
    it = 1200
    t = np.arange(it + 1)
    rows = np.zeros((it + 1, 7), np.uint64)
    rows[:, 0] = (1000 + 500 * np.sin(t / 50.0)).astype(np.uint64)
    rows[:, 1] = 10
    rows[:, 2] = t
    rows[:, 3] = rows[:, 0] * 3
    rows[:, 4] = 70
    rows[:, 5] = 4 * t
    fm = rp.focal_measures(rows, cells=4)
    s = rows[:, 0].astype(float)
    assert fm["sheep.max"] == s.max() and fm["sheep.argmax"] == np.argmax(s)
    assert fm["sheep.min"] == s.min() and fm["sheep.argmin"] == np.argmin(s)
    assert fm["grass.ss_mean"] == pytest.approx(t[1001:].mean())
    assert fm["grass.ss_std"] == pytest.approx(t[1001:].std(ddof=1))
    assert fm["sheep_energy.ss_mean"] == pytest.approx(3.0)
    assert fm["wolf_energy.max"] == 7.0 and fm["countdown.argmax"] == it
    assert len(fm) == 36


def test_holm_rejects_only_shifted_measures():
    rng = np.random.default_rng(3)
    A = [{"x": rng.normal(0, 1), "y": rng.normal(0, 1)} for _ in range(30)]
    B = [{"x": rng.normal(0, 1), "y": rng.normal(3, 1)} for _ in range(30)]
    rep = rp.compare(A, B)
    rej = {r["measure"]: r["reject"] for r in rep["measures"]}
    assert rej == {"x": False, "y": True} and not rep["pass"]


def _oracle_focal(args):
    d, seed, fault = args
    from oracle.pp_oracle import Oracle
    o = Oracle(d, seed=seed)
    if fault:
        o.set_fault(fault)
    row0 = np.array([o.stats()], np.uint64)
    rc, rows = o.run(d["iters"])
    o.close()
    assert rc == 0
    return rp.focal_measures(np.vstack([row0, rows]), d["grid_x"] * d["grid_y"])


def _statref_focal(args):
    d, seed, rules = args
    from oracle.pp_oracle import statref_run
    return rp.focal_measures(statref_run(d, seed, rules=rules),
                             d["grid_x"] * d["grid_y"])


def _samples(d, n=30, fault=0, rules="cells"):
    with Pool(6) as pool:
        ours = pool.map(_oracle_focal, [(d, s, fault) for s in range(1, n + 1)])
        ref = pool.map(_statref_focal,
                       [(d, s, rules) for s in range(1001, 1001 + n)])
    return ours, ref


@pytest.mark.parametrize("name,d,rules", [
    ("v1 100x100", v1(100), "cells"),
    ("reference v2 100x100", v2ref(100), "cells"),
    ("v1 100x100 NetLogo rules", v1(100), "netlogo"),
])
def test_model_equivalent_to_reference_rules(name, d, rules):
    """30 replicas of the model under test (the oracle, which the GPU
    matches bit for bit) against 30 replicas of the independent simulator:
    no focal measure differs after Holm correction."""
    ours, ref = _samples(d, rules=rules)
    rep = rp.compare(ours, ref)
    bad = [(r["measure"], r["test"], r["p"]) for r in rep["measures"]
           if r["reject"]]
    assert rep["pass"], f"{name}: {bad}"


def test_statistical_test_detects_a_planted_mistake():
    """The same comparison rejects a model whose sheep gain is doubled."""
    ours, ref = _samples(v1(100), n=30, fault=4)
    rep = rp.compare(ours, ref)
    assert not rep["pass"]
    assert sum(r["reject"] for r in rep["measures"]) >= 5
