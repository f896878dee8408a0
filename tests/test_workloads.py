"""Workload draws of the benchmark (host only)."""
import collections

import numpy as np

from pphpc_amd.workloads import (BATCH_SIZES, SERVE_SIZES, batch_replicas,
                                 footprint_bytes, reserve_for, serve_requests,
                                 v1, v2s)


def test_parameter_sets_match_the_reference_densities():
    d = v1(100)
    assert (d["init_sheep"], d["init_wolves"], d["grass_restart"]) == (400, 200, 10)
    d = v2s(1600)
    assert (d["init_sheep"], d["init_wolves"]) == (409600, 204800)
    assert reserve_for(d) == 1600 * 1600          # one agent per cell
    assert reserve_for(v1(100)) == 10000


def test_batch_draw_is_fixed_and_covers_the_sizes():
    a = batch_replicas(64)
    b = batch_replicas(64)
    assert [(k, L, s) for k, L, _, s in a] == [(k, L, s) for k, L, _, s in b]
    assert [s for *_, s in a] == list(range(1, 65))
    sides = collections.Counter(L for _, L, _, _ in a)
    assert set(sides) <= set(BATCH_SIZES) and len(sides) == len(BATCH_SIZES)
    kinds = collections.Counter(k for k, *_ in a)
    assert 20 <= kinds["v1"] <= 44


def test_request_stream_is_poisson_at_the_rate():
    s = serve_requests(rate=5.0, duration=2000.0, draw_key=3)
    t = np.array([r["arrival"] for r in s])
    assert np.all(np.diff(t) > 0) and t[-1] < 2000.0
    assert abs(len(s) / 2000.0 - 5.0) < 0.3          # 10000 +- ~100
    gaps = np.diff(t)
    assert abs(gaps.mean() - 0.2) < 0.01 and abs(gaps.std() - 0.2) < 0.015
    sides = collections.Counter(r["L"] for r in s)
    assert set(sides) == set(SERVE_SIZES)
    its = np.array([r["params"]["iters"] for r in s])
    assert its.min() >= 1000 and its.max() <= 4000


def test_server_batch_fits_its_budget():
    from pphpc_amd import Batch
    from pphpc_amd.serve import batch_limits
    for gib in (1, 8, 48):
        a, c = batch_limits(gib << 30)
        assert Batch.footprint(a, c, 256, 15) <= gib << 30
        assert footprint_bytes(v2s(1600)) < (gib << 30)
