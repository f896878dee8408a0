"""Statistical equivalence at the benchmark size: 30 GPU replicas of the dense
v2-style 1600 x 1600 workload, 4000 iterations each, against 30 replicas of
the independent simulator of the Java single-thread rules (committed focal
measures, tests/golden/statref_v2s1600.json, written by
tests/golden/gen_statref.py).

Following PPHPC's replica methodology: per replica the 36 focal measures
(max/argmax, min/argmin, steady-state mean and std of the six outputs), per
measure Welch's t-test or Mann-Whitney U (chosen by Shapiro-Wilk), the family
corrected with Holm's procedure at alpha 0.05. The 30 GPU replicas run as one
batch.
"""
import json
import os
import sys

import pytest

from helpers import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(GOLDEN, "statref_v2s1600.json")
NETLOGO = os.path.join(GOLDEN, "statref_netlogo_v2s1600.json")


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="fixture not written")
def test_v2s_1600_replicas_equivalent_to_reference_rules():
    from pphpc_amd import Batch, Params
    from pphpc_amd.replicas import compare, focal_measures
    from pphpc_amd.workloads import reserve_for
    fix = json.load(open(FIXTURE))
    d = dict(fix["params"])
    assert d["iters"] == fix["iters"] == 4000
    n = len(fix["focal"])
    assert n >= 30
    cells = d["grid_x"] * d["grid_y"]
    _case(f"{n} GPU replicas of 1600x1600 v2s, 4000 iterations, one batch")
    with Batch(n * reserve_for(d), n * cells, n, d["grass_restart"],
               device=0) as b:
        ids = [b.add(Params.from_values(**d), seed, reserve_for(d))
               for seed in range(1, n + 1)]
        while b.resources()["active"]:
            b.run(512)
        focal = []
        for rid in ids:
            st = b.replica(rid)
            assert st["state"] == 2 and st["steps"] == 4000, st
            focal.append(focal_measures(b.rows_array(rid), cells,
                                        fix["steady_from"]))
    assert fix["rules"] == "cells"
    rep = compare(focal, fix["focal"])
    rejected = [m["measure"] for m in rep["measures"] if m["reject"]]
    print(f"[stats] Java rules: {rep['raw_rejections']} of "
          f"{len(rep['measures'])} measures below alpha before Holm; "
          f"rejected after Holm: {rejected}", flush=True)
    report = {"workload": "v2s 1600x1600, 4000 iterations, seeds 1..%d" % n,
              "steady_from": fix["steady_from"], "java_rules": rep}
    if os.path.exists(NETLOGO):
        # Reported, not asserted: the NetLogo rule set differs from the Java
        # one in the wolf's choice of prey, which 30 replicas at this size
        # resolve (steady-state wolf count and energy).
        nl = compare(focal, json.load(open(NETLOGO))["focal"])
        report["netlogo_rules"] = nl
        print(f"[stats] NetLogo rules: rejected after Holm: "
              f"{[m['measure'] for m in nl['measures'] if m['reject']]}",
              flush=True)
    out_dir = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out_dir):
        with open(os.path.join(out_dir, "stats_report_v2s1600.json"), "w") as f:
            json.dump(report, f, indent=1)
    assert rep["pass"], rejected
