"""Replica statistics: PPHPC's focal measures and two-sample tests.

Follows the model-comparison methodology of the PPHPC reference (the paper
cited at netlogo/pphpc.nlogo:599): each replica's six outputs -- sheep,
wolves, grass alive, mean sheep energy, mean wolf energy, mean grass
countdown (the columns of pp_stats_save, opencl/pp/pp_common.c:217-224) --
are reduced to focal measures:

  max, argmax, min, argmin     over iterations 0..m
  steady-state mean and std    over iterations l+1..m, l = STEADY_STATE_FROM

and each focal measure is compared between two samples of replicas with
Welch's t-test when Shapiro-Wilk accepts normality of both samples (alpha
0.05), else the Mann-Whitney U test. The family of 36 comparisons is
corrected with Holm's step-down procedure.

The truncation point l = 1000 (of 4000 iterations) is not in the reference
repository; it is the steady-state start used by the methodology and is
stated here explicitly.
"""
from __future__ import annotations

import math
from typing import Dict, Sequence

import numpy as np

STEADY_STATE_FROM = 1000
OUTPUTS = ("sheep", "wolves", "grass", "sheep_energy", "wolf_energy",
           "countdown")
MEASURES = ("max", "argmax", "min", "argmin", "ss_mean", "ss_std")


def outputs_from_rows(rows: np.ndarray, cells: int) -> np.ndarray:
    """(iters+1, 6) float array of the six outputs from (iters+1, 7) stats
    rows of exact sums (sheep, wolves, grass, sheep_en, wolves_en, grass_en,
    errors); a mean over an empty population is 0."""
    r = np.asarray(rows, dtype=np.float64)
    out = np.empty((r.shape[0], 6))
    out[:, 0] = r[:, 0]
    out[:, 1] = r[:, 1]
    out[:, 2] = r[:, 2]
    with np.errstate(invalid="ignore", divide="ignore"):
        out[:, 3] = np.where(r[:, 0] > 0, r[:, 3] / np.maximum(r[:, 0], 1), 0.0)
        out[:, 4] = np.where(r[:, 1] > 0, r[:, 4] / np.maximum(r[:, 1], 1), 0.0)
    out[:, 5] = r[:, 5] / float(cells)
    return out


def focal_measures(rows: np.ndarray, cells: int,
                   steady_from: int = STEADY_STATE_FROM) -> Dict[str, float]:
    """The 36 focal measures of one replica, keyed 'output.measure'."""
    o = outputs_from_rows(rows, cells)
    if o.shape[0] <= steady_from + 1:
        raise ValueError("run shorter than the steady-state start")
    ss = o[steady_from + 1:]
    fm = {}
    for k, name in enumerate(OUTPUTS):
        col = o[:, k]
        fm[f"{name}.max"] = float(col.max())
        fm[f"{name}.argmax"] = float(np.argmax(col))
        fm[f"{name}.min"] = float(col.min())
        fm[f"{name}.argmin"] = float(np.argmin(col))
        fm[f"{name}.ss_mean"] = float(ss[:, k].mean())
        fm[f"{name}.ss_std"] = float(ss[:, k].std(ddof=1))
    return fm


# ----------------------------------------------------------------- tests

def _betacf(a: float, b: float, x: float) -> float:
    """Continued fraction of the incomplete beta function (modified Lentz)."""
    tiny = 1e-300
    qab, qap, qam = a + b, a + 1.0, a - 1.0
    c, d = 1.0, 1.0 - qab * x / qap
    d = 1.0 / (d if abs(d) > tiny else tiny)
    h = d
    for m in range(1, 400):
        m2 = 2 * m
        aa = m * (b - m) * x / ((qam + m2) * (a + m2))
        d = 1.0 + aa * d
        d = 1.0 / (d if abs(d) > tiny else tiny)
        c = 1.0 + aa / c
        c = c if abs(c) > tiny else tiny
        h *= d * c
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2))
        d = 1.0 + aa * d
        d = 1.0 / (d if abs(d) > tiny else tiny)
        c = 1.0 + aa / c
        c = c if abs(c) > tiny else tiny
        de = d * c
        h *= de
        if abs(de - 1.0) < 1e-15:
            break
    return h


def betainc_reg(a: float, b: float, x: float) -> float:
    """Regularized incomplete beta I_x(a, b)."""
    if x <= 0.0:
        return 0.0
    if x >= 1.0:
        return 1.0
    lbt = (math.lgamma(a + b) - math.lgamma(a) - math.lgamma(b)
           + a * math.log(x) + b * math.log1p(-x))
    if x < (a + 1.0) / (a + b + 2.0):
        return math.exp(lbt) * _betacf(a, b, x) / a
    return 1.0 - math.exp(lbt) * _betacf(b, a, 1.0 - x) / b


def welch_t(a: Sequence[float], b: Sequence[float]):
    """Welch's unequal-variance t-test: (t, df, two-sided p)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    va, vb = a.var(ddof=1) / len(a), b.var(ddof=1) / len(b)
    if va + vb == 0.0:
        return 0.0, float(len(a) + len(b) - 2), \
            (1.0 if a.mean() == b.mean() else 0.0)
    t = (a.mean() - b.mean()) / math.sqrt(va + vb)
    df = (va + vb) ** 2 / (va ** 2 / (len(a) - 1) + vb ** 2 / (len(b) - 1))
    p = betainc_reg(df / 2.0, 0.5, df / (df + t * t))
    return float(t), float(df), float(p)


def mann_whitney(a: Sequence[float], b: Sequence[float]):
    """Mann-Whitney U test, normal approximation with tie and continuity
    corrections: (U of the first sample, two-sided p)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    n1, n2 = len(a), len(b)
    allv = np.concatenate([a, b])
    order = np.argsort(allv, kind="mergesort")
    ranks = np.empty(len(allv))
    sv = allv[order]
    i = 0
    ties = 0.0
    while i < len(sv):
        j = i
        while j + 1 < len(sv) and sv[j + 1] == sv[i]:
            j += 1
        ranks[order[i:j + 1]] = (i + j) / 2.0 + 1.0
        t = j - i + 1
        ties += t ** 3 - t
        i = j + 1
    u1 = ranks[:n1].sum() - n1 * (n1 + 1) / 2.0
    mu = n1 * n2 / 2.0
    n = n1 + n2
    sigma2 = n1 * n2 / 12.0 * ((n + 1) - ties / (n * (n - 1)))
    if sigma2 <= 0:
        return float(u1), 1.0
    z = (abs(u1 - mu) - 0.5) / math.sqrt(sigma2)
    p = math.erfc(max(z, 0.0) / math.sqrt(2.0))
    return float(u1), float(min(1.0, p))


def _normal(x) -> bool:
    from scipy import stats
    x = np.asarray(x, float)
    if np.ptp(x) == 0:
        return False
    return bool(stats.shapiro(x).pvalue > 0.05)


def compare(sample_a: Sequence[Dict[str, float]],
            sample_b: Sequence[Dict[str, float]], alpha: float = 0.05):
    """Compares two samples of focal-measure dicts. Returns a report with
    one entry per focal measure (test, p, reject after Holm) and 'pass' =
    no rejection."""
    keys = list(sample_a[0].keys())
    rows = []
    for k in keys:
        a = [s[k] for s in sample_a]
        b = [s[k] for s in sample_b]
        if _normal(a) and _normal(b):
            _, _, p = welch_t(a, b)
            test = "welch_t"
        else:
            _, p = mann_whitney(a, b)
            test = "mann_whitney"
        rows.append({"measure": k, "test": test, "p": p,
                     "mean_a": float(np.mean(a)), "mean_b": float(np.mean(b))})
    # Holm step-down over the family
    order = sorted(range(len(rows)), key=lambda i: rows[i]["p"])
    m = len(rows)
    rejecting = True
    for rank, i in enumerate(order):
        thr = alpha / (m - rank)
        rejecting = rejecting and rows[i]["p"] <= thr
        rows[i]["holm_threshold"] = thr
        rows[i]["reject"] = bool(rejecting)
    raw_rejects = sum(r["p"] <= alpha for r in rows)
    return {"pass": not any(r["reject"] for r in rows),
            "raw_rejections": int(raw_rejects), "alpha": alpha,
            "measures": rows}
