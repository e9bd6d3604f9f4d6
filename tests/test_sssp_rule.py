"""CPU pins of the SSSP automatic schedule: the fits of
sygraph_amd/heldout_sssp.py (the rule, the bucket width) on hand-worked cases,
the committed fit records (sygraph_amd/data/sssp_rule_fit.json,
sssp_delta_fit.json) against the constants sssp.hip uses, and the held-out
comparison record (sygraph_amd/data/sssp_heldout.json)."""
import json
import math
import os
import re

import pytest

from sygraph_amd import heldout_sssp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "sygraph_amd", "data")


def _row(n, m, plain, bucket_pull):
    return {"n": n, "m": m, "plain": plain, "bucket_pull": bucket_pull}


def test_fit_hand_worked_case():
    # A and C are faster with buckets, B (3 arcs per vertex) with the
    # reference's schedule; every rule with M <= 100 and 4 <= D <= 10 picks
    # the faster schedule everywhere, and the tie goes to the larger M, D
    rows = [_row(10, 100, 5.0, 3.0), _row(100, 300, 2.0, 9.0), _row(10, 1000, 9.0, 4.0)]
    r = heldout_sssp.fit(rows, arcs=[0, 100, 1000], degree=[0, 4, 10])
    assert (r["best"]["min_arcs"], r["best"]["min_degree"]) == (100, 10)
    assert r["best"]["geomean_vs_best"] == pytest.approx(1.0)
    assert r["best"]["total_ms"] == pytest.approx(3.0 + 2.0 + 4.0)
    assert r["plain_total_ms"] == pytest.approx(16.0)
    assert r["per_run_best_total_ms"] == pytest.approx(9.0)
    assert r["plain_geomean_vs_best"] == pytest.approx(math.exp((math.log(5 / 3) + math.log(9 / 4)) / 3))
    assert len(r["grid"]) == 9
    # (0, 0): buckets everywhere -> B at 9 / 2
    g00 = next(c for c in r["grid"] if c["min_arcs"] == 0 and c["min_degree"] == 0)
    assert g00["geomean_vs_best"] == pytest.approx((9 / 2) ** (1 / 3))


def test_fit_objective_is_the_geometric_mean_not_the_total():
    # buckets everywhere saves 10 ms on Y and loses 1 ms (2x) on X: least
    # total time, but the geometric mean prefers the reference's schedule
    rows = [_row(1, 10, 1.0, 2.0), _row(1, 20, 100.0, 90.0)]
    r = heldout_sssp.fit(rows, arcs=[0, 30], degree=[0])
    assert r["best"]["min_arcs"] == 30
    by = {c["min_arcs"]: c for c in r["grid"]}
    assert by[0]["total_ms"] < by[30]["total_ms"]
    assert by[30]["geomean_vs_best"] == pytest.approx(math.sqrt(100 / 90))


def test_shipped_rule_is_the_recorded_fit():
    """The committed record of `python -m sygraph_amd.heldout_sssp --fit`
    (training graphs 100-199, four sources each) refits to the pair it
    records, and that pair is what the automatic schedule in sssp.hip uses."""
    with open(os.path.join(DATA, "sssp_rule_fit.json")) as f:
        rec = json.load(f)
    assert tuple(rec["seeds"]) == (100, 199) and not rec["quick"]
    assert len(rec["runs"]) == 400
    assert all(r["plain"] > 0 and r["bucket_pull"] > 0 for r in rec["runs"])
    best = heldout_sssp.fit(rec["runs"])["best"]
    assert (best["min_arcs"], best["min_degree"]) == (rec["best"]["min_arcs"], rec["best"]["min_degree"])
    with open(os.path.join(ROOT, "sygraph_amd", "csrc", "sssp.hip")) as f:
        src = f.read()
    arcs = int(re.search(r"kAutoBucketArcs = int64_t\(1\) << (\d+);", src).group(1))
    degree = int(re.search(r"kAutoBucketDegree = (\d+);", src).group(1))
    assert (1 << arcs, degree) == (best["min_arcs"], best["min_degree"])
    # better than the reference's schedule everywhere, by the fit's own measure
    assert best["geomean_vs_best"] < rec["plain_geomean_vs_best"]
    assert best["total_ms"] < rec["plain_total_ms"]


def test_shipped_held_out_record():
    """The held-out comparison (graphs 300-349, four sources each, the three
    schedules' distances compared on the device): the defaults at least as
    fast as the reference's schedule in total and in the geometric mean."""
    with open(os.path.join(DATA, "sssp_heldout.json")) as f:
        rec = json.load(f)
    assert tuple(rec["seeds"]) == (300, 349) and not rec["quick"]
    assert len(rec["runs"]) == 200
    tot = rec["summary"]["total_ms"]
    assert tot["adaptive"] <= tot["plain"]
    assert rec["summary"]["adaptive_speedup_vs_plain_geomean"] >= 1.0
    geo = math.exp(sum(math.log(r["plain"] / r["adaptive"]) for r in rec["runs"]) / len(rec["runs"]))
    assert geo == pytest.approx(rec["summary"]["adaptive_speedup_vs_plain_geomean"])


def test_fit_factor_hand_worked_case():
    rows = [{"f0.25": 2.0, "f0.5": 3.0}, {"f0.25": 5.0, "f0.5": 4.0}]
    r = heldout_sssp.fit_factor(rows, [0.25, 0.5])
    assert r["best"]["factor"] == 0.25
    by = {c["factor"]: c for c in r["candidates"]}
    assert by[0.25]["geomean_vs_best"] == pytest.approx(math.sqrt(1.25))
    assert by[0.5]["geomean_vs_best"] == pytest.approx(math.sqrt(1.5))
    assert by[0.5]["total_ms"] == pytest.approx(7.0)


def test_shipped_bucket_width_is_the_recorded_fit():
    """The committed record of `python -m sygraph_amd.heldout_sssp
    --fit-delta` (the training graphs the rule sends to buckets, four sources
    each, widths f x the mean weight) refits to the factor it records, and
    sssp.hip's automatic width uses that factor."""
    with open(os.path.join(DATA, "sssp_delta_fit.json")) as f:
        rec = json.load(f)
    assert tuple(rec["seeds"]) == (100, 199) and not rec["quick"]
    assert len(rec["runs"]) >= 40
    assert all(r["m"] >= heldout_sssp.RULE_ARCS and r["m"] >= heldout_sssp.RULE_DEGREE * r["n"] for r in rec["runs"])
    best = heldout_sssp.fit_factor(rec["runs"])["best"]
    assert best["factor"] == rec["best"]["factor"]
    with open(os.path.join(ROOT, "sygraph_amd", "csrc", "sssp.hip")) as f:
        src = f.read()
    factor = float(re.search(r"kAutoDeltaFactor = ([0-9.]+);", src).group(1))
    assert factor == best["factor"]
    arcs = int(re.search(r"kAutoBucketArcs = int64_t\(1\) << (\d+);", src).group(1))
    degree = int(re.search(r"kAutoBucketDegree = (\d+);", src).group(1))
    assert (1 << arcs, degree) == (heldout_sssp.RULE_ARCS, heldout_sssp.RULE_DEGREE)
