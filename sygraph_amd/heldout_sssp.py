"""SSSP on the held-out graph set: the reference's schedule against this build's.

    python -m sygraph_amd.heldout_sssp [--quick] [--out sygraph_amd/data/sssp_heldout.json]

The graphs are the held-out set of the BFS policy (seeds 300-349, families and
sizes as in train_policy.make_graph), each carrying u32 weights U[1,255]
(weight seed = graph seed), from the same four sources per graph.  Per run the
device time of the whole SSSP (best of three) under

  plain     frontier Bellman-Ford, every round a push (delta < 0: the schedule
            of algorithms/sssp.hpp:152-210);
  buckets   delta-stepping with width 64 (half the mean weight), heavy rounds
            pushing;
  adaptive  the defaults: the automatic schedule (buckets of half the mean
            weight on graphs of at least 2^25 arcs and 8 arcs per vertex,
            plain otherwise) with heavy rounds pulling when cheaper.

The three distance arrays of a run are compared on the device (sg_digest) and
must be equal.  Writes the per-run times and the totals / geometric means.

    python -m sygraph_amd.heldout_sssp --fit [--seeds 100-199] [--out sygraph_amd/data/sssp_rule_fit.json]

fits the automatic schedule instead, on the training graphs (seeds 100-199 by
default): per run it times `plain` against `bucket_pull` (buckets of half the
mean weight with the automatic heavy-round direction, what the automatic
schedule runs when it chooses buckets), and `fit` picks the rule "buckets when
the graph has at least M arcs and at least D arcs per vertex" with the least
geometric mean over the runs of the rule's time / the faster schedule's time
(every graph weighs the same; the total time is dominated by the largest grids
and barely tells arc thresholds apart; ties: the larger M, then the larger D).

    python -m sygraph_amd.heldout_sssp --fit-delta [--out sygraph_amd/data/sssp_delta_fit.json]

fits the bucket width the automatic schedule uses, on the training graphs that
the rule sends to buckets: per run it times widths of f x 128 (128 = the mean
of weights U[1,255]) for each f in FIT_FACTORS, heavy rounds automatic, and
`fit_factor` picks the f with the least geometric mean of time / the run's
best time (ties: the larger f).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import sygraph_amd as sg  # noqa: E402
from sygraph_amd.train_policy import make_graph, sources_of  # noqa: E402

MODES = [("plain", dict(delta=-1.0, heavy_pull=-1)), ("buckets", dict(delta=64.0, heavy_pull=-1)),
         ("adaptive", dict(delta=0.0, heavy_pull=0))]
FIT_MODES = [("plain", dict(delta=-1.0, heavy_pull=-1)), ("bucket_pull", dict(delta=64.0, heavy_pull=0))]
FIT_ARCS = [0] + [1 << k for k in range(18, 29)]
FIT_DEGREE = [0, 2, 4, 6, 8, 10, 12, 16, 24, 32]
FIT_FACTORS = [0.125, 0.25, 0.375, 0.5, 0.75, 1.0]
MEAN_WEIGHT = 128.0  # of weights U[1,255]
RULE_ARCS, RULE_DEGREE = 1 << 25, 8  # the automatic schedule's rule (sssp.hip)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def evaluate(dev, seeds, quick, nsrc=4, reps=3, modes=MODES, keep=None):
    """keep(n, m): whether to time a graph (default: all)."""
    rows = []
    for seed in seeds:
        name, G = make_graph(dev, seed, quick, weight_seed=seed)
        if keep is not None and not keep(G.n, G.m):
            G.destroy()
            continue
        s = sg.SSSP(G)
        for src in sources_of(G, seed, nsrc):
            t, digests = {}, set()
            for key, kw in modes:
                best = None
                for _ in range(reps):
                    s.init(src)
                    st = s.run(**kw)
                    best = st.kernel_ms if best is None else min(best, st.kernel_ms)
                t[key] = best
                digests.add(s.dist.digest(G.n, 4))
            if len(digests) != 1:
                raise RuntimeError(f"seed {seed} source {src}: distances differ between schedules")
            rows.append({"seed": seed, "graph": name, "n": G.n, "m": G.m, "source": src, **t})
        last = rows[-1]
        log(f"  seed {seed}: {name}: " + " ".join(f"{k} {last[k]:.2f}" for k, _ in modes) + " ms")
        G.destroy()
    return rows


def fit(rows, arcs=FIT_ARCS, degree=FIT_DEGREE):
    """The (M, D) of the rule "buckets when m >= M and m / n >= D" with the
    least geometric mean of time / best time over `rows` (plain / bucket_pull
    times per run); every candidate with that mean and its total time, and
    the totals of plain everywhere and of the best choice per run."""
    def pick(r, M, D):
        return r["bucket_pull"] if r["m"] >= M and r["m"] >= D * r["n"] else r["plain"]

    def slowdown(M, D):
        return float(np.exp(np.mean([np.log(pick(r, M, D) / min(r["plain"], r["bucket_pull"])) for r in rows])))
    grid = [{"min_arcs": M, "min_degree": D, "geomean_vs_best": slowdown(M, D),
             "total_ms": sum(pick(r, M, D) for r in rows)} for M in arcs for D in degree]
    best = min(grid, key=lambda c: (round(c["geomean_vs_best"], 6), -c["min_arcs"], -c["min_degree"]))
    return {"best": best, "plain_total_ms": sum(r["plain"] for r in rows),
            "plain_geomean_vs_best": slowdown(1 << 62, 0),
            "per_run_best_total_ms": sum(min(r["plain"], r["bucket_pull"]) for r in rows), "grid": grid}


def factor_key(f: float) -> str:
    return f"f{f:g}"


def fit_factor(rows, factors=FIT_FACTORS):
    """The width factor with the least geometric mean of time / the run's best
    time over `rows` (one time per factor per run, keys factor_key(f)), and
    every factor's mean and total."""
    keys = [factor_key(f) for f in factors]
    best_of = [min(r[k] for k in keys) for r in rows]
    cand = [{"factor": f, "geomean_vs_best": float(np.exp(np.mean([np.log(r[k] / b) for r, b in zip(rows, best_of)]))),
             "total_ms": float(sum(r[k] for r in rows))} for f, k in zip(factors, keys)]
    best = min(cand, key=lambda c: (round(c["geomean_vs_best"], 6), -c["factor"]))
    return {"best": best, "candidates": cand}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "data", "sssp_heldout.json"))
    ap.add_argument("--quick", action="store_true", help="smaller graphs and set (smoke run)")
    ap.add_argument("--seeds", default="300-349", help="graph seeds, first-last (training graphs: 100-299)")
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--fit", action="store_true", help="fit the automatic schedule's rule (training graphs)")
    ap.add_argument("--fit-delta", action="store_true", help="fit the automatic bucket width (training graphs)")
    a = ap.parse_args()
    if (a.fit or a.fit_delta) and a.seeds == "300-349":
        a.seeds = "100-199"
    if a.fit and a.out == ap.get_default("out"):
        a.out = os.path.join(HERE, "data", "sssp_rule_fit.json")
    if a.fit_delta and a.out == ap.get_default("out"):
        a.out = os.path.join(HERE, "data", "sssp_delta_fit.json")
    dev = sg.Device(0)
    lo, hi = (int(x) for x in a.seeds.split("-"))
    seeds = list(range(lo, hi + 1)) if not a.quick else list(range(lo, min(hi, lo + 5) + 1))
    if a.fit_delta:
        modes = [(factor_key(f), dict(delta=f * MEAN_WEIGHT, heavy_pull=0)) for f in FIT_FACTORS]
        rows = evaluate(dev, seeds, a.quick, a.sources, a.reps, modes=modes,
                        keep=lambda n, m: m >= RULE_ARCS and m >= RULE_DEGREE * n)
        res = fit_factor(rows)
        log(json.dumps(res, indent=1))
        with open(a.out, "w") as f:
            json.dump({"seeds": [seeds[0], seeds[-1]], "quick": a.quick,
                       "weights": "u32 U[1,255], weight seed = graph seed", **res, "runs": rows}, f, indent=1)
        print(json.dumps(res["best"]))
        return
    if a.fit:
        rows = evaluate(dev, seeds, a.quick, a.sources, a.reps, modes=FIT_MODES)
        res = fit(rows)
        log(json.dumps({k: v for k, v in res.items() if k != "grid"}, indent=1))
        with open(a.out, "w") as f:
            json.dump({"seeds": [seeds[0], seeds[-1]], "quick": a.quick,
                       "weights": "u32 U[1,255], weight seed = graph seed", **res, "runs": rows}, f, indent=1)
        print(json.dumps(res["best"]))
        return
    rows = evaluate(dev, seeds, a.quick, a.sources, a.reps)
    tot = {k: float(sum(r[k] for r in rows)) for k, _ in MODES}
    geo = lambda num, den: float(np.exp(np.mean([np.log(r[num] / r[den]) for r in rows])))  # noqa: E731
    summary = {"total_ms": tot,
               "adaptive_speedup_vs_plain_total": tot["plain"] / tot["adaptive"],
               "adaptive_speedup_vs_plain_geomean": geo("plain", "adaptive"),
               "adaptive_speedup_vs_buckets_total": tot["buckets"] / tot["adaptive"],
               "adaptive_speedup_vs_buckets_geomean": geo("buckets", "adaptive")}
    log(json.dumps(summary, indent=1))
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump({"seeds": [seeds[0], seeds[-1]], "quick": a.quick, "weights": "u32 U[1,255], weight seed = graph seed",
                   "summary": summary, "runs": rows}, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
