"""The alpha / beta sweep of the hybrid BFS (scripts/hybrid-advance.sh:38-88)
on this engine.

    python -m sygraph_amd.hybrid_sweep [--scale 21] [--sources 10] [--out FILE]

The reference's script runs its bfs example with `--advance hybrid` for every
alpha in {1, 5, 10, ..., 40} and beta in {1, 10, ..., 10^8}, ten fixed sources
per graph, appending the `-V` run details to a log per graph.  Its graphs
(SuiteSparse / DIMACS files) are not available here, so the sweep runs on the
generated Kronecker graph that stands for kron_g500-logn21 in its list: RMAT
scale 21 (by default), edge factor 16, a/b/c = .57/.19/.19, symmetrized, ten
sources drawn from the vertices with arcs (seed = the scale).  Per (alpha,
beta) it reports the device time of the ten traversals (hipEvents on the
library stream, best of two), their GTEPS (reached arcs / time), the levels
pulled, and checks that every depth array equals plain push's (device
digest).  The grid and the best pair go to the JSON file and a table to
stdout.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import sygraph_amd as sg  # noqa: E402

ALPHAS = [1, 5, 10, 15, 20, 25, 30, 35, 40]
BETAS = [1, 10, 100, 1000, 10000, 100000, 1000000, 10000000, 100000000]


def sweep(dev, G, sources, alphas=ALPHAS, betas=BETAS, reps=2):
    """Rows of {alpha, beta, ms, gteps, n_pull}; raises if any depth array
    differs from plain push's."""
    b = sg.BFS(G)
    ref = {}
    for s in sources:
        b.init(s)
        b.run(sg.POLICY_PUSH)
        ref[s] = b.depth.digest(G.n, 4)
    e0, e1 = sg.Event(dev), sg.Event(dev)
    rows = []
    for alpha in alphas:
        for beta in betas:
            opt = sg.bfs_options(sg.POLICY_HYBRID, sg.LB_BUCKETING, float(alpha), float(beta))
            best, edges, n_pull = None, 0, 0
            for _ in range(reps):
                st = sg.BfsStats()
                edges, n_pull = 0, 0
                dev.synchronize()
                e0.record()
                for s in sources:
                    sg.check(sg.lib().sg_bfs(G.h, s, C.byref(opt), C.c_void_p(b.depth.ptr), C.byref(st)))
                    edges += st.reached_edges
                    n_pull += st.n_pull
                e1.record()
                dev.synchronize()
                ms = e0.elapsed_ms(e1)
                best = ms if best is None else min(best, ms)
            for s in sources:  # the depths once more, one source at a time, against plain push
                b.init(s)
                b.run(sg.POLICY_HYBRID, float(alpha), float(beta))
                if b.depth.digest(G.n, 4) != ref[s]:
                    raise RuntimeError(f"alpha {alpha} beta {beta} source {s}: depths differ from plain push")
            rows.append({"alpha": alpha, "beta": beta, "ms": round(best, 4),
                         "gteps": round(edges / (best / 1e3) / 1e9, 2), "n_pull": n_pull})
    return rows


def table(rows, alphas=ALPHAS, betas=BETAS) -> str:
    """GTEPS as a grid: one line per alpha, one column per beta."""
    by = {(r["alpha"], r["beta"]): r for r in rows}
    out = ["alpha\\beta " + " ".join(f"{b:>8g}" for b in betas)]
    for a in alphas:
        out.append(f"{a:>10d} " + " ".join(f"{by[(a, b)]['gteps']:>8.1f}" for b in betas))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=21)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=10)
    ap.add_argument("--out", default=None, help="JSON file for the grid (default: stdout only)")
    a = ap.parse_args()
    dev = sg.Device(0)
    G = sg.Graph.rmat(dev, a.scale, a.edge_factor, 0.57, 0.19, 0.19, seed=a.scale)
    G.prepare()
    off, _, _ = G.download()
    cand = np.nonzero(np.diff(off) > 0)[0]
    sources = [int(x) for x in np.random.default_rng(a.scale).choice(cand, a.sources, replace=False)]
    rows = sweep(dev, G, sources)
    best = max(rows, key=lambda r: r["gteps"])
    res = {"graph": f"RMAT scale {a.scale} edge factor {a.edge_factor} a/b/c .57/.19/.19 seed {a.scale}",
           "vertices": G.n, "arcs": G.m, "sources": sources, "alphas": ALPHAS, "betas": BETAS, "grid": rows,
           "best": best}
    print(table(rows))
    print(f"best: alpha {best['alpha']} beta {best['beta']}: {best['gteps']} GTEPS ({best['ms']} ms for "
          f"{len(sources)} sources)")
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    G.destroy()


if __name__ == "__main__":
    main()
