"""BFS on the held-out graph set: plain, heuristic, learned and batched.

    python -m sygraph_amd.heldout_bfs [--quick] [--out sygraph_amd/data/bfs_heldout.json]

The graphs and sources are the held-out set of the BFS policy (seeds 300-349,
train_policy.make_graph / sources_of: four sources per graph).  Per graph the
device time (hipEvents on the library stream, best of three) of traversing
from all four sources under

  plain            push with the bucketing balancer, one source after
                   another (the reference's plain method on this engine);
  heuristic        the reference's direction rule (alpha 15, beta 24), one
                   source after another;
  learned          the shipped learned policy, one source after another;
  batched          the four sources as bit-lanes of one multi-source BFS
                   (sg_bfs_multi; its default direction rule);
  learned_batched  the same batch with the learned policy choosing each
                   level's direction on the batch's totals (sg_bfs_multi_ex).

The depths of every source are compared across the four (device digests) and
must be equal.  Writes the per-graph times and the totals / geometric means.
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
from sygraph_amd.train_policy import make_graph, sources_of  # noqa: E402

KEYS = ["plain", "heuristic", "learned", "batched", "learned_batched"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _timed(dev, fn, reps=3):
    e0, e1 = sg.Event(dev), sg.Event(dev)
    best = None
    for _ in range(reps):
        dev.synchronize()
        e0.record()
        fn()
        e1.record()
        dev.synchronize()
        ms = e0.elapsed_ms(e1)
        best = ms if best is None else min(best, ms)
    return best


def evaluate(dev, pol, seeds, quick):
    rows = []
    for seed in seeds:
        name, G = make_graph(dev, seed, quick)
        srcs = sources_of(G, seed)
        b = sg.BFS(G)
        digests = {}
        t = {}
        for key, kw in [("plain", dict(policy=sg.POLICY_PUSH)), ("heuristic", dict(policy=sg.POLICY_HYBRID)),
                        ("learned", dict(policy=sg.POLICY_LEARNED, learned=pol))]:
            opt = sg.bfs_options(kw["policy"], sg.LB_BUCKETING, 15.0, 24.0, learned=kw.get("learned"))

            def run_all():
                for s in srcs:
                    sg.check(sg.lib().sg_bfs(G.h, s, C.byref(opt), C.c_void_p(b.depth.ptr), None))
            b.init(srcs[0])
            t[key] = _timed(dev, run_all)
            for s in srcs:  # the depths of each source, once more, for the comparison
                b.init(s)
                b.run(alpha=15.0, beta=24.0, **kw)
                digests.setdefault(s, set()).add(b.depth.digest(G.n, 4))
        src_arr = np.ascontiguousarray(np.asarray(srcs, dtype=np.int32))
        out = sg.DeviceBuffer(dev, 4 * len(srcs) * G.n, dtype=np.int32)
        for key, mo in (("batched", sg.msbfs_options()),
                        ("learned_batched", sg.msbfs_options(sg.POLICY_LEARNED, learned=pol))):
            def run_batched():
                sg.check(sg.lib().sg_bfs_multi_ex(G.h, len(srcs), src_arr.ctypes.data_as(C.c_void_p), C.byref(mo),
                                                  C.c_void_p(out.ptr)))
            t[key] = _timed(dev, run_batched)
            for i, s in enumerate(srcs):
                h = C.c_uint64()
                sg.check(sg.lib().sg_digest(dev.h, C.c_void_p(out.ptr + 4 * i * G.n), G.n, 4, C.byref(h)))
                digests[s].add(h.value)
        if any(len(d) != 1 for d in digests.values()):
            raise RuntimeError(f"seed {seed}: depths differ between the methods")
        out.free()
        rows.append({"seed": seed, "graph": name, "sources": srcs, **t})
        log(f"  seed {seed}: {name}: " + " ".join(f"{k} {t[k]:.2f}" for k in KEYS) + " ms")
        G.destroy()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "data", "bfs_heldout.json"))
    ap.add_argument("--quick", action="store_true", help="smaller graphs and set (smoke run)")
    ap.add_argument("--policy", default=os.path.join(HERE, "data", "bfs_policy.f32"))
    a = ap.parse_args()
    dev = sg.Device(0)
    pol = sg.Policy(dev, np.fromfile(a.policy, dtype=np.float32))
    seeds = list(range(300, 350)) if not a.quick else list(range(300, 306))
    rows = evaluate(dev, pol, seeds, a.quick)
    tot = {k: float(sum(r[k] for r in rows)) for k in KEYS}
    geo = lambda num, den: float(np.exp(np.mean([np.log(r[num] / r[den]) for r in rows])))  # noqa: E731
    summary = {"total_ms": tot, "learned_speedup_vs_heuristic_total": tot["heuristic"] / tot["learned"],
               "learned_speedup_vs_heuristic_geomean": geo("heuristic", "learned")}
    for k in ("heuristic", "learned", "batched", "learned_batched"):
        summary[f"{k}_speedup_vs_plain_total"] = tot["plain"] / tot[k]
        summary[f"{k}_speedup_vs_plain_geomean"] = geo("plain", k)
    log(json.dumps(summary, indent=1))
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump({"seeds": [seeds[0], seeds[-1]], "quick": a.quick, "summary": summary, "runs": rows}, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
