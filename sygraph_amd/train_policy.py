"""Train the learned per-level BFS policy on the device and evaluate it.

    python -m sygraph_amd.train_policy [--out sygraph_amd/data] [--quick]

1. Graph sets.  Training: seeds 100-299; held-out: seeds 300-349.  The seed
   fixes the family and size: RMAT (scale 16-22, a in [0.45, 0.65],
   b = c = (0.95 - a) / 2 so d = 0.05, edge factor 4-32), 4-neighbour grids
   (side 512, 1024, 2048 or 4096, 0-20 % of edges deleted) and Erdos-Renyi
   graphs (2^16-2^21 vertices, average degree 4-64).  All are generated on
   the device.
2. Per-level costs.  From four non-isolated sources per graph, the BFS runs
   once per action with that action forced at every level (push or pull x
   work mapping variant 0/1/2).  The frontier sequence of a BFS does not
   depend on the actions, so level L of every run sees the same state and
   cost(L, a) = traversal + conversion time of level L in the run forced to a
   (bursts off, so every level is timed on its own).
   Each level L >= 1 becomes one sample: the features the device computes
   when it decides level L (policy.hpp), labelled with argmin_a cost(L, a).
3. Training: minibatch Adam on the device (sg_policy_train_step) on
   samples resampled in proportion to the level's stake (slowest minus fastest
   action time), so the levels that dominate traversal time dominate the loss.
4. Held-out evaluation: the summed device time of the BFS under the plain
   policy (push, bucketing), the reference heuristic (alpha 15, beta 24), the
   learned policy, and the per-level best action (a lower bound), all with
   bursts off; the three policies also with the default bursts.

Writes <out>/bfs_policy.f32 (1,542 float32 parameters) and
<out>/bfs_policy.json (sets, sample counts, training loss, held-out times).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import sygraph_amd as sg  # noqa: E402

ACTIONS = [(sg.POLICY_PUSH, sg.LB_BUCKETING), (sg.POLICY_PUSH, sg.LB_WORKGROUP), (sg.POLICY_PUSH, sg.LB_SUBGROUP),
           (sg.POLICY_PULL, sg.LB_BUCKETING), (sg.POLICY_PULL, sg.LB_WORKGROUP), (sg.POLICY_PULL, sg.LB_SUBGROUP)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_graph(dev, seed: int, quick: bool, weight_seed: int = 0):
    """The graph of a seed (family and size drawn from the seed); with
    weight_seed != 0 it carries u32 weights U[1,255].  quick: smaller sizes
    of the same families (smoke runs)."""
    rng = np.random.default_rng(seed)
    kind = rng.choice(["rmat", "rmat", "rmat", "grid", "er"])
    if kind == "rmat":
        scale = int(rng.integers(16, (19 if quick else 22) + 1))
        ef = int(rng.integers(4, 33))
        a = float(rng.uniform(0.45, 0.65))
        bc = (0.95 - a) / 2.0
        return f"rmat s{scale} ef{ef} a{a:.3f}", sg.Graph.rmat(dev, scale, ef, a, bc, bc, seed=seed,
                                                              weight_seed=weight_seed)
    if kind == "grid":
        side = int(rng.choice([128, 256, 512] if quick else [512, 1024, 2048, 4096]))
        p = float(rng.uniform(0.0, 0.2))
        return f"grid {side}x{side} p{p:.2f}", sg.Graph.grid(dev, side, side, p, seed=seed, weight_seed=weight_seed)
    lg = int(rng.integers(16, (19 if quick else 21) + 1))
    deg = int(rng.integers(4, 65))
    n = 1 << lg
    return f"er 2^{lg} d{deg}", sg.Graph.erdos_renyi(dev, n, n * deg // 2, seed=seed, weight_seed=weight_seed)


def sources_of(G, seed: int, k: int = 4):
    off, _, _ = G.download()
    cand = np.nonzero(np.diff(off) > 0)[0]
    rng = np.random.default_rng(seed + 7)
    return [int(x) for x in rng.choice(cand, min(k, len(cand)), replace=False)]


NO_BURSTS = {"burst_arcs": -1}  # every level its own launches, so each level's time is its own


def level_costs(G, src: int):
    """cost[L][a] for every action a forced at every level, and the stats of
    the first run (the frontier sequence is the same in all runs).  Bursts
    are off: a burst times its levels together, so their costs could not be
    told apart."""
    b = sg.BFS(G)
    costs, stats0 = [], None
    for pol, lb in ACTIONS:
        b.init(src)
        st = b.run(pol, 15.0, 24.0, load_balancer=lb, tuning=NO_BURSTS)
        nrec = min(st.levels, sg.MAX_LEVELS_RECORDED)  # levels past the record limit give no samples
        c = [st.level[i].kernel_ms + st.level[i].convert_ms for i in range(nrec)]
        costs.append(c)
        if stats0 is None:
            stats0 = [(st.level[i].frontier_vertices, st.level[i].frontier_edges, st.level[i].unexplored_edges,
                       st.level[i].max_degree) for i in range(nrec)]
    L = min(len(c) for c in costs)
    return np.array([[costs[a][i] for a in range(len(ACTIONS))] for i in range(L)]), stats0[:L]


def level_features(stats0, n: int, m: int):
    """Features of the decision for level L (L >= 1), as policy.hpp computes
    them on the device: the frontier entering L, the arcs still unexplored
    after L-1, the largest degree discovered at L-1."""
    out = []
    for L in range(1, len(stats0)):
        fv, fe, _, _ = stats0[L]
        _, _, unexplored, maxdeg = stats0[L - 1]
        out.append([np.log2(1 + fv) / 24, np.log2(1 + fe) / 32, np.log2(1 + unexplored) / 32, np.log2(1 + L) / 12,
                    np.log2(1 + fe / max(fv, 1)) / 16, np.log2(1 + maxdeg) / 24, np.log2(1 + m / max(n, 1)) / 8,
                    np.log2(max(n, 1)) / 32])
    return np.array(out, dtype=np.float32).reshape(-1, 8)


MAX_LEVELS_PER_SOURCE = 1024  # long traversals (grids) give at most this many samples, spread evenly


def collect(dev, seeds, quick):
    X, Y, C = [], [], []
    for seed in seeds:
        t0 = time.time()
        name, G = make_graph(dev, seed, quick)
        for src in sources_of(G, seed):
            cost, stats0 = level_costs(G, src)
            if len(stats0) < 2:
                continue
            x, c = level_features(stats0, G.n, G.m), cost[1:]
            if len(c) > MAX_LEVELS_PER_SOURCE:
                keep = np.unique(np.linspace(0, len(c) - 1, MAX_LEVELS_PER_SOURCE).astype(np.int64))
                # each kept level stands for len(c) / len(keep) levels: its costs are scaled so
                # that the stake-weighted sampling of the training keeps the source's weight
                x, c = x[keep], c[keep] * (len(c) / len(keep))
            X.append(x)
            Y.append(np.argmin(c, axis=1).astype(np.int32))
            C.append(c)
        log(f"  seed {seed}: {name} n={G.n} m={G.m} ({time.time() - t0:.1f} s)")
        G.destroy()
    return np.concatenate(X), np.concatenate(Y), np.concatenate(C)


def train(dev, X, Y, C, steps: int, lr: float, batch: int = 8192, seed: int = 0):
    """Cost-sensitive training: samples are drawn in proportion to how much
    the choice matters at that level (slowest minus fastest action), so the
    loss is dominated by the levels that dominate the traversal time.  Adam on
    minibatches of a shuffled pool that lives on the device."""
    rng = np.random.default_rng(seed)
    stake = C.max(axis=1) - C.min(axis=1)
    prob = (stake + 1e-4) / (stake + 1e-4).sum()
    pool = rng.choice(len(Y), size=max(2 * len(Y), batch), replace=True, p=prob)
    Xb, Yb = dev.to_device(X[pool]), dev.to_device(Y[pool])
    scale = np.concatenate([np.full(256, np.sqrt(2 / 8)), np.zeros(32), np.full(1024, np.sqrt(2 / 32)), np.zeros(32),
                            np.full(192, np.sqrt(1 / 32)), np.zeros(6)])
    pol = sg.Policy(dev, (rng.normal(0, 1, sg.POLICY_PARAMS) * scale).astype(np.float32))
    losses = []
    nb = len(pool) // batch
    for step in range(steps):
        losses.append(pol.train_step(Xb, Yb, batch, lr, first=(step % nb) * batch))
        if step % 500 == 0 or step == steps - 1:
            log(f"  step {step}: loss {losses[-1]:.4f}")
    return pol, losses


def evaluate(dev, pol, seeds, quick):
    """Device time of whole traversals (best of three) under each policy,
    without bursts (the setting of the per-level costs, so per_level_best
    bounds them) and, as `<policy>_bursts`, with the default bursts."""
    rows = []
    for seed in seeds:
        name, G = make_graph(dev, seed, quick)
        b = sg.BFS(G)
        for src in sources_of(G, seed):
            t = {}
            for key, kw in [("plain", dict(policy=sg.POLICY_PUSH)), ("heuristic", dict(policy=sg.POLICY_HYBRID)),
                            ("learned", dict(policy=sg.POLICY_LEARNED, learned=pol))]:
                for suffix, tuning in (("", NO_BURSTS), ("_bursts", None)):
                    best = None
                    for _ in range(3):
                        b.init(src)
                        st = b.run(alpha=15.0, beta=24.0, tuning=tuning, **kw)
                        best = st.kernel_ms if best is None else min(best, st.kernel_ms)
                    t[key + suffix] = best
            cost, _ = level_costs(G, src)
            t["per_level_best"] = float(cost.min(axis=1).sum())
            rows.append({"seed": seed, "graph": name, "source": src, **t})
        last = rows[-1]
        log(f"  held-out seed {seed}: {name}: plain {last['plain']:.2f} heuristic {last['heuristic']:.2f} "
            f"learned {last['learned']:.2f} best {last['per_level_best']:.2f} ms")
        G.destroy()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "data"))
    ap.add_argument("--quick", action="store_true", help="smaller graphs and sets (smoke run)")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--candidates", type=int, default=4,
                    help="networks trained from different initialisations; the one with the smallest time "
                         "regret on the training levels is kept")
    ap.add_argument("--no-eval", action="store_true", help="skip the held-out evaluation")
    ap.add_argument("--samples", default=None, help="also save the training levels (npz) here")
    a = ap.parse_args()
    dev = sg.Device(0)
    train_seeds = list(range(100, 300)) if not a.quick else list(range(100, 116))
    test_seeds = list(range(300, 350)) if not a.quick else list(range(300, 306))
    t0 = time.time()
    log("collecting training levels ...")
    X, Y, C = collect(dev, train_seeds, a.quick)
    log(f"{len(Y)} samples, class counts {np.bincount(Y, minlength=6).tolist()} ({time.time() - t0:.1f}s)")
    if a.samples:
        np.savez(a.samples, X=X, Y=Y, C=C)
    best = None
    for cand in range(a.candidates):
        p_, l_ = train(dev, X, Y, C, a.steps, a.lr, seed=cand)
        pred = np.argmax(p_.forward(X), axis=1)
        reg = float((C[np.arange(len(Y)), pred] - C.min(axis=1)).sum() / C.min(axis=1).sum())
        log(f"candidate {cand}: loss {l_[0]:.3f} -> {l_[-1]:.3f}, time regret {reg:.3%}")
        if best is None or reg < best[0]:
            best = (reg, cand, p_, l_, pred)
    regret, chosen, pol, losses, pred = best
    acc = float((pred == Y).mean())
    log(f"train: candidate {chosen}, loss {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy {acc:.3f}, "
        f"time regret {regret:.3%}")
    os.makedirs(a.out, exist_ok=True)
    pol.get_params().astype(np.float32).tofile(os.path.join(a.out, "bfs_policy.f32"))
    meta = {"train_seeds": [train_seeds[0], train_seeds[-1]], "held_out_seeds": [test_seeds[0], test_seeds[-1]],
            "quick": a.quick, "samples": int(len(Y)), "class_counts": np.bincount(Y, minlength=6).tolist(),
            "steps": a.steps, "lr": a.lr, "candidates": a.candidates, "chosen": chosen,
            "loss_first": losses[0], "loss_last": losses[-1], "train_accuracy": acc, "train_time_regret": regret}
    if a.no_eval:
        with open(os.path.join(a.out, "bfs_policy.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print(json.dumps({k: meta[k] for k in ("samples", "train_accuracy", "train_time_regret")}))
        return
    rows = evaluate(dev, pol, test_seeds, a.quick)
    tot = {k: float(sum(r[k] for r in rows)) for k in ["plain", "heuristic", "learned", "per_level_best",
                                                        "plain_bursts", "heuristic_bursts", "learned_bursts"]}
    geo = lambda num, den: float(np.exp(np.mean([np.log(r[num] / r[den]) for r in rows])))  # noqa: E731
    summary = {"held_out_total_ms": tot,
               "learned_speedup_vs_plain_total": tot["plain"] / tot["learned"],
               "learned_speedup_vs_heuristic_total": tot["heuristic"] / tot["learned"],
               "learned_speedup_vs_plain_geomean": geo("plain", "learned"),
               "learned_speedup_vs_heuristic_geomean": geo("heuristic", "learned"),
               "with_bursts": {"learned_speedup_vs_plain_total": tot["plain_bursts"] / tot["learned_bursts"],
                               "learned_speedup_vs_heuristic_total": tot["heuristic_bursts"] / tot["learned_bursts"],
                               "learned_speedup_vs_plain_geomean": geo("plain_bursts", "learned_bursts"),
                               "learned_speedup_vs_heuristic_geomean": geo("heuristic_bursts", "learned_bursts")}}
    log(json.dumps(summary, indent=1))
    meta.update({"held_out": summary, "held_out_runs": rows})
    with open(os.path.join(a.out, "bfs_policy.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
