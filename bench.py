"""Benchmark: BFS GTEPS on one MI355X, RMAT scale 24 (edge factor 16,
a/b/c/d = .57/.19/.19/.05, seed 2, int64 offsets / int32 ids), 16 sources.

One step = one BFS from each of the 16 sources (the committed source list of
tests/golden/rmat24_seed2.json, drawn from the largest component).  The graph
is generated on the device before the timed region; depth arrays are device
resident.  value = sum over sources of the TEPS numerator (sum of degrees of
reached vertices, algorithms/bfs.hpp:88-99) / wall time of the steps.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS on 1 MI355X (BFS/SSSP RMAT s24), % HBM peak; learned+batched vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def level_bytes(ls, w0_bytes: int, off_bytes: int = 8) -> float:
    """Algorithmic bytes of one BFS level's traversal kernel.  Push: 2*w_o per
    frontier vertex, 8 B per arc (4 B id + 4 B status probe), three bitmap
    passes (the discoveries' depths are written by the count pass).  Pull: per
    scanned vertex 2*w_o (its offsets) + 8 B (hint id and its probe); per probe
    of the adjacency list after a hint miss 8 B (edges_inspected counts the
    hint probe too); 4 B per discovered vertex; three bitmap passes."""
    if ls.direction == 0:
        # the depths of a push level's discoveries are written by the count pass, in vertex order
        return 2 * off_bytes * ls.frontier_vertices + 8 * ls.frontier_edges + 3 * w0_bytes
    return (2 * off_bytes * ls.vertices_scanned + 8 * ls.edges_inspected + 4 * ls.discovered + 3 * w0_bytes)


def run_policy(sg, G, sources, policy, alpha, beta, steps, warmup, collect, learned=None):
    """The timed steps (one traversal per source each), then, with collect, one
    more untimed pass over the sources with each step's kernels timed on their
    own (an extra event per step, a stream barrier, kept out of the timed
    steps) for the per-kernel figures."""
    b = sg.BFS(G)
    edges_total = 0
    # -> [bytes, ms, launches]; "step rest": the count / scan / write launches
    # of each step and the bursts that run inside them (timed together)
    per_kernel = {0: [0.0, 0.0, 0], 1: [0.0, 0.0, 0], "burst": [0.0, 0.0, 0], "step rest": [0.0, 0.0, 0]}
    w0_bytes = ((G.n + 63) // 64) * 8
    kernel_ms_total = 0.0
    per_trav = []  # device-time GTEPS of each timed traversal
    # a step: one traversal per source, back to back on the device (sg_bfs_many);
    # the rows stay on the device
    rows = G.dev.alloc(4 * len(sources) * G.n)
    for _ in range(warmup):
        sg.bfs_many(G, sources, policy, alpha, beta, out=rows, learned=learned)
    G.dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = sg.bfs_many(G, sources, policy, alpha, beta, out=rows, learned=learned)
        edges_total += st.reached_edges
        kernel_ms_total += st.kernel_ms
    G.dev.synchronize()
    wall = time.perf_counter() - t0
    rows.free()
    if collect:
        for s in sources:
            b.init(s)
            st = b.run(policy, alpha, beta, learned=learned, record_levels=True)
            if st.kernel_ms > 0:
                per_trav.append(st.reached_edges / (st.kernel_ms / 1e3) / 1e9)
            for L in range(min(st.levels, sg.MAX_LEVELS_RECORDED)):
                ls = st.level[L]
                # levels run inside a burst are not level_kernel launches
                k = per_kernel["burst" if ls.burst else ls.direction]
                k[0] += level_bytes(ls, w0_bytes)
                k[1] += ls.kernel_ms
                k[2] += 1
                if not ls.burst:
                    per_kernel["step rest"][1] += ls.convert_ms
                    per_kernel["step rest"][2] += 1
    return edges_total, wall, kernel_ms_total, per_kernel, per_trav


def run_policy_lb(sg, G, sources, policy, lb):
    """One BFS per source under a fixed direction and work mapping (push:
    bucketing bins / edge-balanced / wave-split large lists; pull: lane-serial
    head of 16 / 4 / 64 in-arcs), after one untimed pass."""
    b = sg.BFS(G)
    for s in sources[:2]:
        b.init(s)
        b.run(policy, 15.0, 24.0, load_balancer=lb)
    G.dev.synchronize()
    t0 = time.perf_counter()
    e = 0
    for s in sources:
        b.init(s)
        e += b.run(policy, 15.0, 24.0, load_balancer=lb).reached_edges
    G.dev.synchronize()
    return e, time.perf_counter() - t0


def copy_ceiling(sg, dev, mib: int = 1024, reps: int = 5):
    """Read+write bandwidth of a plain 16-byte streaming copy (sg_copy_device)
    of `mib` MiB, best of `reps`, timed with hipEvents on the library stream."""
    import ctypes as C

    nbytes = mib << 20
    src, dst = dev.alloc(nbytes), dev.alloc(nbytes)
    src.fill(0x5A, 1, nbytes)
    lib = sg.lib()
    e0, e1 = dev.event(), dev.event()
    best = None
    for _ in range(reps):
        e0.record()
        sg.check(lib.sg_copy_device(dev.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes))
        e1.record()
        dev.synchronize()
        ms = e0.elapsed_ms(e1)
        best = ms if best is None else min(best, ms)
    src.free()
    dst.free()
    return 2.0 * nbytes / (best / 1e3) / 1e9


def gather_calibration(sg, dev, src_gib: int = 8, count: int = 1 << 24):
    """One random 4-byte gather (sg_gather_device) of `count` words from a
    `src_gib` GiB source, far larger than the caches, so nearly every word
    brings in a line of its own: the reference point of the HBM counters for
    random probes (the streaming copy is the other).  Returns the gather's
    rate in G words/s (hipEvents on the library stream)."""
    import ctypes as C

    words = (src_gib << 30) // 4
    src, dst = dev.alloc(4 * words), dev.alloc(4 * count)
    lib = sg.lib()
    e0, e1 = dev.event(), dev.event()
    e0.record()
    sg.check(lib.sg_gather_device(dev.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), words, count, 7))
    e1.record()
    dev.synchronize()
    ms = e0.elapsed_ms(e1)
    src.free()
    dst.free()
    return count / (ms / 1e3) / 1e9


def measured_traffic(kernel: str, summary: str = "bfs_hybrid_rmat24_traffic.json"):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter read on
    the calibration copy kernel), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def batched_vs_plain(sg, G, sources16):
    """64 sources (the 16 committed ones + 48 seeded non-isolated vertices):
    batched multi-source BFS (64 bit-lanes) against one plain push BFS per
    source and one hybrid BFS per source.  Depths must agree exactly."""
    off, _, _ = G.download()
    rng = np.random.default_rng(64)
    cand = np.nonzero(np.diff(off) > 0)[0]
    srcs = list(sources16) + [int(x) for x in rng.choice(cand, 48, replace=False)]
    b = sg.BFS(G)
    out = {}
    edges = 0
    rows = {}
    for name, pol in [("plain_push", sg.POLICY_PUSH), ("hybrid", sg.POLICY_HYBRID)]:
        G.dev.synchronize()
        t0 = time.perf_counter()
        e = 0
        for s in srcs:
            b.init(s)
            st = b.run(pol, 15.0, 24.0, record_levels=False)
            e += st.reached_edges
            if name == "plain_push" and s in (srcs[0], srcs[-1]):
                rows[s] = b.get_distances()
        G.dev.synchronize()
        out[name] = e / (time.perf_counter() - t0) / 1e9
        edges = e
    import ctypes as C

    buf = G.dev.alloc(4 * len(srcs) * G.n)  # depth rows stay on the device
    sa = np.ascontiguousarray(np.asarray(srcs, dtype=np.int32))
    G.dev.synchronize()
    t0 = time.perf_counter()
    sg.check(sg.lib().sg_bfs_multi(G.h, len(srcs), sa.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr)))
    G.dev.synchronize()
    t = time.perf_counter() - t0
    out["batched"] = edges / t / 1e9
    d = buf.download(np.int32, len(srcs) * G.n).reshape(len(srcs), G.n)
    assert np.array_equal(d[0], rows[srcs[0]]) and np.array_equal(d[-1], rows[srcs[-1]])
    buf.free()
    return {k: round(v, 3) for k, v in out.items()}


def secondary(sg, dev):
    """CC, TC and BC on the other configurations of the path (best of 3,
    graphs generated on the device): CC on the 4096 x 4096 grid with 10 % of
    the edges deleted and on RMAT scale 20 edge factor 32; BFS on that grid
    with bursts of small levels (single-workgroup and grid-wide), with the
    single-workgroup ones only, and with none; TC on that RMAT; BC from 64
    sources on RMAT scale 20 edge factor 48, a = 0.55, one source at a time
    and in a 64-lane batch, and 256 sources in four batches."""
    out = {}

    def best(fn, reps=3):
        t = []
        for _ in range(reps):
            dev.synchronize()
            t0 = time.perf_counter()
            fn()
            dev.synchronize()
            t.append(time.perf_counter() - t0)
        return min(t)

    G = sg.Graph.grid(dev, 4096, 4096, 0.1, seed=4)
    cc = sg.CC(G)
    cc.init()
    t = best(cc.run)
    out["cc_grid4096_ms"] = round(t * 1e3, 3)
    out["cc_grid4096_gteps"] = round(G.m / t / 1e9, 3)
    # BFS on the grid (diameter ~8,000 levels) from the first non-isolated
    # vertex (a corner), with small levels run in single-workgroup bursts
    # (default) and without
    off, _, _ = G.download()
    src = int(np.nonzero(np.diff(off) > 0)[0][0])
    b = sg.BFS(G)
    for key, burst, grid in (("bfs_grid4096_ms", 0, 0), ("bfs_grid4096_no_grid_burst_ms", 0, -1),
                             ("bfs_grid4096_no_burst_ms", -1, -1)):
        def one():
            b.init(src)
            b.run(sg.POLICY_HYBRID, 15.0, 24.0, tuning={"burst_arcs": burst, "grid_burst_arcs": grid})
        out[key] = round(best(one, reps=2) * 1e3, 3)
    out["bfs_grid4096_levels"] = int(b.stats.levels)
    G.destroy()
    G = sg.Graph.rmat(dev, 20, 32, seed=5)
    cc = sg.CC(G)
    cc.init()
    t = best(cc.run)
    out["cc_rmat20_ms"] = round(t * 1e3, 3)
    off, cols, _ = G.download()
    b_cc = cc_bytes(off, cc.get_labels())
    out["cc_rmat20_roofline"] = {"alg_bytes": round(b_cc), "achieved_gbs": round(b_cc / t / 1e9, 1),
                                 "frac": round(b_cc / t / 1e9 / HBM_PEAK_GBS, 4), "time": "wall, whole call"}
    tc = sg.TC(G)
    t = best(tc.run)
    out["tc_rmat20_ms"] = round(t * 1e3, 3)
    out["tc_rmat20_triangles"] = tc.count
    b_tc = tc_bytes(off, cols)
    out["tc_rmat20_roofline"] = {"alg_bytes": round(b_tc), "achieved_gbs": round(b_tc / t / 1e9, 1),
                                 "frac": round(b_tc / t / 1e9 / HBM_PEAK_GBS, 4), "time": "wall, whole call"}
    del off, cols
    G.destroy()
    G = sg.Graph.rmat(dev, 20, 48, 0.55, 0.17, 0.17, seed=6)
    off, _, _ = G.download()
    rng = np.random.default_rng(6)
    srcs = [int(x) for x in rng.choice(np.nonzero(np.diff(off) > 0)[0], 64, replace=False)]
    bc = sg.BC(G)
    bc.init(srcs)
    t = best(bc.run, reps=1)
    out["bc_rmat20_64src_ms"] = round(t * 1e3, 3)
    bc.init(srcs)
    t = best(lambda: bc.run(batched=True), reps=2)
    out["bc_rmat20_64src_batched_ms"] = round(t * 1e3, 3)
    srcs256 = [int(x) for x in rng.choice(np.nonzero(np.diff(off) > 0)[0], 256, replace=False)]
    bc.init(srcs256)
    t = best(lambda: bc.run(batched=True), reps=1)
    out["bc_rmat20_256src_batched_ms"] = round(t * 1e3, 3)
    G.destroy()
    out.update(c4(sg, dev, best))
    return out


def c4(sg, dev, best):
    """The two batched-BC graphs with the committed fixture's sources
    (tests/golden/c4_bc.json): RMAT scale 20 edge factor 56, a/b/c/d =
    .55/.20/.20/.05, seed 6 (1,048,576 vertices, average degree 101.5), all
    256 sources in four 64-lane batches; the 4096 x 4096 grid, the first
    64-lane batch (one batch takes seconds there).  Batched BC and
    multi-source BFS, best of the timed calls after a first call that sizes
    the buffers."""
    import ctypes as C
    path = os.path.join(ROOT, "tests", "golden", "c4_bc.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        fx = json.load(f)
    out = {}
    for key, nsrc in (("rmat", 256), ("grid", 64)):
        rec = fx[key]
        p = rec["params"]
        if p["kind"] == "rmat":
            G = sg.Graph.rmat(dev, p["scale"], p["edge_factor"], p["a"], p["b"], p["c"], seed=p["seed"])
        else:
            G = sg.Graph.grid(dev, p["rows"], p["cols"], p["p_delete"], seed=p["seed"])
        srcs = rec["sources"][:nsrc]
        bc = sg.BC(G)
        bc.init(srcs)
        bc.run(batched=True)
        t = best(lambda: bc.run(batched=True), reps=1 if key == "grid" else 2)
        tag = "c4a_rmat20_ef56" if key == "rmat" else "c4b_grid4096"
        out[f"{tag}_bc_{nsrc}src_batched_ms"] = round(t * 1e3, 3)
        buf = dev.alloc(4 * nsrc * G.n)
        sa = np.ascontiguousarray(np.asarray(srcs, dtype=np.int32))
        run = lambda: sg.check(sg.lib().sg_bfs_multi(G.h, nsrc, sa.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr)))
        run()
        t = best(run, reps=1 if key == "grid" else 2)
        out[f"{tag}_msbfs_{nsrc}src_ms"] = round(t * 1e3, 3)
        buf.free()
        G.destroy()
    return out


def sssp_bytes(sg, st, n: int, off_bytes: int = 8) -> tuple[float, float]:
    """Algorithmic bytes of one SSSP run: 4 B per vertex (distance init) and
    per relaxing round three bitmap passes, 2*w_o per frontier vertex and 12 B
    per relaxed arc (4 B id + 4 B weight + 4 B distance atomic-min), from the
    run's per-round counters; and of its relax launches alone (2*w_o per
    frontier vertex, 12 B per relaxed arc)."""
    w0 = ((n + 63) // 64) * 8
    b, br = 4.0 * n, 0.0
    for r in range(min(st.rounds, sg.MAX_LEVELS_RECORDED)):
        relax = 2 * off_bytes * st.round_vertices[r] + 12 * st.round_edges[r]
        b += 3 * w0 + relax
        br += relax
    return b, br


def run_sssp(sg, G, sources, delta=0.0, model=None):
    """Wall time of one SSSP per source; `model` (a dict) receives the summed
    algorithmic bytes and device time of the runs."""
    s = sg.SSSP(G)
    s.init(sources[0])
    s.run(delta=delta)  # workspace and mean weight outside the timed loop
    edges = 0
    byts, dev_ms = 0.0, 0.0
    G.dev.synchronize()
    t0 = time.perf_counter()
    for src in sources:
        s.init(src)
        st = s.run(delta=delta)
        edges += st.reached_edges
        if model is not None:
            byts += sssp_bytes(sg, st, G.n)[0]
            dev_ms += st.kernel_ms
    G.dev.synchronize()
    wall = time.perf_counter() - t0
    if model is not None:
        model["bytes"], model["ms"] = byts, dev_ms
        # the dominant kernel on its own: every relax launch timed by hipEvents on the library
        # stream (a separate pass, as the events are stream barriers)
        rb, rms, rl = 0.0, 0.0, 0
        for src in sources:
            s.init(src)
            st = s.run(delta=delta, record_relax=True)
            rb += sssp_bytes(sg, st, G.n)[1]
            rms += st.relax_ms
            rl += st.relax_launches
        model["relax_bytes"], model["relax_ms"], model["relax_launches"] = rb, rms, rl
    return edges, wall


def tc_bytes(off: np.ndarray, cols: np.ndarray, off_bytes: int = 8) -> float:
    """Algorithmic bytes of TC: orientation (offsets, ids and the heads'
    offsets read, DAG ids written) plus the intersections, sum over u of
    (2*w_o + 4*d+(u)) and over DAG arcs (u,v) of (2*w_o + 4*d+(v))."""
    deg = np.diff(off).astype(np.int64)
    n, m = len(deg), len(cols)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    dst = cols.astype(np.int64)
    up = (deg[src] < deg[dst]) | ((deg[src] == deg[dst]) & (src < dst))
    dplus = np.bincount(src[up], minlength=n).astype(np.int64)
    m_dag = int(up.sum())
    orient = 2 * off_bytes * n + 4 * m + 2 * off_bytes * m + 4 * m_dag
    inter = 2 * off_bytes * n + 4 * m_dag + 2 * off_bytes * m_dag + 4 * int(dplus[dst[up]].sum())
    return float(orient + inter)


def cc_bytes(off: np.ndarray, labels: np.ndarray, off_bytes: int = 8) -> float:
    """Algorithmic bytes of undirected CC (sampling union-find): every vertex
    links its first two arcs (offsets, ids, two parent reads each), two
    flattens (parent read and written), the outside pass (offsets, parent),
    and the vertices outside the largest component link their other arcs."""
    deg = np.diff(off).astype(np.int64)
    n = len(deg)
    vals, counts = np.unique(labels, return_counts=True)
    outside = labels != vals[np.argmax(counts)]
    rest = int(np.maximum(deg[outside] - 2, 0).sum())
    return float(2 * off_bytes * n + 12 * int(np.minimum(deg, 2).sum()) + 16 * n + (2 * off_bytes + 4) * n
                 + 12 * rest)


def cpu_baseline(oracle, G, sources, budget_s: float):
    off, cols, _ = G.download()
    g = oracle.CSRGraph(G.n, off, cols)
    threads = oracle.omp_max_threads()
    deg = g.degrees()
    edges = 0
    t_used = 0.0
    done = 0
    for s in sources:
        t0 = time.perf_counter()
        d, _ = oracle.bfs_plain_omp(g, s, oracle.BFS_HYBRID, 15.0, 24.0)
        t_used += time.perf_counter() - t0
        edges += int(deg[d != g.n + 1].sum())
        done += 1
        if t_used > budget_s:
            break
    return {"value": edges / t_used / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"{done} of {len(sources)} sources, OpenMP restatement of the reference plain method "
                      f"(2-level bitmap frontier, hybrid alpha=15 beta=24), {t_used:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--policy", default="hybrid", choices=["push", "pull", "hybrid"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the push-only / pull-only side measurements")
    ap.add_argument("--algo", default="bfs", choices=["bfs", "sssp", "secondary"],
                    help="sssp: time only SSSP from the 16 sources; secondary: only the CC / TC / BC "
                         "configurations (for profiling those paths)")
    ap.add_argument("--delta", type=float, default=0.0, help="SSSP bucket width with --algo sssp (0 automatic)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported (the traversal core runs on a "
              f"single device; sharding over several GPUs is a design, not built).", file=sys.stderr)
        sys.exit(2)

    import sygraph_amd as sg

    dev = sg.Device(0)
    info = dev.info()
    log(f"device: {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")
    if a.algo == "secondary":
        print(json.dumps({"secondary": secondary(sg, dev)}), flush=True)
        return
    t0 = time.perf_counter()
    G = sg.Graph.rmat(dev, a.scale, 16, 0.57, 0.19, 0.19, seed=2, weight_seed=3)
    log(f"graph: RMAT s{a.scale} n={G.n} m={G.m} generated on device in {time.perf_counter() - t0:.2f}s")
    prep_ms = G.prepare()  # in-degree-0 bitmap and pull hint: per-graph preprocessing, outside the timed steps
    log(f"graph preprocessing (in-degree-0 bitmap, pull hint): {prep_ms:.2f} ms device time")
    golden = os.path.join(ROOT, "tests", "golden", "rmat24_seed2.json")
    if a.scale == 24 and os.path.exists(golden):
        with open(golden) as f:
            sources = json.load(f)["sources"]
    else:
        rng = np.random.default_rng(2)
        off, _, _ = G.download()
        cand = np.nonzero(np.diff(off) > 0)[0]
        sources = [int(x) for x in rng.choice(cand, 16, replace=False)]
    pol = {"push": sg.POLICY_PUSH, "pull": sg.POLICY_PULL, "hybrid": sg.POLICY_HYBRID}[a.policy]
    if a.algo == "sssp":
        copy_ceiling(sg, dev, mib=256, reps=2)  # the counters' calibration kernels, in the same process
        gather_calibration(sg, dev)
        for _ in range(a.warmup):
            run_sssp(sg, G, sources, a.delta)
        e, w = 0, 0.0
        for _ in range(a.steps):
            e1, w1 = run_sssp(sg, G, sources, a.delta)
            e, w = e + e1, w + w1
        print(json.dumps({"metric": "SSSP GTEPS on 1 MI355X (RMAT s24, 16 sources)", "value": round(e / w / 1e9, 3),
                          "unit": "GTEPS", "steps": a.steps, "delta": a.delta}), flush=True)
        return

    edges, wall, kms, per_kernel, per_trav = run_policy(sg, G, sources, pol, 15.0, 24.0, a.steps, a.warmup, True)
    gteps = edges / wall / 1e9
    gteps_kernel = edges / (kms / 1e3) / 1e9 if kms > 0 else None
    log(f"{a.policy}: {gteps:.2f} GTEPS wall, {gteps_kernel:.2f} GTEPS device, {wall / a.steps * 1e3:.2f} ms/step")

    # the traversal of every level outside a burst is one kernel
    # (level_kernel, push or pull body chosen on the device): the roofline
    # covers all its launches
    byts = per_kernel[0][0] + per_kernel[1][0]
    ms = per_kernel[0][1] + per_kernel[1][1]
    launches = per_kernel[0][2] + per_kernel[1][2]
    achieved = byts / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    traffic = measured_traffic("level_kernel")
    avg_s = ms / max(launches, 1) / 1e3
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "basis": "achieved and frac: algorithmic bytes per launch / average launch time; measured_*: the "
                     "committed PMC traffic per launch / the same time",
            "kernel": "level_kernel",
            "launches": launches, "avg_launch_us": round(avg_s * 1e6, 2),
            "alg_bytes_per_launch": round(byts / max(launches, 1)),
            "measured_achieved": round(traffic / avg_s / 1e9, 1) if traffic and avg_s > 0 else None,
            "measured_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic and avg_s > 0 else None}
    for k, name in {0: "level_kernel, push levels", 1: "level_kernel, pull levels"}.items():
        if per_kernel[k][2]:
            b_, m_, n_ = per_kernel[k]
            log(f"  {name}: {n_} levels, avg {m_ / n_ * 1e3:.1f} us, "
                f"{b_ / (m_ / 1e3) / 1e9:.0f} GB/s algorithmic")
    rest, burst = per_kernel["step rest"], per_kernel["burst"]
    if rest[2]:
        log(f"  count / scan / write passes and bursts: {rest[2]} steps, avg {rest[1] / rest[2] * 1e3:.1f} us; "
            f"{burst[2]} levels ran inside bursts")

    extra = {}
    cc = copy_ceiling(sg, dev)
    roof["copy_ceiling_gbs"] = round(cc, 1)
    log(f"copy kernel ceiling: {cc:.0f} GB/s (read + write)")
    gw = gather_calibration(sg, dev)
    roof["random_gather_gwords_s"] = round(gw, 2)
    log(f"random 4-byte gather: {gw:.2f} G words/s")
    if not a.no_extra:
        bp = batched_vs_plain(sg, G, sources)
        extra["bfs_64_sources_gteps"] = bp
        log(f"64 sources: batched {bp['batched']:.1f}, plain push {bp['plain_push']:.1f}, "
            f"hybrid {bp['hybrid']:.1f} GTEPS")
        model = {}
        e2, w2 = run_sssp(sg, G, sources, model=model)
        extra["gteps_sssp"] = round(e2 / w2 / 1e9, 3)
        log(f"sssp: {e2 / w2 / 1e9:.2f} GTEPS wall (16 sources, u32 weights U[1,255] seed 3, automatic delta)")
        s_run = model["bytes"] / (model["ms"] / 1e3) / 1e9 if model["ms"] > 0 else 0.0
        s_ach = model["relax_bytes"] / (model["relax_ms"] / 1e3) / 1e9 if model["relax_ms"] > 0 else 0.0
        s_traffic = measured_traffic("relax_kernel", "sssp_rmat24_traffic.json")
        nl = max(model["relax_launches"], 1)
        avg_s = model["relax_ms"] / nl / 1e3
        extra["roofline_sssp"] = {
            "bound": "hbm", "achieved": round(s_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(s_ach / HBM_PEAK_GBS, 4), "traffic": s_traffic, "kernel": "relax_kernel",
            "launches": model["relax_launches"], "avg_launch_us": round(avg_s * 1e6, 2),
            "alg_bytes_per_launch": round(model["relax_bytes"] / nl),
            "basis": "achieved and frac: algorithmic bytes of the relax launches (2*w_o per frontier vertex, 12 B "
                     "per relaxed arc, from the per-round counters) / their hipEvent time; whole_run_*: every "
                     "launch of the runs (selection and queue writes included)",
            "whole_run_achieved": round(s_run, 1), "whole_run_frac": round(s_run / HBM_PEAK_GBS, 4),
            "measured_achieved": round(s_traffic / avg_s / 1e9, 1) if s_traffic and avg_s > 0 else None,
            "measured_frac": round(s_traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if s_traffic and avg_s > 0 else None}
        log(f"sssp roofline: relax_kernel {s_ach:.0f} GB/s algorithmic ({s_ach / HBM_PEAK_GBS:.3f} of peak), "
            f"{s_run:.0f} GB/s over whole runs")
        by_lb = {}
        for pname, p in (("push", sg.POLICY_PUSH), ("pull", sg.POLICY_PULL)):
            for lname, lbv in (("bucketing", sg.LB_BUCKETING), ("workgroup_mapped", sg.LB_WORKGROUP),
                               ("subgroup_mapped", sg.LB_SUBGROUP)):
                e3, w3 = run_policy_lb(sg, G, sources, p, lbv)
                by_lb[f"{pname}_{lname}"] = round(e3 / w3 / 1e9, 3)
        extra["gteps_by_load_balancer"] = by_lb
        log(f"by load balancer (16 sources, one step each): {by_lb}")
        sweep = {}
        for d in [-1.0, 16.0, 32.0, 64.0, 128.0, 256.0]:
            e3, w3 = run_sssp(sg, G, sources[:4], d)
            sweep["bellman_ford" if d < 0 else f"delta_{int(d)}"] = round(e3 / w3 / 1e9, 3)
        extra["gteps_sssp_by_delta"] = sweep
        log(f"sssp by delta (4 sources): {sweep}")
        for name, p in [("push", sg.POLICY_PUSH), ("pull", sg.POLICY_PULL), ("hybrid", sg.POLICY_HYBRID)]:
            if name == a.policy:
                continue
            e2, w2, k2, _, _ = run_policy(sg, G, sources, p, 15.0, 24.0, 1, 0, False)
            extra[f"gteps_{name}"] = round(e2 / w2 / 1e9, 3)
            log(f"{name}: {e2 / w2 / 1e9:.2f} GTEPS wall")
        weights = os.path.join(ROOT, "sygraph_amd", "data", "bfs_policy.f32")
        if os.path.exists(weights):
            pol = sg.Policy(dev, np.fromfile(weights, dtype=np.float32))
            e2, w2, _, _, _ = run_policy(sg, G, sources, sg.POLICY_LEARNED, 15.0, 24.0, a.steps, 1, False, learned=pol)
            extra["gteps_learned"] = round(e2 / w2 / 1e9, 3)
            log(f"learned: {e2 / w2 / 1e9:.2f} GTEPS wall (sygraph_amd/data/bfs_policy.f32)")

    cpu = None
    if not a.no_cpu_baseline:
        import oracle

        log("cpu baseline ...")
        cpu = cpu_baseline(oracle, G, sources, a.cpu_budget)
        log(f"cpu: {cpu['value']:.3f} GTEPS on {cpu['cores']} threads")

    if not a.no_extra:
        G.destroy()  # frees the scale-24 graph before the secondary graphs
        extra["secondary"] = secondary(sg, dev)
        log(f"secondary: {extra['secondary']}")

    out = {
        "metric": METRIC, "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic (RMAT generated on device)",
        "config": {"workload": f"BFS {a.policy} (alpha=15, beta=24) from 16 sources, RMAT scale {a.scale} edge factor "
                               f"16 a/b/c/d=.57/.19/.19/.05 seed 2, int64 offsets, int32 ids",
                   "vertices": G.n, "arcs": G.m, "sources": len(sources), "policy": a.policy,
                   "graph_preprocessing_ms": round(prep_ms, 3),
                   "gteps_device_time": round(gteps_kernel, 3) if gteps_kernel else None,
                   # per traversal, device time (Graph500 style: harmonic mean over the sources)
                   "gteps_per_source": {
                       "hmean": round(len(per_trav) / sum(1.0 / x for x in per_trav), 3) if per_trav else None,
                       "min": round(min(per_trav), 3) if per_trav else None,
                       "max": round(max(per_trav), 3) if per_trav else None},
                   **extra},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
