"""Autotuner: the choice is a pure function of the timings (CPU), and on the
GPU every tuned point reproduces the default outputs exactly, the choice
comes from the grid and a cached graph is not re-measured."""
import json

import numpy as np
import pytest


def test_select_is_deterministic_argmin():
    from sygraph_amd import autotune

    assert autotune.select([3.0, 1.0, 2.0]) == 1
    assert autotune.select([1.0, 1.0, 0.5, 0.5]) == 2  # ties go to the earliest point
    assert autotune.select([5.0]) == 0


def test_bfs_grid_covers_every_combination():
    from sygraph_amd import autotune

    pts = autotune.bfs_points()
    assert len(pts) == 243 and len({tuple(sorted(p.items())) for p in pts}) == 243
    assert autotune.bfs_points({"chunk": [64, 128]}) == [{"chunk": 64}, {"chunk": 128}]


@pytest.mark.gpu
def test_tune_bfs_exact_and_cached(sg, device, oracle_mod, tmp_path):
    from sygraph_amd import autotune

    g = oracle_mod.rmat(12, 16, seed=51)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    srcs = [int(np.argmax(g.degrees())), 7]
    grid = {"t_medium": [16, 64], "t_large": [256, 1024], "chunk": [64, 1024], "blocks_per_cu": [2, 8]}
    cache = str(tmp_path / "tune.json")
    choice = autotune.tune_bfs(G, srcs, grid=grid, repeats=1, cache=cache)
    assert choice in autotune.bfs_points(grid)
    rec = next(iter(json.load(open(cache)).values()))
    assert len(rec["ms"]) == 16 and rec["choice"] == autotune.bfs_points(grid)[autotune.select(rec["ms"])]
    # the tuned point is exact, and a second call reads the cache
    b = sg.BFS(G)
    b.init(srcs[0])
    b.run(sg.POLICY_HYBRID, 15.0, 24.0, tuning=choice)
    assert np.array_equal(b.get_distances(), oracle_mod.bfs(g, srcs[0])[0])
    assert autotune.tune_bfs(G, srcs, grid={"chunk": [7]}, cache=cache) == choice


@pytest.mark.gpu
def test_tune_sssp_exact(sg, device, oracle_mod, tmp_path):
    from sygraph_amd import autotune

    g = oracle_mod.rmat(11, 16, seed=52)
    oracle_mod.fill_weights(g, 4)
    G = sg.Graph.from_csr(device, g.offsets, g.cols, weights=g.weights)
    d = autotune.tune_sssp(G, [0, 3], deltas=[-1.0, 3.0, 50.0], repeats=1, cache=str(tmp_path / "t.json"))
    assert d in (-1.0, 3.0, 50.0)


@pytest.mark.gpu
def test_tuning_out_of_range_rejected(sg, device, oracle_mod):
    g = oracle_mod.rmat(8, 4, seed=1)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    b = sg.BFS(G)
    b.init(0)
    for bad in [{"chunk": 8}, {"blocks_per_cu": 100}, {"t_medium": 100, "t_large": 50}]:
        with pytest.raises(sg.SygraphError):
            b.run(sg.POLICY_HYBRID, 15.0, 24.0, tuning=bad)
