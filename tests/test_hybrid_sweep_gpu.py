"""The alpha / beta sweep of the hybrid BFS (scripts/hybrid-advance.sh:38-88)
on a small graph: every (alpha, beta) gives plain push's depths (the sweep
compares device digests and raises otherwise) and the grid has one row per
pair."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hybrid_sweep_small_grid(sg, device):
    from sygraph_amd import hybrid_sweep
    G = sg.Graph.rmat(device, 12, 16, 0.57, 0.19, 0.19, seed=12)
    G.prepare()
    off, _, _ = G.download()
    srcs = [int(x) for x in np.nonzero(np.diff(off) > 0)[0][:3]]
    alphas, betas = [1, 15, 40], [1, 24, 100000000]
    rows = hybrid_sweep.sweep(device, G, srcs, alphas, betas, reps=1)
    assert [(r["alpha"], r["beta"]) for r in rows] == [(a, b) for a in alphas for b in betas]
    assert all(r["ms"] > 0 and r["gteps"] > 0 for r in rows)
    assert any(r["n_pull"] > 0 for r in rows)  # some pairs pull
    assert len(hybrid_sweep.table(rows, alphas, betas).splitlines()) == 1 + len(alphas)
    G.destroy()
