"""The held-out comparison scripts on one small graph each: every method
gives the same depths / distances (the scripts compare device digests and
raise otherwise) and the timings are positive."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_heldout_bfs_one_graph(sg, device):
    from sygraph_amd import heldout_bfs
    pol = sg.Policy(device, np.fromfile(heldout_bfs.os.path.join(heldout_bfs.HERE, "data", "bfs_policy.f32"),
                                        dtype=np.float32))
    rows = heldout_bfs.evaluate(device, pol, [301], quick=True)
    assert len(rows) == 1 and len(rows[0]["sources"]) == 4
    assert all(rows[0][k] > 0 for k in heldout_bfs.KEYS)


def test_heldout_sssp_one_graph(sg, device):
    from sygraph_amd import heldout_sssp
    rows = heldout_sssp.evaluate(device, [302], quick=True, nsrc=2, reps=1)
    assert len(rows) == 2
    assert all(r[k] > 0 for r in rows for k, _ in heldout_sssp.MODES)


@pytest.mark.parametrize("seed", [301, 329])
def test_heldout_bfs_full_size_graph_against_oracle(sg, device, oracle_mod, seed):
    """One full-size held-out graph of each kind in the comparison (seed 301:
    RMAT scale 18 edge factor 8; seed 329: the 512 x 512 grid with 5 % of its
    edges deleted): the four sources' depths under plain push, the heuristic,
    the learned policy, the batch of four with its default direction rule and
    with the learned policy all equal the oracle's."""
    import os
    from sygraph_amd.train_policy import make_graph, sources_of
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pol = sg.Policy(device, np.fromfile(os.path.join(root, "sygraph_amd", "data", "bfs_policy.f32"),
                                        dtype=np.float32))
    name, G = make_graph(device, seed, quick=False)
    srcs = sources_of(G, seed)
    off, cols, _ = G.download()
    g = oracle_mod.CSRGraph(G.n, off, cols)
    ref = [oracle_mod.bfs(g, s)[0] for s in srcs]
    b = sg.BFS(G)
    for kw in (dict(policy=sg.POLICY_PUSH), dict(policy=sg.POLICY_HYBRID),
               dict(policy=sg.POLICY_LEARNED, learned=pol)):
        for s, r in zip(srcs, ref):
            b.init(s)
            b.run(alpha=15.0, beta=24.0, **kw)
            assert np.array_equal(b.get_distances(), r), (name, s, kw["policy"])
    for policy in (sg.MSBFS_AUTO, sg.POLICY_LEARNED):
        d = sg.bfs_multi(G, srcs, policy=policy, learned=pol if policy == sg.POLICY_LEARNED else None)
        for i, r in enumerate(ref):
            assert np.array_equal(d[i], r), (name, srcs[i], policy)
    G.destroy()
