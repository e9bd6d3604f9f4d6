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
