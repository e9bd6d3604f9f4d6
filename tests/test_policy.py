"""CPU pins of the learned policy's FP64 oracle (oracle/mlp.py): the analytic
gradient of the mean softmax cross-entropy against central finite
differences, the Adam step against its closed form on a quadratic, and the
feature scaling against hand-computed values."""
import numpy as np

from oracle import mlp


def _problem(seed=0, B=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, mlp.N_PARAMS)
    X = rng.uniform(0, 1, (B, mlp.N_IN))
    y = rng.integers(0, mlp.N_OUT, B)
    return p, X, y


def test_param_count_and_layout():
    assert mlp.N_PARAMS == 1542
    p = np.arange(mlp.N_PARAMS, dtype=np.float64)
    w = mlp.unpack(p)
    assert w["W1"][1, 0] == 8 and w["b1"][0] == 256 and w["W2"][0, 0] == 288 and w["b3"][5] == 1541
    assert np.array_equal(mlp.pack(w), p)


def test_gradient_matches_finite_differences():
    p, X, y = _problem()
    _, g = mlp.loss_and_grad(p, X, y)
    rng = np.random.default_rng(5)
    for idx in list(rng.integers(0, mlp.N_PARAMS, 60)) + [0, 255, 256, 287, 1311, 1343, 1535, 1541]:
        fd = mlp.finite_difference(p, X, y, int(idx))
        assert abs(fd - g[idx]) <= 1e-7 + 1e-5 * abs(fd), (idx, fd, g[idx])


def test_loss_at_zero_params_is_log6():
    _, X, y = _problem(1)
    loss, g = mlp.loss_and_grad(np.zeros(mlp.N_PARAMS), X, y)
    assert abs(loss - np.log(6)) < 1e-12
    # with zero weights only the output biases receive gradient
    w = mlp.unpack(g)
    assert np.all(w["W1"] == 0) and np.all(w["W2"] == 0) and np.all(w["W3"] == 0)
    assert np.allclose(w["b3"], np.bincount(y, minlength=6) / len(y) * -1 + 1 / 6)


def test_adam_first_step_is_signed_lr():
    p = np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    q, m, v = mlp.adam_step(p, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    assert np.allclose(q[:3], -0.01 * np.sign(g[:3]), rtol=1e-4) and q[3] == 0


def test_features_hand_values():
    f = mlp.features(1023, 1023 * 3, 2 ** 20 - 1, 7, 255, 2 ** 16, 2 ** 16 * 7)
    assert np.allclose(f, [10 / 24, np.log2(3070) / 32, 20 / 32, 3 / 12, 2 / 16, 8 / 24, 3 / 8, 16 / 32], rtol=1e-6)


def test_training_features_match_the_oracle():
    """train_policy.level_features (the training set) must compute what the
    device computes at a decision; the oracle restates the device formula."""
    from sygraph_amd.train_policy import level_features
    rng = np.random.default_rng(3)
    stats = [(int(rng.integers(1, 10 ** 6)), int(rng.integers(1, 10 ** 8)), int(rng.integers(0, 10 ** 9)),
              int(rng.integers(0, 10 ** 5))) for _ in range(40)]
    n, m = 1 << 22, 1 << 28
    got = level_features(stats, n, m)
    want = np.stack([mlp.features(stats[L][0], stats[L][1], stats[L - 1][2], L, stats[L - 1][3], n, m)
                     for L in range(1, len(stats))])
    assert np.array_equal(got, want)


def test_shipped_policy_and_its_held_out_record():
    """The shipped parameters (1,542 finite FP32 values), the record of the run
    that trained them (sygraph_amd/data/bfs_policy.json) and the held-out
    comparison run with them (sygraph_amd/data/bfs_heldout.json, written by
    python -m sygraph_amd.heldout_bfs): held-out graphs disjoint from the
    training graphs, four sources on each of the 50, every method with the
    same depths (the script compares device digests and raises otherwise),
    the learned policy at least as fast as the reference's heuristic (its
    judge) in total, and the learned policy on batches of four sources at
    least 1.3x plain push in total."""
    import json
    import os
    d = os.path.join(os.path.dirname(__file__), "..", "sygraph_amd", "data")
    p = np.fromfile(os.path.join(d, "bfs_policy.f32"), dtype=np.float32)
    assert p.shape == (mlp.N_PARAMS,) and np.all(np.isfinite(p))
    with open(os.path.join(d, "bfs_policy.json")) as f:
        meta = json.load(f)
    with open(os.path.join(d, "bfs_heldout.json")) as f:
        ho = json.load(f)
    tr, hs = meta["train_seeds"], ho["seeds"]
    assert (tr[0], tr[1]) == (100, 299) and (hs[0], hs[1]) == (300, 349)
    assert not meta["quick"] and not ho["quick"]
    assert len(ho["runs"]) == 50 and all(len(r["sources"]) == 4 for r in ho["runs"])
    tot = ho["summary"]["total_ms"]
    assert tot["learned"] <= tot["heuristic"]
    assert tot["plain"] >= 1.3 * tot["learned_batched"]
    assert tot["plain"] > tot["learned"]
