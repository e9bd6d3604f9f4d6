"""The learned BFS policy on the MI355X: HIP forward, gradient and Adam
kernels against the FP64 oracle (oracle/mlp.py, itself pinned by finite
differences), a training run that must fit a learnable labelling, and the
learned / forced per-level actions inside the BFS, which must leave every
depth bit-exact.  FP32 tolerances are stated per check."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _problem(seed=0, B=300):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, mlp.N_PARAMS).astype(np.float32)
    X = rng.uniform(0, 1, (B, mlp.N_IN)).astype(np.float32)
    y = rng.integers(0, mlp.N_OUT, B).astype(np.int32)
    return p, X, y


def test_forward_matches_oracle(sg, device):
    p, X, _ = _problem()
    pol = sg.Policy(device, p)
    got = pol.forward(X)
    ref, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)  # FP32 vs FP64
    assert np.array_equal(pol.get_params(), p)


def test_gradient_matches_oracle(sg, device):
    p, X, y = _problem(1)
    pol = sg.Policy(device, p)
    loss, g = pol.gradient(X, y)
    rl, rg = mlp.loss_and_grad(p.astype(np.float64), X.astype(np.float64), y)
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    scale = np.abs(rg).max()
    assert np.all(np.abs(g - rg) <= 1e-5 * scale + 1e-4 * np.abs(rg)), np.abs(g - rg).max()
    # deterministic: the gradient kernel sums each parameter in sample order
    assert np.array_equal(pol.gradient(X, y)[1], g)


def test_adam_steps_match_oracle(sg, device):
    p, X, y = _problem(2)
    pol = sg.Policy(device, p)
    xb, yb = device.to_device(X), device.to_device(y)
    q, m, v = p.astype(np.float64), np.zeros(mlp.N_PARAMS), np.zeros(mlp.N_PARAMS)
    for t in range(1, 4):
        _, rg = mlp.loss_and_grad(q, X.astype(np.float64), y)
        q, m, v = mlp.adam_step(q, rg, m, v, t, lr=1e-3)
        pol.train_step(xb, yb, len(X), 1e-3)
    assert np.allclose(pol.get_params(), q, rtol=1e-4, atol=2e-6)


def test_training_fits_a_learnable_labelling(sg, device):
    rng = np.random.default_rng(3)
    X = rng.uniform(0, 1, (4096, mlp.N_IN)).astype(np.float32)
    A = rng.normal(0, 1, (mlp.N_OUT, mlp.N_IN))
    y = np.argmax(X @ A.T, axis=1).astype(np.int32)
    pol = sg.Policy(device, rng.normal(0, 0.3, mlp.N_PARAMS).astype(np.float32))
    xb, yb = device.to_device(X), device.to_device(y)
    first = pol.train_step(xb, yb, len(X), 1e-2)
    for _ in range(400):
        last = pol.train_step(xb, yb, len(X), 1e-2)
    acc = float((np.argmax(pol.forward(X), 1) == y).mean())
    assert last < 0.5 * first and acc > 0.85, (first, last, acc)


def _bias_only(action):
    p = np.zeros(mlp.N_PARAMS, dtype=np.float32)
    p[mlp.N_PARAMS - mlp.N_OUT + action] = 1.0  # b3: every level scores `action` highest
    return p


@pytest.mark.parametrize("action", range(6))
def test_learned_bfs_follows_the_policy_and_stays_exact(sg, device, oracle_mod, action):
    g = oracle_mod.rmat(12, 16, seed=33)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    src = int(np.argmax(g.degrees()))
    ref, levels = oracle_mod.bfs(g, src)
    pol = sg.Policy(device, _bias_only(action))
    b = sg.BFS(G)
    b.init(src)
    st = b.run(sg.POLICY_LEARNED, learned=pol, tuning={"learned_gate": -1})  # the network decides every level
    assert np.array_equal(b.get_distances(), ref) and st.levels == levels
    acts = [st.level[i].action for i in range(st.levels)]
    assert acts[0] == 0 and all(a == action for a in acts[1:]), acts
    # with the default gate, the levels whose frontier has fewer arcs than 1/1024 of the
    # graph's arcs push with the bucketing mapping and the others follow the network
    b.init(src)
    st = b.run(sg.POLICY_LEARNED, learned=pol)
    assert np.array_equal(b.get_distances(), ref)
    lv = [st.level[i] for i in range(st.levels)]
    for L in range(1, st.levels):
        gated = lv[L].frontier_edges * 1024 < g.m
        assert lv[L].action == (0 if gated else action), (L, lv[L].action, gated)


@pytest.mark.parametrize("lb", ["LB_WORKGROUP", "LB_SUBGROUP", "LB_WORKITEM"])
@pytest.mark.parametrize("policy", ["POLICY_PUSH", "POLICY_PULL", "POLICY_HYBRID"])
def test_work_mapping_variants_exact(sg, device, oracle_mod, lb, policy):
    g = oracle_mod.rmat(12, 16, seed=34)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    src = int(np.argmax(g.degrees()))
    b = sg.BFS(G)
    b.init(src)
    st = b.run(getattr(sg, policy), 15.0, 24.0, load_balancer=getattr(sg, lb))
    assert np.array_equal(b.get_distances(), oracle_mod.bfs(g, src)[0])
    variant = {"LB_WORKGROUP": 1, "LB_SUBGROUP": 2, "LB_WORKITEM": 0}[lb]
    assert all(st.level[i].action % 3 == variant for i in range(st.levels))


def test_learned_without_policy_is_rejected(sg, device, oracle_mod):
    g = oracle_mod.rmat(8, 4, seed=1)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    b = sg.BFS(G)
    b.init(0)
    with pytest.raises(sg.SygraphError):
        b.run(sg.POLICY_LEARNED)


@pytest.mark.parametrize("kind", ["rmat", "grid"])
def test_in_situ_decisions_match_the_oracle_forward(sg, device, oracle_mod, kind):
    """Every level's action under a random policy equals the FP64 oracle's
    argmax on the features of the logged level statistics (decisions whose
    two best scores lie within FP32 rounding of each other are skipped)."""
    g = oracle_mod.rmat(14, 16, seed=35) if kind == "rmat" else oracle_mod.grid(96, 96, 0.1, seed=36)
    G = sg.Graph.from_csr(device, g.offsets, g.cols)
    rng = np.random.default_rng(37)
    checked = set()
    for trial in range(4):
        p = rng.normal(0, 1.0, mlp.N_PARAMS).astype(np.float32)
        pol = sg.Policy(device, p)
        src = int(rng.choice(np.nonzero(g.degrees() > 0)[0]))
        b = sg.BFS(G)
        b.init(src)
        st = b.run(sg.POLICY_LEARNED, learned=pol, tuning={"learned_gate": -1})
        assert np.array_equal(b.get_distances(), oracle_mod.bfs(g, src)[0])
        lv = [st.level[i] for i in range(st.levels)]
        for L in range(1, st.levels):
            f = mlp.features(lv[L].frontier_vertices, lv[L].frontier_edges, lv[L - 1].unexplored_edges, L,
                             lv[L - 1].max_degree, g.n, g.m)
            s, _ = mlp.forward(p.astype(np.float64), f[None, :].astype(np.float64))
            top = np.sort(s[0])
            if top[-1] - top[-2] <= 1e-4 * (1 + abs(top[-1])):
                continue
            assert lv[L].action == int(np.argmax(s[0])), (trial, L, s[0], lv[L].action)
            checked.add(lv[L].action)
    assert len(checked) >= 2, checked  # the random policies exercised several actions
