"""Pins the CPU reference of the graph network: FP64 gradients against central finite differences,
the bf16 rounding helper against exact representable cases, aggregate_t as the transpose of
aggregate."""
import numpy as np

from oracle import gnn_ref as R


def _graph(n=40, seed=0):
    rng = np.random.default_rng(seed)
    # random symmetric graph with up to 3 neighbours per node (face table with -1 padding)
    nbr = -np.ones((n, 3), dtype=np.int64)
    deg = np.zeros(n, dtype=int)
    for _ in range(2 * n):
        a, b = rng.integers(n, size=2)
        if a != b and deg[a] < 3 and deg[b] < 3 and b not in nbr[a]:
            nbr[a, deg[a]] = b
            nbr[b, deg[b]] = a
            deg[a] += 1
            deg[b] += 1
    return nbr


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0])
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to even (1.0); 1 + 3*2^-8 rounds up
    np.testing.assert_array_equal(R.bf16(x), [1.0, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, -2.5, 0.0])


def test_aggregate_transpose():
    nbr = _graph()
    nb, mask, deg = R.neighbours(nbr)
    rng = np.random.default_rng(1)
    H, G = rng.normal(size=(40, 5)), rng.normal(size=(40, 5))
    lhs = np.sum(R.aggregate(H, nb, mask, deg) * G)
    rhs = np.sum(H * R.aggregate_t(G, nb, mask, deg))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_gradients_match_finite_differences():
    layers = 2
    nbr = _graph()
    rng = np.random.default_rng(2)
    p = R.init_params(layers, 3).astype(np.float64)
    p[-65:-1] = rng.normal(0, 0.3, 64)  # non-trivial output layer
    X = rng.normal(size=(40, R.NIN))
    t = rng.normal(size=40)
    w = rng.uniform(0.5, 1.5, 40)
    loss, g, _ = R.loss_and_grad(p, X, nbr, t, layers, w, bf=False)
    idx = rng.choice(len(p), 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = R.loss_and_grad(pp, X, nbr, t, layers, w, bf=False)[0]
        lm = R.loss_and_grad(pm, X, nbr, t, layers, w, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_matches_closed_form_first_step():
    p = np.array([1.0, -2.0], dtype=np.float32)
    g = np.array([0.5, -0.25], dtype=np.float32)
    p1, m, v = R.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first bias-corrected step moves every parameter by lr * sign(g)
    np.testing.assert_allclose(p1, p - 0.1 * np.sign(g), rtol=1e-6)
