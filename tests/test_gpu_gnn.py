"""GPU: BF16-MFMA graph network kernels against the CPU reference (oracle/gnn_ref.py): the fused
inference layers (forward) and the training passes (gradients, which run the training forward).

The reference rounds every GEMM operand to bfloat16 exactly as the kernels do and accumulates in
FP64; the kernels accumulate in FP32, so outputs and gradients agree to ~1e-4 relative (the
tolerance below), far tighter than the ~1e-2 a BF16 rounding mismatch or a transposed operand
would produce."""
import numpy as np
import pytest

from oracle import gnn_ref as R
from staticbtesolver_amd import capi
from staticbtesolver_amd.gnn import Gnn
from tests.cases import case_kwargs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def graph():
    c = capi.Case(**case_kwargs("2DM2DG"))
    return c.geometry()["nbr"]


def _setup(graph, layers, seed):
    rng = np.random.default_rng(seed)
    p = R.init_params(layers, seed)
    p[-65:-1] = rng.normal(0, 0.2, 64).astype(np.float32)
    X = rng.normal(size=(graph.shape[0], R.NIN)).astype(np.float32)
    X[:, 12:] = 0  # padded features
    return p, X, rng


@pytest.mark.parametrize("layers", [1, 4])
def test_forward_matches_reference(graph, layers):
    say("gnn forward", layers)
    p, X, _ = _setup(graph, layers, 1)
    g = Gnn(layers=layers)
    g.set_graph(graph)
    g.set_params(p)
    y = g.forward(X)
    yr, _ = R.forward(p, X, graph, layers, bf=True)
    assert np.max(np.abs(y - yr)) <= 2e-4 * np.max(np.abs(yr)) + 1e-6


def test_forward_matches_reference_on_a_large_mesh(tmp_path):
    """40k nodes (2,514 row tiles over 629 work-groups); the last tile is partial (the node count is
    not a multiple of 16)."""
    from staticbtesolver_amd import meshgen

    say("gnn forward, large mesh")
    p = meshgen.make_square_case(str(tmp_path), 40000, 7, 5, 7, L=1e-6)
    nbr = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-6, Ly=1e-6).geometry()["nbr"]
    assert nbr.shape[0] % 16 != 0
    par, X, _ = _setup(nbr, 4, 5)
    g = Gnn(layers=4)
    g.set_graph(nbr)
    g.set_params(par)
    y = g.forward(X)
    yr, _ = R.forward(par, X, nbr, 4, bf=True)
    assert np.max(np.abs(y - yr)) <= 2e-4 * np.max(np.abs(yr)) + 1e-6


@pytest.mark.parametrize("layers", [1, 4])
def test_gradients_match_reference(graph, layers):
    say("gnn backward", layers)
    p, X, rng = _setup(graph, layers, 2)
    t = rng.normal(size=graph.shape[0]).astype(np.float32)
    w = rng.uniform(0.5, 1.5, graph.shape[0]).astype(np.float32)
    g = Gnn(layers=layers)
    g.set_graph(graph)
    g.set_params(p)
    loss = g.backward(X, t, w)
    gg = g.grads()
    lr, gr, _ = R.loss_and_grad(p, X, graph, t.astype(np.float64), layers, w.astype(np.float64), bf=True)
    assert abs(loss - lr) <= 1e-4 * lr
    assert np.linalg.norm(gg - gr) <= 2e-3 * np.linalg.norm(gr)
    # every parameter block individually (catches a transposed block that the global norm hides)
    off = [0, R.NIN * R.HID, R.NIN * R.HID + R.HID]
    for l in range(layers):
        b = off[-1]
        off += [b + 4096, b + 8192, b + 8192 + 64]
    off += [off[-1] + 64, off[-1] + 65]
    for a, b in zip(off[:-1], off[1:]):
        ref = gr[a:b]
        assert np.linalg.norm(gg[a:b] - ref) <= 5e-3 * np.linalg.norm(ref) + 1e-7, (a, b)


def test_adam_step_matches_reference(graph):
    say("gnn adam")
    p, X, rng = _setup(graph, 2, 3)
    t = rng.normal(size=graph.shape[0]).astype(np.float32)
    g = Gnn(layers=2)
    g.set_graph(graph)
    g.set_params(p)
    g.backward(X, t)
    gg = g.grads()
    g.adam(1e-3)
    p1, _, _ = R.adam(p, gg, np.zeros_like(p), np.zeros_like(p), 1, lr=1e-3)
    np.testing.assert_allclose(g.get_params(), p1, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss_on_a_smooth_target(graph):
    """Target = neighbour average of a feature (a function a 1-layer graph network represents)."""
    say("gnn training")
    rng = np.random.default_rng(4)
    N = graph.shape[0]
    X = np.zeros((N, R.NIN), dtype=np.float32)
    X[:, :4] = rng.normal(size=(N, 4))
    nb, mask, deg = R.neighbours(graph)
    t = (R.aggregate(X[:, :1].astype(np.float64), nb, mask, deg)[:, 0] + 0.5 * X[:, 1]).astype(np.float32)
    g = Gnn(layers=2, seed=5)
    g.set_graph(graph)
    l0 = g.backward(X, t)
    for _ in range(300):
        g.backward(X, t)
        g.adam(3e-3)
    l1 = g.backward(X, t)
    say("loss", l0, "->", l1)
    assert l1 < 0.05 * l0
