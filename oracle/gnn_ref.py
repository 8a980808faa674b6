"""CPU reference of the mesh graph network used for the learned outer-iteration correction.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network kernels (forward, backward, Adam) in
staticbtesolver_amd/csrc/bte_gnn.hip.  The product never imports it.

Network (no counterpart in the reference): cells are nodes, interior faces are edges.
    H0      = relu(X W_in + b_in)                               X: [N, 16], hidden 64
    M_l     = mean over the interior neighbours of H_l
    H_{l+1} = H_l + relu(H_l W1_l + M_l W2_l + b_l)             l = 0..L-1
    y       = H_L w_out + b_out
Loss: sum_c w_c (y_c - t_c)^2 / sum_c w_c.
With bf16=True every GEMM operand is rounded to bfloat16 (round to nearest even) before an FP64
product, which is what the GPU's BF16 MFMA kernels compute up to FP32 accumulation order.
Parity: the FP64 gradients are pinned by central finite differences (tests/test_gnn_ref.py).
"""
from __future__ import annotations

import numpy as np

NIN = 16
HID = 64


def n_params(layers: int) -> int:
    return NIN * HID + HID + layers * (2 * HID * HID + HID) + HID + 1


def unpack(p, layers):
    p = np.asarray(p, dtype=np.float64)
    o = 0

    def take(shape):
        nonlocal o
        n = int(np.prod(shape))
        a = p[o:o + n].reshape(shape)
        o += n
        return a

    P = {"W_in": take((NIN, HID)), "b_in": take((HID,)), "W1": [], "W2": [], "b": []}
    for _ in range(layers):
        P["W1"].append(take((HID, HID)))
        P["W2"].append(take((HID, HID)))
        P["b"].append(take((HID,)))
    P["w_out"] = take((HID,))
    P["b_out"] = take((1,))
    return P


def pack(P):
    parts = [P["W_in"].ravel(), P["b_in"]]
    for l in range(len(P["W1"])):
        parts += [P["W1"][l].ravel(), P["W2"][l].ravel(), P["b"][l]]
    parts += [P["w_out"], P["b_out"]]
    return np.concatenate(parts)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def neighbours(nbr):
    """Interior neighbour lists and degrees from the face table (negative = boundary)."""
    nbr = np.asarray(nbr)
    mask = nbr >= 0
    deg = mask.sum(axis=1)
    return nbr, mask, deg


def aggregate(H, nbr, mask, deg):
    idx = np.where(mask, nbr, 0)
    S = (H[idx] * mask[..., None]).sum(axis=1)
    return S / np.maximum(deg, 1)[:, None]


def aggregate_t(G, nbr, mask, deg):
    """Transpose of aggregate: out[j] = sum over c with j in nb(c) of G[c] / deg(c)."""
    out = np.zeros_like(G)
    Gd = G / np.maximum(deg, 1)[:, None]
    N, F = nbr.shape
    for f in range(F):
        m = mask[:, f]
        np.add.at(out, nbr[m, f], Gd[m])
    return out


def forward(p, X, nbr, layers, bf=True):
    P = unpack(p, layers)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    nb, mask, deg = neighbours(nbr)
    X = np.asarray(X, dtype=np.float64)
    cache = {"X": X}
    Z0 = r(X) @ r(P["W_in"]) + P["b_in"]
    H = np.maximum(Z0, 0)
    cache["Z0"], cache["H"], cache["M"], cache["Z"] = Z0, [H], [], []
    for l in range(layers):
        M = aggregate(H, nb, mask, deg)
        Z = r(H) @ r(P["W1"][l]) + r(M) @ r(P["W2"][l]) + P["b"][l]
        H = H + np.maximum(Z, 0)
        cache["M"].append(M)
        cache["Z"].append(Z)
        cache["H"].append(H)
    y = H @ P["w_out"] + P["b_out"][0]
    return y, cache


def loss_and_grad(p, X, nbr, target, layers, weight=None, bf=True):
    P = unpack(p, layers)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    nb, mask, deg = neighbours(nbr)
    y, cache = forward(p, X, nbr, layers, bf)
    w = np.ones_like(y) if weight is None else np.asarray(weight, dtype=np.float64)
    sw = w.sum()
    loss = float(np.sum(w * (y - target) ** 2) / sw)
    dy = 2 * w * (y - target) / sw
    G = {"W1": [None] * layers, "W2": [None] * layers, "b": [None] * layers}
    HL = cache["H"][layers]
    G["w_out"] = HL.T @ dy
    G["b_out"] = np.array([dy.sum()])
    dH = np.outer(dy, P["w_out"])
    for l in reversed(range(layers)):
        dZ = dH * (cache["Z"][l] > 0)
        Hl, Ml = cache["H"][l], cache["M"][l]
        G["W1"][l] = r(Hl).T @ r(dZ)
        G["W2"][l] = r(Ml).T @ r(dZ)
        G["b"][l] = dZ.sum(axis=0)
        G1 = r(dZ) @ r(P["W1"][l]).T
        G2 = r(dZ) @ r(P["W2"][l]).T
        dH = dH + G1 + aggregate_t(G2, nb, mask, deg)
    dZ0 = dH * (cache["Z0"] > 0)
    G["W_in"] = r(cache["X"]).T @ r(dZ0)
    G["b_in"] = dZ0.sum(axis=0)
    return loss, pack(G), y


def init_params(layers, seed):
    """He-style initialisation, zero biases, small output layer."""
    rng = np.random.default_rng(seed)
    P = {"W_in": rng.normal(0, np.sqrt(2.0 / NIN), (NIN, HID)), "b_in": np.zeros(HID), "W1": [], "W2": [], "b": []}
    for _ in range(layers):
        P["W1"].append(rng.normal(0, np.sqrt(1.0 / HID) * 0.5, (HID, HID)))
        P["W2"].append(rng.normal(0, np.sqrt(1.0 / HID) * 0.5, (HID, HID)))
        P["b"].append(np.zeros(HID))
    P["w_out"] = rng.normal(0, 0.01, HID)
    P["b_out"] = np.zeros(1)
    return pack(P).astype(np.float32)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (t counts from 1); FP32 state like the GPU kernel."""
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    g = g + np.float32(wd) * p
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    mh = m / np.float32(1 - b1 ** t)
    vh = v / np.float32(1 - b2 ** t)
    p = p - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
    return p, m, v
