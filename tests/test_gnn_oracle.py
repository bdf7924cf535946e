"""CPU pins of the surrogate's FP64 oracle: its analytic gradient against
central finite differences, and the flat parameter layout."""
from __future__ import annotations

import numpy as np
import pytest


def _graph(n, rng, radius=0.3):
    pts = rng.random((n, 2))
    d = np.linalg.norm(pts[:, None] - pts[None], axis=-1)
    adj = (d < radius) & ~np.eye(n, dtype=bool)
    off = np.concatenate([[0], np.cumsum(adj.sum(1))]).astype(np.int32)
    recv, send = np.nonzero(adj)
    return off, send.astype(np.int32), recv.astype(np.int32)


def _problem(layers, n, seed=0):
    from tmp_sph_amd.surrogate import gnn_layout
    rng = np.random.default_rng(seed)
    lay = gnn_layout(layers)
    npar = int(lay[-5 + 3] + lay[-5 + 1])
    params = np.zeros(npar, np.float32)
    for i in range(lay.size // 5):
        w, N, K, b = lay[5 * i: 5 * i + 4]
        params[w: w + N * K] = (rng.uniform(-1, 1, N * K) * np.sqrt(6.0 / K)).astype(np.float32)
        params[b: b + N] = rng.uniform(-0.1, 0.1, N).astype(np.float32)
    off, send, recv = _graph(n, rng)
    Xn = np.zeros((n, 128), np.float32)
    Xn[:, :15] = rng.standard_normal((n, 15))
    Xe = np.zeros((send.size, 128), np.float32)
    Xe[:, :3] = rng.standard_normal((send.size, 3))
    target = rng.standard_normal((n, 3)).astype(np.float32)
    typ = (rng.random(n) < 0.8).astype(np.int32)
    return lay, params, off, send, recv, Xn, Xe, target, typ


def test_layout_covers_parameter_vector():
    from tmp_sph_amd.surrogate import gnn_layout
    lay = gnn_layout(10).reshape(-1, 5)
    assert lay.shape[0] == 3 * (2 * 10 + 3)
    ends = lay[:, 3] + lay[:, 1]
    starts = lay[:, 0]
    assert starts[0] == 0 and np.all(starts[1:] == ends[:-1])
    assert np.all(lay[:, 3] == lay[:, 0] + lay[:, 1] * lay[:, 2])
    # 1.58 M parameters at latent 128 and 10 blocks
    assert 1_500_000 < ends[-1] < 1_700_000


@pytest.mark.parametrize("layers", [1, 2])
def test_oracle_gradient_matches_finite_differences(orc, layers):
    lay, params, off, send, recv, Xn, Xe, target, typ = _problem(layers, 40)
    loss, G = orc.gnn_loss(layers, params, lay, Xn, Xe, off, send, recv, target, typ)
    assert loss > 0
    rng = np.random.default_rng(5)
    nlin = lay.size // 5
    checked = 0
    for li in rng.choice(nlin, size=min(nlin, 9), replace=False):
        w, N, K, b = lay[5 * li: 5 * li + 4]
        for idx in (w + rng.integers(0, N * K), b + rng.integers(0, N)):
            if li == nlin - 1 and (idx - (w if idx < b else b)) // (K if idx < b else 1) >= 3:
                continue  # decoder outputs >= 3 do not reach the loss
            eps = 2e-5  # small enough that no ReLU changes side between the two evaluations
            p1 = params.copy()
            p1[idx] += eps
            p2 = params.copy()
            p2[idx] -= eps
            h = float(p1[idx]) - float(p2[idx])
            l1, _ = orc.gnn_loss(layers, p1, lay, Xn, Xe, off, send, recv, target, typ, grad=False)
            l2, _ = orc.gnn_loss(layers, p2, lay, Xn, Xe, off, send, recv, target, typ, grad=False)
            fd = (l1 - l2) / h
            assert fd == pytest.approx(G[idx], rel=1e-5, abs=1e-8 * max(1.0, abs(loss))), (li, idx)
            checked += 1
    assert checked >= 10


def test_bf16_emulation_is_close_to_fp64(orc):
    lay, params, off, send, recv, Xn, Xe, target, typ = _problem(2, 40, seed=3)
    l64, g64 = orc.gnn_loss(2, params, lay, Xn, Xe, off, send, recv, target, typ)
    lb, gb = orc.gnn_loss(2, params, lay, Xn, Xe, off, send, recv, target, typ, bf16=True)
    assert lb == pytest.approx(l64, rel=0.05)
    cos = float(g64 @ gb / np.linalg.norm(g64) / np.linalg.norm(gb))
    assert cos > 0.99


def _bf16_bits(x):
    """Round to nearest even BF16 (through FP32); returns the 16-bit patterns."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _from_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _q(x):
    return _from_bits(_bf16_bits(x))


def _forward_recording(layers, params, lay, Xn, Xe, off, send, recv):
    """An independent NumPy statement of the BF16 forward pass that records
    what the device keeps: h1/h2 of every MLP call, latents and messages."""
    L = lay.reshape(-1, 5)
    H = []

    def lin(li, x):
        w, N, K, b, _ = L[li]
        Wm = params[w: w + N * K].reshape(N, K)
        return _q(x) @ _q(Wm).T + params[b: b + N].astype(np.float64)

    def mlp(first, x):
        b1 = _bf16_bits(np.maximum(lin(first, x), 0.0))
        b2 = _bf16_bits(np.maximum(lin(first + 1, _from_bits(b1)), 0.0))
        H.extend([b1.ravel(), b2.ravel()])
        return lin(first + 2, _from_bits(b2))

    n = Xn.shape[0]
    V = [mlp(0, Xn.astype(np.float64))]
    E = [mlp(3, Xe.astype(np.float64))]
    A = []
    for l in range(layers):
        E.append(E[l] + mlp(6 + 6 * l, np.concatenate([E[l], V[l][recv], V[l][send]], axis=1)))
        agg = np.zeros((n, 128))
        np.add.at(agg, recv, E[l + 1])
        A.append(agg)
        V.append(V[l] + mlp(9 + 6 * l, np.concatenate([V[l], agg], axis=1)))
    Y = mlp(6 + 6 * layers, V[layers])
    return np.stack(V), np.stack(E), np.stack(A), np.concatenate(H), Y


@pytest.mark.parametrize("layers", [1, 3])
def test_backward_at_recorded_activations_matches_oracle(orc, layers):
    """The gradient the oracle computes at recorded activations (the check the
    device backward pass is held to) equals its own end-to-end BF16 gradient
    when the activations come from an independent NumPy forward pass; the two
    forward statements agree on the outputs."""
    lay, params, off, send, recv, Xn, Xe, target, typ = _problem(layers, 60, seed=4)
    loss, G, Yo = orc.gnn_loss(layers, params, lay, Xn, Xe, off, send, recv, target, typ, bf16=True,
                               outputs=True)
    V, E, A, H, Y = _forward_recording(layers, params, lay, Xn, Xe, off, send, recv)
    np.testing.assert_allclose(Y[:, :3], Yo, rtol=1e-6, atol=1e-7)
    l2, G2 = orc.gnn_backward_at(layers, params, lay, Xn, Xe, send, recv, target, typ, V, E, A, H, Y)
    assert l2 == pytest.approx(loss, rel=1e-6)
    assert np.linalg.norm(G2 - G) <= 1e-6 * np.linalg.norm(G)
    # a mask taken from the wrong activation is caught: zero one hidden layer
    H_bad = H.copy()
    H_bad[: 60 * 128] = 0
    _, G3 = orc.gnn_backward_at(layers, params, lay, Xn, Xe, send, recv, target, typ, V, E, A, H_bad, Y)
    assert np.linalg.norm(G3 - G) > 1e-3 * np.linalg.norm(G)
