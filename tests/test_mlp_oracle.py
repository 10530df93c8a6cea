"""Pinning of the surrogate's CPU oracle (no GPU).

* back-propagation against central finite differences of the loss (FP64, tiny network),
  and a planted mistake in the activation derivative that the same check must catch;
* the feature map against closed forms and an independent rotation routine;
* bfloat16 rounding against torch's float32 -> bfloat16 conversion and hand-computed ties.
"""
import numpy as np
import pytest
import torch

import oracle.mlp_oracle as mo
from pinocchio_amd import models as M


def tiny(seed=0, nin=5, hidden=7, nh=2, nout=3, B=11):
    rng = np.random.default_rng(seed)
    dims = [nin] + [hidden] * nh + [nout]
    layers = [(rng.normal(size=(dims[k + 1], dims[k])) * 0.7, rng.normal(size=dims[k + 1]) * 0.3)
              for k in range(len(dims) - 1)]
    X = rng.normal(size=(B, nin))
    Y = rng.normal(size=(B, nout))
    return layers, X, Y


def fd_check(layers, X, Y, h=1e-6):
    _, grads = mo.loss_and_grad(layers, X, Y)
    worst = 0.0
    for k, (W, b) in enumerate(layers):
        for P, G in ((W, grads[k][0]), (b, grads[k][1])):
            flat, gflat = P.reshape(-1), G.reshape(-1)
            for i in range(flat.size):
                old = flat[i]
                flat[i] = old + h
                lp, _ = mo.loss_and_grad(layers, X, Y)
                flat[i] = old - h
                lm, _ = mo.loss_and_grad(layers, X, Y)
                flat[i] = old
                fd = (lp - lm) / (2 * h)
                worst = max(worst, abs(fd - gflat[i]) / max(abs(fd), 1e-3))
    return worst


def test_backprop_matches_finite_differences():
    layers, X, Y = tiny()
    assert fd_check(layers, X, Y) < 1e-6


def test_planted_activation_derivative_mistake_is_caught(monkeypatch):
    layers, X, Y = tiny(seed=1)
    monkeypatch.setattr(mo, "dsilu", lambda z: 1.0 / (1.0 + np.exp(-z)))  # sigmoid alone: wrong
    assert fd_check(layers, X, Y) > 1e-2


def test_pack_roundtrip():
    layers, _, _ = tiny(seed=2)
    p = mo.pack(layers)
    back = mo.unpack(p, [W.shape for W, _ in layers])
    for (W, b), (W2, b2) in zip(layers, back):
        assert np.array_equal(W, W2) and np.array_equal(b, b2)


def test_features_closed_forms():
    from scipy.spatial.transform import Rotation
    d = M.quadruped18()
    q, v, a, tau = M.sample_inputs(d, 257, M.INPUT_SEED_BASE + 7, tau_sigma=20.0)
    X = mo.features(d, q, v, tau)
    assert X.shape == (257, 73)
    s, c = X[:, :12], X[:, 12:24]
    assert np.allclose(s ** 2 + c ** 2, 1.0, atol=1e-14)
    assert np.allclose(np.arctan2(s, c), q[7:19].T, atol=1e-12)  # revolute coordinates in (-pi, pi)
    quat = X[:, 24:28]
    assert np.array_equal(quat, q[3:7].T)
    R = X[:, 28:37].reshape(-1, 3, 3)
    assert np.allclose(R, Rotation.from_quat(quat).as_matrix(), atol=1e-12)
    assert np.array_equal(X[:, 37:55], v.T) and np.array_equal(X[:, 55:73], tau.T)


def test_bf16_rounding():
    one = 1.0
    assert mo.bf16_round(one) == 1.0
    assert mo.bf16_round(1 + 2.0 ** -8) == 1.0                       # tie -> even mantissa
    assert mo.bf16_round(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6         # tie -> even mantissa
    assert mo.bf16_round(1 + 2.0 ** -8 + 2.0 ** -20) == 1 + 2.0 ** -7  # above the tie
    x = np.random.default_rng(3).normal(size=10000) * 10.0 ** np.random.default_rng(4).integers(-8, 8, 10000)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    assert np.array_equal(mo.bf16_round(x), ref)


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_shapes_and_linear_limit(bf16):
    # a network whose hidden pre-activations stay tiny is linear: silu(z) ~ z / 2
    layers, X, _ = tiny(seed=5, nh=1)
    W0, b0 = layers[0]
    layers[0] = (W0 * 1e-4, b0 * 0)
    out, Z, A = mo.forward(layers, X, bf16=False)
    W1, b1 = layers[1]
    lin = X @ (W0 * 1e-4).T / 2 @ W1.T + b1
    assert np.allclose(out, lin, atol=1e-7)
    assert len(Z) == 1 and len(A) == 2
