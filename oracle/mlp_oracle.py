"""CPU oracle of the forward-dynamics surrogate: features, MLP forward pass, MSE loss and
its gradient by back-propagation, in FP64 numpy.

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the benchmark's CPU
baseline as the checker of the GPU surrogate; the library never calls it.

The surrogate has no counterpart in the reference; what it must reproduce is fixed by
its own definition (the feature map, y = W x + b layers with SiLU, the normalised MSE).
`bf16=True` rounds operands to bfloat16 (round to nearest even) exactly where the GPU
kernels do -- normalised inputs, weights, activations, output and hidden gradients --
so the GPU can be compared at FP32-accumulation accuracy.  The back-propagation is pinned
by central finite differences of the loss (tests/test_mlp_oracle.py), the feature map by
closed forms (sin^2 + cos^2 = 1, the rotation matrix of a quaternion from an independent
library routine).
"""
from __future__ import annotations

import numpy as np


def bf16_round(x):
    """float64 -> nearest bfloat16 value (ties to even), via float32 as the GPU converts."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def quat_to_rot(q):
    """Rotation matrices (..., 3, 3) of quaternions [..., x, y, z, w] (no renormalisation)."""
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def features(desc, q, v, tau):
    """(B, nin) raw features; q (nq, B), v / tau (nv, B).  Order: sin of every revolute
    coordinate, their cos, then (free-flyer root) quaternion and rotation matrix row-major,
    then v and tau."""
    from pinocchio_amd.models import JT_FREEFLYER
    iq = desc.idx_q
    rev = [int(iq[j]) for j in range(1, desc.njoints) if desc.jtype[j] != JT_FREEFLYER]
    ff = [j for j in range(1, desc.njoints) if desc.jtype[j] == JT_FREEFLYER]
    cols = [np.sin(q[rev]).T, np.cos(q[rev]).T]
    if ff:
        j = ff[0]
        quat = q[iq[j] + 3:iq[j] + 7].T
        cols += [quat, quat_to_rot(quat).reshape(-1, 9)]
    cols += [v.T, tau.T]
    return np.concatenate(cols, axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params, shapes):
    """Packed parameters (per layer W row-major then b) -> [(W, b)]."""
    out, o = [], 0
    for (n_out, n_in) in shapes:
        W = params[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = params[o:o + n_out]
        o += n_out
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    assert o == len(params)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def normalise_inputs(X, xmean, xstd, bf16=False):
    inv = (1.0 / np.asarray(xstd, np.float32)).astype(np.float64) if bf16 else 1.0 / np.asarray(xstd, np.float64)
    Xn = (X - xmean) * inv
    return bf16_round(Xn) if bf16 else Xn


def forward(layers, Xn, bf16=False):
    """Returns (output, pre-activations Z, activations A) with A[0] = Xn."""
    r = bf16_round if bf16 else (lambda x: x)
    A, Z = [Xn], []
    h = Xn
    for k, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        if k < len(layers) - 1:
            Z.append(z)
            h = r(silu(z))
            A.append(h)
        else:
            return z, Z, A


def loss_and_grad(layers, Xn, Yn, bf16=False):
    """Mean squared error over real outputs and its gradient [(dW, db)] per layer."""
    r = bf16_round if bf16 else (lambda x: x)
    out, Z, A = forward(layers, Xn, bf16)
    B, nv = Yn.shape
    diff = out - Yn
    loss = float((diff ** 2).sum() / (B * nv))
    d = r(diff * (2.0 / (B * nv)))
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        grads[k] = (d.T @ A[k], d.sum(0))
        if k > 0:
            d = r((d @ r(W)) * dsilu(Z[k - 1]))
    return loss, grads
