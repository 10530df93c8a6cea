"""Learned forward-dynamics surrogate (rbd_mlp_* in include/rbd.h) and its training loop.

A multilayer perceptron maps the features of (q, v, tau) -- sin and cos of every revolute
coordinate, the free-flyer quaternion and its rotation matrix, v and tau (73 inputs on the
18-DoF quadruped) -- through hidden SiLU layers (3 x 512 by default) to ddq.  Products run
on BF16 matrix cores with FP32 accumulation; master weights are FP32 and trained with Adam
on the mean squared error of normalised ddq.  `forward_dynamics` keeps the network's ddq
for a configuration only if the RNEA residual ||rnea(q, v, ddq) - tau|| <= tol ||tau||
and recomputes the others with ABA.

Arrays are FP32 torch tensors on the context's GPU in the batch-fastest layout: q is
(nq, B), v / tau / ddq are (nv, B).
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import Context, Engine, Model, _check, _i64, lib

RBD_ASYNC = 0x2
RBD_MLP_LAYERED = 0x40


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _ld(t, rows, name):
    import torch
    if t.dtype != torch.float32 or t.dim() != 2 or t.shape[0] != rows or t.stride(1) != 1:
        raise ValueError(f"{name}: expected a float32 ({rows}, B) tensor with unit batch stride")
    return _i64(t.stride(0) if t.shape[1] > 0 else max(t.shape[1], 1))


def sample_inputs_device(model: Model, B: int, seed: int, tau_sigma=1.0, base_range=1.0, with_a=False):
    """Seeded inputs generated on the GPU (the host sampler's stream, draw for draw)."""
    import torch
    dev = torch.device("cuda", model.ctx.device)
    q = torch.empty((model.nq, B), dtype=torch.float32, device=dev)
    v = torch.empty((model.nv, B), dtype=torch.float32, device=dev)
    tau = torch.empty((model.nv, B), dtype=torch.float32, device=dev)
    a = torch.empty((model.nv, B), dtype=torch.float32, device=dev) if with_a else None
    ld = _i64(max(B, 1))
    _check(lib().rbd_sample_inputs(model.ctx.handle, model.handle, _i64(B), ctypes.c_uint64(seed),
                                   ctypes.c_double(tau_sigma), ctypes.c_double(base_range), _p(q), ld, _p(v), ld,
                                   _p(a) if a is not None else None, ld, _p(tau), ld, ctypes.c_uint32(0)))
    return (q, v, a, tau) if with_a else (q, v, tau)


class Surrogate:
    def __init__(self, model: Model, hidden: int = 512, nhidden: int = 3, seed: int = 0x5EED5EED):
        self.model = model
        self.ctx = model.ctx
        h = ctypes.c_void_p()
        _check(lib().rbd_mlp_create(self.ctx.handle, model.handle, int(hidden), int(nhidden), ctypes.c_uint64(seed),
                                    ctypes.byref(h)))
        self.handle = h
        nin, nout, hid, nh = (ctypes.c_int32() for _ in range(4))
        npar = ctypes.c_int64()
        _check(lib().rbd_mlp_dims(h, ctypes.byref(nin), ctypes.byref(nout), ctypes.byref(hid), ctypes.byref(nh),
                                  ctypes.byref(npar)))
        self.nin, self.nout, self.hidden, self.nhidden = nin.value, nout.value, hid.value, nh.value
        self.nparams = npar.value
        self.norm = (np.zeros(self.nin, np.float32), np.ones(self.nin, np.float32), np.zeros(self.nout, np.float32),
                     np.ones(self.nout, np.float32))

    def __del__(self):
        if getattr(self, "handle", None) is not None:
            try:
                lib().rbd_mlp_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    # parameters -------------------------------------------------------------------
    def layer_shapes(self):
        dims = [self.nin] + [self.hidden] * self.nhidden + [self.nout]
        return [(dims[k + 1], dims[k]) for k in range(len(dims) - 1)]

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        _check(lib().rbd_mlp_get_params(self.handle, out.ctypes.data_as(ctypes.c_void_p), _i64(self.nparams)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(lib().rbd_mlp_set_params(self.handle, p.ctypes.data_as(ctypes.c_void_p), _i64(self.nparams)))

    def set_normalization(self, xmean, xstd, ymean, ystd):
        arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (xmean, xstd, ymean, ystd)]
        _check(lib().rbd_mlp_set_normalization(self.handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
        self.norm = tuple(arrs)

    # device calls -----------------------------------------------------------------
    def features(self, q, v, tau):
        import torch
        B = q.shape[1]
        X = torch.empty((B, self.nin), dtype=torch.float32, device=q.device)
        _check(lib().rbd_mlp_features(self.ctx.handle, self.handle, _i64(B), _p(q), _ld(q, self.model.nq, "q"), _p(v),
                                      _ld(v, self.model.nv, "v"), _p(tau), _ld(tau, self.model.nv, "tau"), _p(X),
                                      ctypes.c_uint32(0)))
        return X

    def fit_normalization(self, q, v, tau, ddq, n=262144):
        """Per-feature and per-output mean / std from (up to) n samples."""
        X = self.features(q[:, :n].contiguous(), v[:, :n].contiguous(), tau[:, :n].contiguous()).double()
        Y = ddq[:, :n].double()
        xs = X.std(0).clamp_min(1e-6)
        ys = Y.std(1).clamp_min(1e-6)
        self.set_normalization(X.mean(0).cpu().numpy(), xs.cpu().numpy(), Y.mean(1).cpu().numpy(), ys.cpu().numpy())

    def predict(self, q, v, tau, out=None, async_=False, layered=False):
        """ddq = network(q, v, tau).  Hidden widths 128, 256 and 512 run as one fused kernel;
        layered=True evaluates layer by layer with the training path's kernels."""
        import torch
        B = q.shape[1]
        out = torch.empty((self.nout, B), dtype=torch.float32, device=q.device) if out is None else out
        flags = (RBD_ASYNC if async_ else 0) | (RBD_MLP_LAYERED if layered else 0)
        _check(lib().rbd_mlp_predict(self.ctx.handle, self.handle, _i64(B), _p(q), _ld(q, self.model.nq, "q"), _p(v),
                                     _ld(v, self.model.nv, "v"), _p(tau), _ld(tau, self.model.nv, "tau"), _p(out),
                                     _ld(out, self.nout, "ddq"), ctypes.c_uint32(flags)))
        return out

    def train_step(self, q, v, tau, ddq, lr, loss_acc=None, async_=True):
        B = q.shape[1]
        _check(lib().rbd_mlp_train_step(self.ctx.handle, self.handle, _i64(B), _p(q), _ld(q, self.model.nq, "q"),
                                        _p(v), _ld(v, self.model.nv, "v"), _p(tau), _ld(tau, self.model.nv, "tau"),
                                        _p(ddq), _ld(ddq, self.nout, "ddq"), ctypes.c_float(lr),
                                        _p(loss_acc) if loss_acc is not None else None,
                                        ctypes.c_uint32(RBD_ASYNC if async_ else 0)))

    def gradients(self, q, v, tau, ddq):
        """(loss, gradient of every parameter in packed order) for one minibatch."""
        B = q.shape[1]
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_float()
        _check(lib().rbd_mlp_gradients(self.ctx.handle, self.handle, _i64(B), _p(q), _ld(q, self.model.nq, "q"), _p(v),
                                       _ld(v, self.model.nv, "v"), _p(tau), _ld(tau, self.model.nv, "tau"), _p(ddq),
                                       _ld(ddq, self.nout, "ddq"), g.ctypes.data_as(ctypes.c_void_p),
                                       _i64(self.nparams), ctypes.byref(loss)))
        return loss.value, g

    def forward_dynamics(self, q, v, tau, tol=1e-3, out=None, want_mask=False):
        """(ddq, number of ABA fallbacks[, accepted mask, surrogate residual])."""
        import torch
        B = q.shape[1]
        out = torch.empty((self.nout, B), dtype=torch.float32, device=q.device) if out is None else out
        acc = torch.empty(B, dtype=torch.uint8, device=q.device) if want_mask else None
        res = torch.empty(B, dtype=torch.float32, device=q.device) if want_mask else None
        nfb = ctypes.c_int64()
        _check(lib().rbd_mlp_forward_dynamics(self.ctx.handle, self.handle, _i64(B), _p(q), _ld(q, self.model.nq, "q"),
                                              _p(v), _ld(v, self.model.nv, "v"), _p(tau),
                                              _ld(tau, self.model.nv, "tau"), ctypes.c_float(tol), _p(out),
                                              _ld(out, self.nout, "ddq"), ctypes.byref(nfb),
                                              _p(acc) if acc is not None else None,
                                              _p(res) if res is not None else None))
        return (out, nfb.value, acc, res) if want_mask else (out, nfb.value)


def make_dataset(model: Model, n: int, seed: int, tau_sigma=20.0, chunk=1 << 20):
    """n samples (q, v, tau) drawn on the GPU with targets ddq from the GPU ABA."""
    import torch
    eng = Engine(model, "batch")
    dev = torch.device("cuda", model.ctx.device)
    Q = torch.empty((model.nq, n), dtype=torch.float32, device=dev)
    V = torch.empty((model.nv, n), dtype=torch.float32, device=dev)
    T = torch.empty((model.nv, n), dtype=torch.float32, device=dev)
    D = torch.empty((model.nv, n), dtype=torch.float32, device=dev)
    for k, s in enumerate(range(0, n, chunk)):
        m = min(chunk, n - s)
        q, v, tau = sample_inputs_device(model, m, seed + k, tau_sigma=tau_sigma)
        Q[:, s:s + m] = q
        V[:, s:s + m] = v
        T[:, s:s + m] = tau
        D[:, s:s + m] = eng.aba(q, v, tau)
    model.ctx.synchronize()
    return Q, V, T, D


def train(sur: Surrogate, data, epochs=10, batch=4096, lr=1e-3, lr_min=1e-5, seed=0, log=None):
    """Adam over shuffled minibatches with a cosine learning-rate schedule; returns the
    mean normalised MSE of each epoch and the wall time."""
    import torch
    Q, V, T, D = data
    n = Q.shape[1]
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    gen = torch.Generator(device=Q.device)
    gen.manual_seed(seed)
    loss_acc = torch.zeros(epochs, dtype=torch.float32, device=Q.device)
    t0 = time.perf_counter()
    step = 0
    for e in range(epochs):
        perm = torch.randperm(n, device=Q.device, generator=gen)
        Qs, Vs, Ts, Ds = (x[:, perm] for x in (Q, V, T, D))
        for s in range(steps_per_epoch):
            c = 0.5 * (1 + math.cos(math.pi * step / max(total, 1)))
            sl = slice(s * batch, (s + 1) * batch)
            sur.train_step(Qs[:, sl], Vs[:, sl], Ts[:, sl], Ds[:, sl], lr_min + (lr - lr_min) * c,
                           loss_acc=loss_acc[e:e + 1])
            step += 1
        if log:
            sur.ctx.synchronize()
            log(e, float(loss_acc[e].item()) / steps_per_epoch)
    sur.ctx.synchronize()
    return (loss_acc / max(steps_per_epoch, 1)).cpu().numpy(), time.perf_counter() - t0
