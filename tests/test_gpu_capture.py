"""Stream capture of engine calls (rbd_capture_begin / _end, rbd_graph_launch): a recorded
sequence replays to the same results as the eager calls, reading its inputs again at
every replay, and calls that cannot be recorded are refused without breaking the capture.

Each graph is replayed several times with new inputs written in place and eager work on
the same stream between the replays; every replay is compared with eager calls on the
same inputs (bit for bit: the same kernels on the same data).
"""
import ctypes

import numpy as np
import pytest
import torch

import pinocchio_amd as pa
from pinocchio_amd import models as M

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float32):
    return torch.tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda").contiguous()


def inputs(desc, B, seed):
    q, v, a, tau = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + seed)
    return q, v, a, tau


def test_step_graph_replays_match_eager(gpu_ctx):
    """ABA + CRBA + FP64 tree Cholesky on the humanoid, batch-fastest, device buffers."""
    desc = M.humanoid36()
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, "batch")
    B, nv = 4096, desc.nv
    q, v, _, tau = inputs(desc, B, 800)
    Q, V, T = dev(q), dev(v), dev(tau)
    ddq = torch.empty((nv, B), device="cuda")
    Mb = torch.empty((nv * nv, B), device="cuda")
    nnz = len(model.cholesky_pattern()[0])
    U = torch.empty((nnz, B), dtype=torch.float64, device="cuda")
    D = torch.empty((nv, B), dtype=torch.float64, device="cuda")
    L = pa.lib()

    def step():
        eng.aba(Q, V, T, out=ddq, async_=True)
        eng.crba(Q, out=Mb, async_=True)
        pa._check(L.rbd_cholesky_decompose(gpu_ctx.handle, model.handle, ctypes.c_int64(B), 0,
                                           ctypes.c_void_p(Mb.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_void_p(U.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_void_p(D.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_uint32(pa.RBD_LAYOUT_BATCH | pa.RBD_ASYNC)))

    step()
    gpu_ctx.synchronize()
    with gpu_ctx.capture() as cap:
        step()
    g = cap.graph
    for r in range(3):
        q2, v2, _, t2 = inputs(desc, B, 801 + r)
        Q.copy_(dev(q2))
        V.copy_(dev(v2))
        T.copy_(dev(t2))
        torch.cuda.synchronize()
        # eager reference on the same inputs
        step()
        gpu_ctx.synchronize()
        ref = [x.clone() for x in (ddq, Mb, U, D)]
        for x in (ddq, Mb, U, D):
            x.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        for name, x, y in zip(("ddq", "M", "U", "D"), (ddq, Mb, U, D), ref):
            assert torch.equal(x, y), (r, name)
    g.close()


def test_aba_derivatives_and_runtime_topology_graphs(gpu_ctx):
    """The ABA-derivative pipeline (workspace reserved by an eager call) and a
    run-time-topology tree (lane memory reserved by an eager call), captured together."""
    desc = M.humanoid36()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), "batch")
    tree = M.random_tree(22, M.MODEL_SEED_BASE + 106, free_flyer=True, compact=True)
    ert = pa.Engine(pa.Model(gpu_ctx, tree), "batch")
    B = 2048
    q, v, _, tau = inputs(desc, B, 810)
    qt, vt, at, tt = inputs(tree, B, 811)
    Q, V, T, Qt, Vt, At, Tt = (dev(x) for x in (q, v, tau, qt, vt, at, tt))
    nv = desc.nv
    outs = [torch.empty((nv * nv, B), device="cuda") for _ in range(3)] + [torch.empty((nv, B), device="cuda")]
    ddq_t = torch.empty((tree.nv, B), device="cuda")
    tau_t = torch.empty((tree.nv, B), device="cuda")

    def calls():
        eng.aba_derivatives(Q, V, T, out=outs[:3], ddq_out=outs[3], async_=True)
        ert.aba(Qt, Vt, Tt, out=ddq_t, async_=True)
        ert.rnea(Qt, Vt, At, out=tau_t, async_=True)

    calls()
    gpu_ctx.synchronize()
    with gpu_ctx.capture() as cap:
        calls()
    for r in range(3):
        for X, x in ((Q, inputs(desc, B, 812 + r)[0]), (Qt, inputs(tree, B, 815 + r)[0])):
            X.copy_(dev(x))
        torch.cuda.synchronize()
        calls()
        gpu_ctx.synchronize()
        ref = [x.clone() for x in outs + [ddq_t, tau_t]]
        for x in outs + [ddq_t, tau_t]:
            x.zero_()
        torch.cuda.synchronize()
        cap.graph.launch()
        for k, (x, y) in enumerate(zip(outs + [ddq_t, tau_t], ref)):
            assert torch.equal(x, y), (r, k)
    cap.graph.close()


def test_refusals_leave_the_context_usable():
    """On a fresh context (nothing reserved yet): host arrays, a first call of a size,
    frame tables and a nested capture are refused; the capture still ends cleanly and the
    context keeps running eagerly."""
    ctx = pa.Context(0)
    desc = M.humanoid36()
    model = pa.Model(ctx, desc)
    eng = pa.Engine(model, "batch")
    B = 512
    q, v, _, tau = inputs(desc, B, 820)
    Q, V, T = dev(q), dev(v), dev(tau)
    out = torch.empty((desc.nv, B), device="cuda")
    eng.aba(Q, V, T, out=out)
    L = pa.lib()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    qh, vh, th = (np.asfortranarray(x) for x in (q, v, tau))
    oh = np.zeros((desc.nv, B), order="F")
    with pytest.raises(ValueError):  # host arrays: staged through device buffers
        with ctx.capture():
            pa._check(L.rbd_aba(ctx.handle, model.handle, ctypes.c_int64(B), pa.RBD_F64, p(qh),
                                ctypes.c_int64(desc.nq), p(vh), ctypes.c_int64(desc.nv), p(th),
                                ctypes.c_int64(desc.nv), p(oh), ctypes.c_int64(desc.nv),
                                ctypes.c_uint32(pa.RBD_HOST_PTRS)))
    outs = [torch.empty((desc.nv ** 2, B), device="cuda") for _ in range(3)]
    with pytest.raises(ValueError):  # its workspace was never reserved
        with ctx.capture():
            eng.aba_derivatives(Q, V, T, out=outs, async_=True)
    with pytest.raises(ValueError):  # host-side frame tables
        with ctx.capture():
            eng.frames(Q, [7], np.concatenate([np.eye(3).ravel(), np.zeros(3)])[None])
    with ctx.capture() as cap:
        with pytest.raises(ValueError):  # nested
            pa._check(L.rbd_capture_begin(ctx.handle))
        eng.aba(Q, V, T, out=out, async_=True)
    ref = eng.aba(Q, V, T)
    out.zero_()
    torch.cuda.synchronize()
    cap.graph.launch()
    assert torch.equal(out, ref)
    cap.graph.close()
    eng.aba_derivatives(Q, V, T, out=outs)  # eager calls still run
