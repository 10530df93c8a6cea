"""rbd_aba_derivatives_noalloc (the allocation-free variant of rbd_aba_derivatives, its
factor and ddq in a file-scope device array) through the C ABI: two calls in a row, the
second with a smaller batch, each against the allocating entry point on the same inputs
(the same kernels run, so the outputs must agree bit for bit), and a batch beyond the
scratch size rejected with an argument error before any work is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

import pinocchio_amd as pa
from pinocchio_amd import models as M

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _call(name, ctx, model, B, Q, V, T, outs, ddq):
    fn = getattr(pa.lib(), name)
    return fn(ctx.handle, model.handle, ctypes.c_int64(B), pa.RBD_F32, _p(Q), ctypes.c_int64(B), _p(V),
              ctypes.c_int64(B), _p(T), ctypes.c_int64(B), *[_p(o) for o in outs], ctypes.c_int64(B), _p(ddq),
              ctypes.c_int64(B), ctypes.c_uint32(pa.RBD_LAYOUT_BATCH))


def _inputs(desc, B, seed):
    q, v, _, tau = M.sample_inputs(desc, B, seed)
    return [torch.tensor(x.astype(np.float32), device="cuda").contiguous() for x in (q, v, tau)]


def test_noalloc_matches_allocating_entry_point(gpu_ctx):
    desc = M.humanoid36()
    model = pa.Model(gpu_ctx, desc)
    nv = desc.nv
    for B, seed in ((5000, M.INPUT_SEED_BASE + 700), (1200, M.INPUT_SEED_BASE + 701)):
        print(f"noalloc vs allocating: humanoid36 B={B}", flush=True)
        Q, V, T = _inputs(desc, B, seed)
        ref = [torch.empty((nv * nv, B), device="cuda") for _ in range(3)]
        ref_ddq = torch.empty((nv, B), device="cuda")
        assert _call("rbd_aba_derivatives", gpu_ctx, model, B, Q, V, T, ref, ref_ddq) == 0, pa.lib().rbd_last_error()
        got = [torch.full((nv * nv, B), float("nan"), device="cuda") for _ in range(3)]
        got_ddq = torch.full((nv, B), float("nan"), device="cuda")
        assert _call("rbd_aba_derivatives_noalloc", gpu_ctx, model, B, Q, V, T, got, got_ddq) == 0, \
            pa.lib().rbd_last_error()
        for r, g in zip(ref + [ref_ddq], got + [got_ddq]):
            assert torch.equal(r, g)
        # ddq omitted: it goes to the scratch too
        got2 = [torch.full((nv * nv, B), float("nan"), device="cuda") for _ in range(3)]
        assert _call("rbd_aba_derivatives_noalloc", gpu_ctx, model, B, Q, V, T, got2, None) == 0
        for r, g in zip(ref, got2):
            assert torch.equal(r, g)


def test_noalloc_rejects_batches_beyond_the_scratch(gpu_ctx):
    desc = M.humanoid36()
    model = pa.Model(gpu_ctx, desc)
    nv = desc.nv
    B = 50000  # 64 MiB holds 44,979 humanoid configurations
    Q, V, T = (torch.zeros((r, B), device="cuda") for r in (desc.nq, nv, nv))
    outs = [torch.zeros((nv * nv, B), device="cuda") for _ in range(3)]
    rc = _call("rbd_aba_derivatives_noalloc", gpu_ctx, model, B, Q, V, T, outs, None)
    assert rc == 1 and b"scratch" in pa.lib().rbd_last_error()
    assert all(int(torch.count_nonzero(o)) == 0 for o in outs)  # nothing was enqueued
