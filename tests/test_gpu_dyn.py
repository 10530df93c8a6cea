"""GPU parity of batched RNEA and ABA against the FP64 oracle (through the C ABI).

Bars: FP64 RNEA within 1e-12 relative (per configuration, inf-norm); FP32 RNEA within
1e-5; FP32 ABA within 1e-4 (the stated FP32 target), the oracle being run in FP64 on
the FP32-rounded inputs.  Cases: every bundled topology, both array layouts, ragged
batches (0, 1, 63, 65, 1000), the full 65536-configuration humanoid batch, the host-
pointer drop-in path (rneaInParallel / abaInParallel) with the reference's argument
errors, and run-to-run determinism.
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols

pytestmark = pytest.mark.gpu

MODELS = ["pendulum", "free_body", "manipulator6", "arm7", "quadruped18", "humanoid36", "random60"]


def build(name):
    return M.SAMPLE_MODELS[name]() if name in M.SAMPLE_MODELS else getattr(M, name)()


def to_dev(x, layout, dtype=torch.float32):
    a = x.T if layout == "col" else x
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def to_host(t, layout):
    a = t.double().cpu().numpy()
    return a.T if layout == "col" else a


_models = {}


def gpu_model(ctx, name):
    if name not in _models:
        desc = build(name)
        _models[name] = (desc, pa.Model(ctx, desc))
    return _models[name]


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", MODELS)
def test_rnea_aba_f32(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 1000
    q, v, a, tau = f32_round(*inputs(desc, B, k=1))
    ref_tau = oracle.rnea(desc, q, v, a)
    ref_ddq = oracle.aba(desc, q, v, tau)
    Q, V, A, T = (to_dev(x, layout) for x in (q, v, a, tau))
    tau_g = to_host(eng.rnea(Q, V, A), layout)
    ddq_g = to_host(eng.aba(Q, V, T), layout)
    e_r = rel_err_cols(tau_g, ref_tau)
    e_a = rel_err_cols(ddq_g, ref_ddq)
    print(f"{name} {layout}: rnea max {e_r.max():.3e}  aba max {e_a.max():.3e} p50 {np.median(e_a):.3e}")
    assert e_r.max() < 1e-5
    assert e_a.max() < 1e-4


@pytest.mark.parametrize("name", ["pendulum", "manipulator6", "quadruped18", "humanoid36", "random60"])
def test_rnea_f64(gpu_ctx, name):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, "col")
    B = 4096
    q, v, a, _ = inputs(desc, B, k=2)
    ref = oracle.rnea(desc, q, v, a)
    out = to_host(eng.rnea(*(to_dev(x, "col", torch.float64) for x in (q, v, a))), "col")
    err = rel_err_cols(out, ref)
    print(f"{name} f64 rnea max rel err {err.max():.3e}")
    assert err.max() < 1e-12


@pytest.mark.parametrize("B", [0, 1, 63, 65])
def test_ragged_batches(gpu_ctx, B):
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    eng = pa.Engine(model, "col")
    q, v, a, tau = f32_round(*inputs(desc, max(B, 1), k=3))
    q, v, a, tau = (x[:, :B] for x in (q, v, a, tau))
    Q, V, A, T = (to_dev(x, "col") for x in (q, v, a, tau))
    tau_g = eng.rnea(Q, V, A)
    ddq_g = eng.aba(Q, V, T)
    if B == 0:
        assert tau_g.numel() == 0 and ddq_g.numel() == 0
        return
    assert rel_err_cols(to_host(tau_g, "col"), oracle.rnea(desc, q, v, a)).max() < 1e-5
    assert rel_err_cols(to_host(ddq_g, "col"), oracle.aba(desc, q, v, tau)).max() < 1e-4


def test_humanoid_full_batch(gpu_ctx):
    """The 65536-configuration humanoid workload, every configuration checked."""
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    B = 65536
    q, v, a, tau = f32_round(*M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 1))
    for layout in ("col", "batch"):
        eng = pa.Engine(model, layout)
        ddq = to_host(eng.aba(*(to_dev(x, layout) for x in (q, v, tau))), layout)
        ref = oracle.aba(desc, q, v, tau)
        e = rel_err_cols(ddq, ref)
        print(f"humanoid36 B={B} {layout}: aba max {e.max():.3e} p50 {np.median(e):.3e} p99.9 {np.quantile(e, 0.999):.3e}")
        assert e.max() < 1e-4


def test_inverse_relation_on_gpu(gpu_ctx):
    """aba(q, v, rnea(q, v, a)) == a, both on the GPU (aba.cpp:124-184 of the reference)."""
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    eng = pa.Engine(model, "batch")
    q, v, a, _ = f32_round(*inputs(desc, 2048, k=4))
    Q, V, A = (to_dev(x, "batch") for x in (q, v, a))
    a2 = to_host(eng.aba(Q, V, eng.rnea(Q, V, A)), "batch")
    assert rel_err_cols(a2, a).max() < 1e-3


def test_determinism(gpu_ctx):
    desc, model = gpu_model(gpu_ctx, "quadruped18")
    eng = pa.Engine(model, "col")
    q, v, a, tau = f32_round(*inputs(desc, 777, k=5))
    Q, V, T = (to_dev(x, "col") for x in (q, v, tau))
    x1 = eng.aba(Q, V, T).clone()
    x2 = eng.aba(Q, V, T)
    assert torch.equal(x1, x2)


def test_dropin_host_api(gpu_ctx):
    """rneaInParallel / abaInParallel mirror: numpy in/out, one column per configuration."""
    desc = M.humanoid36()
    pool = pa.ModelPool(desc, size=4)
    q, v, a, tau = inputs(desc, 300, k=6)
    tau_g = pa.rneaInParallel(4, pool, q, v, a)
    assert rel_err_cols(tau_g, oracle.rnea(desc, q, v, a)).max() < 1e-12
    qf, vf, tf = f32_round(q, v, tau)
    ddq_g = pa.abaInParallel(2, pool, qf, vf, tf, dtype=np.float32)
    assert rel_err_cols(ddq_g, oracle.aba(desc, qf, vf, tf)).max() < 1e-4
    # FP64 by default, like the reference (batch == serial to 1e-12, parallel-aba.cpp:54)
    ddq_64 = pa.abaInParallel(2, pool, q, v, tau)
    assert rel_err_cols(ddq_64, oracle.aba(desc, q, v, tau)).max() < 1e-12
    # output-argument overloads (bindings/python/algorithm/parallel/{rnea,aba}.cpp:54-55)
    out = np.zeros((desc.nv, 300), order="F")
    assert pa.rneaInParallel(4, pool, q, v, a, out) is None
    assert np.array_equal(out, tau_g)
    out2 = np.zeros((desc.nv, 300))
    assert pa.abaInParallel(4, pool, q, v, tau, out2) is None
    assert np.array_equal(out2, ddq_64)
    # argument errors of algorithm/parallel/{rnea,aba}.hpp:53-69
    with pytest.raises(ValueError):
        pa.rneaInParallel(5, pool, q, v, a)  # more threads than the pool size
    with pytest.raises(ValueError):
        pa.rneaInParallel(1, pool, q[:-1], v, a)  # wrong q rows
    with pytest.raises(ValueError):
        pa.abaInParallel(1, pool, q, v, tau[:, :-1])  # column mismatch (also checked for tau)
    with pytest.raises(ValueError):
        pa.rneaInParallel(1, pool, q, v, a, np.zeros((desc.nv, 299)))  # output column mismatch


def test_bad_arguments_c_abi(gpu_ctx):
    import ctypes
    desc, model = gpu_model(gpu_ctx, "arm7")
    L = pa.lib()
    x = torch.zeros((10, 7), device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    # leading dimension smaller than the row count
    rc = L.rbd_rnea(gpu_ctx.handle, model.handle, ctypes.c_int64(10), 0, p, ctypes.c_int64(3), p, ctypes.c_int64(7), p,
                    ctypes.c_int64(7), p, ctypes.c_int64(7), ctypes.c_uint32(0))
    assert rc == 1 and b"leading dimension" in L.rbd_last_error()
    # negative batch
    rc = L.rbd_rnea(gpu_ctx.handle, model.handle, ctypes.c_int64(-1), 0, p, ctypes.c_int64(7), p, ctypes.c_int64(7), p,
                    ctypes.c_int64(7), p, ctypes.c_int64(7), ctypes.c_uint32(0))
    assert rc == 1


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "arm7"])
def test_nonlinear_effects_and_gravity(gpu_ctx, name, dtype, layout):
    """nonLinearEffects = rnea(q, v, 0) and computeGeneralizedGravity = rnea(q, 0, 0)
    (rnea.hpp:105, 133), with the zero inputs taken from the engine's stride-0 zero view:
    bit-identical to rbd_rnea with explicit zero arrays, and against the FP64 oracle."""
    desc = getattr(M, name)()
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, layout)
    B = 301
    q, v, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 78)
    if dtype == torch.float32:
        q, v = f32_round(q, v)
    dev = lambda x: torch.tensor(np.ascontiguousarray(x.T if layout == "col" else x), dtype=dtype, device="cuda")
    Q, V = dev(q), dev(v)
    Z = torch.zeros_like(V)
    nle = eng.nonlinear_effects(Q, V)
    g = eng.generalized_gravity(Q)
    assert torch.equal(nle, eng.rnea(Q, V, Z))
    assert torch.equal(g, eng.rnea(Q, Z, Z))
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for out, vv in ((nle, v), (g, np.zeros_like(v))):
        ref = oracle.rnea(desc, q, vv, np.zeros_like(v))
        o = out.double().cpu().numpy()
        o = o.T if layout == "col" else o
        assert rel_err_cols(o, ref).max() < tol
