"""GPU parity of CRBA, the FP64 tree Cholesky and the analytical derivatives against the
FP64 oracle (through the C ABI).

Bars (per configuration, max-abs relative to the reference matrix's max-abs):
  CRBA FP32 M                           1e-5
  cholesky::decompose FP64 (of an FP64 M) 1e-10 on U and D; solve 1e-10
  computeRNEADerivatives FP32           1e-5 on dtau_dq, dtau_dv, M
  computeABADerivatives FP32            1e-4 on ddq_dq, ddq_dv, Minv (the stated FP32 target)
The oracle always runs in FP64 on the FP32-rounded inputs.  Both array layouts, ragged
batches, the full 16384-configuration derivative workload and the structural zeros
(blocks of unrelated branches) are covered.
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols, rel_err_mats

pytestmark = pytest.mark.gpu

COMPACT = ["pendulum", "free_body", "manipulator6", "arm7", "quadruped18", "humanoid36"]


def build(name):
    return M.SAMPLE_MODELS[name]() if name in M.SAMPLE_MODELS else getattr(M, name)()


def to_dev(x, layout, dtype=torch.float32):
    a = x.T if layout == "col" else x
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda").contiguous()


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
@pytest.mark.parametrize("name", COMPACT)
def test_crba(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    q, _, _, _ = f32_round(*inputs(desc, 333, k=11))
    Mg = pa.Engine.mats_to_numpy(eng.crba(to_dev(q, layout)), layout, desc.nv)
    ref = oracle.crba(desc, q)
    e = rel_err_mats(Mg, ref)
    print(f"{name} {layout}: crba max {e.max():.3e}")
    assert e.max() < 1e-5
    assert np.array_equal(Mg, Mg.transpose(0, 2, 1))


def test_crba_upper_only(gpu_ctx):
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    eng = pa.Engine(model, "col")
    q, _, _, _ = f32_round(*inputs(desc, 70, k=12))
    out = torch.full((70, 36, 36), 7.0, device="cuda")
    Mg = pa.Engine.mats_to_numpy(eng.crba(to_dev(q, "col"), out=out, upper_only=True), "col", 36)
    ref = oracle.crba(desc, q, symmetrize=False)
    low = np.tril(np.ones((36, 36)), -1).astype(bool)
    low[:6, :6] = False  # the free-flyer's row block is written whole, as the reference does
    assert np.all(Mg[:, low] == 7.0)  # untouched
    assert rel_err_mats(np.where(low, 0, Mg), ref).max() < 1e-5


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["arm7", "quadruped18", "humanoid36"])
def test_cholesky_fp64(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    q, _, _, tau = inputs(desc, 200, k=13)
    Mref = oracle.crba(desc, q)  # FP64 M
    if layout == "col":
        Md = torch.tensor(np.ascontiguousarray(Mref.transpose(0, 2, 1)), dtype=torch.float64, device="cuda")
    else:
        Md = torch.tensor(np.ascontiguousarray(Mref.transpose(2, 1, 0).reshape(desc.nv * desc.nv, -1)),
                          dtype=torch.float64, device="cuda")
    U, D = eng.cholesky_decompose(Md)
    Uo, Do, _, x_ref = oracle.cholesky(desc, Mref, rhs=tau)
    rows, cols = model.cholesky_pattern()
    Uh = U.cpu().numpy() if layout == "col" else U.cpu().numpy().T
    Dh = D.cpu().numpy() if layout == "col" else D.cpu().numpy().T
    np.testing.assert_allclose(Dh, Do, rtol=1e-10)
    np.testing.assert_allclose(Uh, Uo[:, rows, cols], rtol=1e-9, atol=1e-10)
    # the strictly-upper pattern is exactly where the oracle's U is non-zero
    full = np.zeros_like(Uo[0], dtype=bool)
    full[rows, cols] = True
    assert not np.any(np.abs(Uo[:, ~full & ~np.eye(desc.nv, dtype=bool)]) > 0)
    x = to_dev(tau, layout, torch.float64)
    eng.cholesky_solve(U, D, x)
    np.testing.assert_allclose(to_host(x, layout), x_ref, rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["arm7", "humanoid36", "random_tree"])
def test_cholesky_products(gpu_ctx, name, layout):
    """cholesky::Uv / Utv / Uiv / Utiv / Mv with the GPU factor, against dense products of
    the oracle's U and D -- and Mv against M x with CRBA's M (the factor closes the loop)"""
    if name == "random_tree":  # a tree without prebuilt kernels: the factor from the
        desc = M.random_tree(20, M.MODEL_SEED_BASE + 107, free_flyer=True, compact=True)  # run-time kernels
        model = pa.Model(gpu_ctx, desc)
    else:
        desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 257
    q, _, _, tau = inputs(desc, B, k=15)
    Mref = oracle.crba(desc, q)
    if layout == "col":
        Md = torch.tensor(np.ascontiguousarray(Mref.transpose(0, 2, 1)), dtype=torch.float64, device="cuda")
    else:
        Md = torch.tensor(np.ascontiguousarray(Mref.transpose(2, 1, 0).reshape(desc.nv * desc.nv, -1)),
                          dtype=torch.float64, device="cuda")
    U, D = eng.cholesky_decompose(Md)
    Uo, Do, _, _ = oracle.cholesky(desc, Mref)
    want = {"Uv": np.einsum("bij,jb->ib", Uo, tau), "Utv": np.einsum("bji,jb->ib", Uo, tau),
            "Uiv": np.stack([np.linalg.solve(Uo[b], tau[:, b]) for b in range(B)], 1),
            "Utiv": np.stack([np.linalg.solve(Uo[b].T, tau[:, b]) for b in range(B)], 1),
            "Mv": np.einsum("bij,jb->ib", Mref, tau)}
    for op, w in want.items():
        x = to_dev(tau, layout, torch.float64)
        eng.cholesky_apply(op, U, D, x)
        e = rel_err_cols(to_host(x, layout), w)
        print(f"{name} {layout} {op}: {e.max():.2e}")
        assert e.max() < 1e-11, op
    with pytest.raises(ValueError):
        eng.cholesky_apply("Dv", U, D, to_dev(tau, layout, torch.float64))


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["arm7", "quadruped18", "humanoid36", "free_body", "pendulum"])
def test_rnea_derivatives(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 300
    q, v, a, _ = f32_round(*inputs(desc, B, k=14))
    outs = eng.rnea_derivatives(to_dev(q, layout), to_dev(v, layout), to_dev(a, layout))
    dq, dv, Mm = (pa.Engine.mats_to_numpy(o, layout, desc.nv) for o in outs)
    rdq, rdv, rMu, _ = oracle.rnea_derivatives(desc, q, v, a)
    rM = oracle.crba(desc, q)
    for n, x, r in (("dtau_dq", dq, rdq), ("dtau_dv", dv, rdv), ("M", Mm, rM)):
        e = rel_err_mats(x, r)
        print(f"{name} {layout}: {n} max {e.max():.3e}")
        assert e.max() < 1e-5, n
        # structural zeros are exact zeros
        assert np.all(x[:, np.all(r == 0, axis=0)] == 0), n


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["arm7", "quadruped18", "humanoid36", "free_body", "pendulum"])
def test_aba_derivatives(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 300
    q, v, _, tau = f32_round(*inputs(desc, B, k=15))
    Xq, Xv, Mi, ddq = eng.aba_derivatives(to_dev(q, layout), to_dev(v, layout), to_dev(tau, layout))
    xq, xv, mi = (pa.Engine.mats_to_numpy(o, layout, desc.nv) for o in (Xq, Xv, Mi))
    rq, rv, rmi, rddq = oracle.aba_derivatives(desc, q, v, tau)
    for n, x, r in (("ddq_dq", xq, rq), ("ddq_dv", xv, rv), ("Minv", mi, rmi)):
        e = rel_err_mats(x, r)
        print(f"{name} {layout}: {n} max {e.max():.3e} p50 {np.median(e):.3e}")
        assert e.max() < 1e-4, n
    ed = np.abs(to_host(ddq, layout) - rddq).max(0) / np.abs(rddq).max(0)
    assert ed.max() < 1e-4


def test_derivatives_full_batch(gpu_ctx):
    """The 16384-configuration humanoid derivative workload, every configuration checked."""
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    B = 16384
    q, v, a, tau = f32_round(*M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 2))
    eng = pa.Engine(model, "batch")
    Q, V, A, T = (to_dev(x, "batch") for x in (q, v, a, tau))
    dq, dv, Mm = (pa.Engine.mats_to_numpy(o, "batch", 36) for o in eng.rnea_derivatives(Q, V, A))
    rdq, rdv, _, _ = oracle.rnea_derivatives(desc, q, v, a)
    for n, x, r in (("dtau_dq", dq, rdq), ("dtau_dv", dv, rdv)):
        e = rel_err_mats(x, r)
        print(f"humanoid36 B={B}: {n} max {e.max():.3e} p50 {np.median(e):.3e} p99.9 {np.quantile(e, 0.999):.3e}")
        assert e.max() < 1e-5
    del dq, dv, Mm, rdq, rdv
    Xq, Xv, Mi, _ = eng.aba_derivatives(Q, V, T)
    rq, rv, rmi, _ = oracle.aba_derivatives(desc, q, v, tau)
    for n, o, r in (("ddq_dq", Xq, rq), ("ddq_dv", Xv, rv), ("Minv", Mi, rmi)):
        e = rel_err_mats(pa.Engine.mats_to_numpy(o, "batch", 36), r)
        print(f"humanoid36 B={B}: {n} max {e.max():.3e} p50 {np.median(e):.3e} p99.9 {np.quantile(e, 0.999):.3e}")
        assert e.max() < 1e-4


def test_derivative_argument_errors(gpu_ctx):
    desc = M.arm7()
    desc.gravity = np.array([0, 0, -9.81, 0.0, 0.2, 0.0])
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, "col")
    x = torch.zeros((4, 7), device="cuda")
    with pytest.raises(ValueError, match="gravity"):
        eng.rnea_derivatives(x, x, x)
    r = M.random60()
    rm = pa.Model(gpu_ctx, r)
    with pytest.raises(NotImplementedError, match="compact"):
        pa.Engine(rm, "col").crba(torch.zeros((2, 61), device="cuda"))


@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "random60"])
def test_wave_split_variants_agree(gpu_ctx, name):
    """The sweeps pick 1, 2 or 4 waves per 64-configuration tile from the batch size
    (subtree split for ABA, output split for the RNEA derivatives).  The same
    configurations evaluated inside batches of 70000 (1 wave), 40000 (2) and 1000 (4) must
    agree with each other and with the oracle."""
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, "batch")
    n = 1000
    outs = {}
    q0, v0, a0, t0 = f32_round(*M.sample_inputs(desc, 70000, M.INPUT_SEED_BASE + 31))
    for B in (70000, 40000, n):
        q, v, tau = (np.ascontiguousarray(x[:, :B]) for x in (q0, v0, t0))
        Q, V, T = (to_dev(x, "batch") for x in (q, v, tau))
        outs[B] = eng.aba(Q, V, T)[:, :n].double().cpu().numpy()
        if B == n:
            ref = oracle.aba(desc, q, v, tau)
    for B, o in outs.items():
        e = np.abs(o - ref).max(0) / np.maximum(np.abs(ref).max(0), 1e-30)
        assert e.max() < 1e-4, (B, e.max())
    d = np.abs(outs[70000] - outs[n]).max(0) / np.maximum(np.abs(outs[n]).max(0), 1e-30)
    assert d.max() < 2e-5
    if model.compact and name != "random60":
        res = {}
        q0, v0, a0, _ = f32_round(*M.sample_inputs(desc, 65536, M.INPUT_SEED_BASE + 32))
        for B in (65536, n):
            Q, V, A = (to_dev(np.ascontiguousarray(x[:, :B]), "batch") for x in (q0, v0, a0))
            res[B] = [x[:, :n].cpu() for x in eng.rnea_derivatives(Q, V, A)]
        for x1, x4 in zip(res[65536], res[n]):
            # the same arithmetic per entry, only the emitting wave differs
            a1, a4 = x1.double().numpy(), x4.double().numpy()
            assert (np.abs(a1 - a4).max(0) / np.maximum(np.abs(a4).max(0), 1e-30)).max() < 1e-6


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "manipulator6", "pendulum"])
def test_compute_minverse(gpu_ctx, name, layout):
    """computeMinverse against the oracle's M^-1 (FP32 1e-4, FP64 1e-10) and Minv M = I
    (aba.cpp:265-387)."""
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 333
    q, v, a, tau = inputs(desc, B, k=70)
    qf = f32_round(q)[0]
    _, _, rmi32, _ = oracle.aba_derivatives(desc, qf, v, tau)
    Mi32 = pa.Engine.mats_to_numpy(eng.compute_minverse(to_dev(qf, layout)), layout, desc.nv)
    assert rel_err_mats(Mi32, rmi32).max() < 1e-4
    _, _, rmi64, _ = oracle.aba_derivatives(desc, q, v, tau)
    Mi64 = pa.Engine.mats_to_numpy(eng.compute_minverse(to_dev(q, layout, torch.float64)), layout, desc.nv)
    assert rel_err_mats(Mi64, rmi64).max() < 1e-10
    Mref = oracle.crba(desc, q)
    eye = np.einsum("bij,bjk->bik", Mi64, Mref)
    assert np.abs(eye - np.eye(desc.nv)).max() < 1e-9


def test_urdf_simple_humanoid_all_algorithms(gpu_ctx):
    """pinocchio's sample robot (tests/golden/simple_humanoid.urdf) read with a free-flyer
    root runs on its prebuilt topology; every algorithm against the oracle."""
    import os
    from pinocchio_amd import urdf
    desc = urdf.parse(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simple_humanoid.urdf"),
                      root="freeflyer")
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, "batch")
    B = 500
    q, v, a, tau = f32_round(*inputs(desc, B, k=80))
    Q, V, A, T = (to_dev(x, "batch") for x in (q, v, a, tau))
    from helpers import rel_err_cols
    assert rel_err_cols(to_host(eng.rnea(Q, V, A), "batch"), oracle.rnea(desc, q, v, a)).max() < 1e-5
    assert rel_err_cols(to_host(eng.aba(Q, V, T), "batch"), oracle.aba(desc, q, v, tau)).max() < 1e-4
    nv = desc.nv
    assert rel_err_mats(pa.Engine.mats_to_numpy(eng.crba(Q), "batch", nv), oracle.crba(desc, q)).max() < 1e-5
    rdq, rdv, _, _ = oracle.rnea_derivatives(desc, q, v, a)
    for x, r in zip(eng.rnea_derivatives(Q, V, A)[:2], (rdq, rdv)):
        assert rel_err_mats(pa.Engine.mats_to_numpy(x, "batch", nv), r).max() < 1e-5
    rq, rv, rmi, _ = oracle.aba_derivatives(desc, q, v, tau)
    for x, r in zip(eng.aba_derivatives(Q, V, T)[:3], (rq, rv, rmi)):
        assert rel_err_mats(pa.Engine.mats_to_numpy(x, "batch", nv), r).max() < 1e-4


def test_largest_batches_use_64bit_offsets(gpu_ctx):
    """A batch whose matrix output exceeds 4 GiB (CRBA, 1,000,003 humanoid configurations,
    5.2 GB of M) and a million-configuration ABA: sampled configurations -- first, last,
    across the 2^31-byte boundary, random -- against the oracle."""
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    eng = pa.Engine(model, "batch")
    B = 1_000_003
    q, v, a, tau = f32_round(*M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 33))
    Q, V, T = (to_dev(x, "batch") for x in (q, v, tau))
    rng = np.random.default_rng(5)
    nv = desc.nv
    cross = (2 ** 31) // (4 * B) + 1  # first matrix row whose byte offset exceeds 2^31
    cols = np.unique(np.concatenate([[0, 1, 63, 64, B - 65, B - 2, B - 1], rng.integers(0, B, 200)]))
    Mg = eng.crba(Q)
    assert Mg.shape == (nv * nv, B) and cross < nv * nv
    sub = Mg[:, torch.tensor(cols, device="cuda")].double().cpu().numpy()
    ref = oracle.crba(desc, q[:, cols])
    assert rel_err_mats(pa.Engine.mats_to_numpy(torch.tensor(sub), "batch", nv), ref).max() < 1e-5
    del Mg
    torch.cuda.empty_cache()
    dd = eng.aba(Q, V, T)[:, torch.tensor(cols, device="cuda")].double().cpu().numpy()
    from helpers import rel_err_cols
    assert rel_err_cols(dd, oracle.aba(desc, q[:, cols], v[:, cols], tau[:, cols])).max() < 1e-4


def test_compute_minverse_host_pointers_padded(gpu_ctx):
    """rbd_compute_minverse through the raw C ABI on host arrays (RBD_HOST_PTRS), Eigen
    column-major layout with a leading dimension larger than nv * nv: the gaps stay as the
    caller left them (FP32 and FP64)."""
    import ctypes
    desc, model = gpu_model(gpu_ctx, "quadruped18")
    nv = desc.nv
    B = 97
    q, v, _, tau = inputs(desc, B, k=71)
    L = pa.lib()
    for code, dt, bar in ((pa.RBD_F32, np.float32, 1e-4), (pa.RBD_F64, np.float64, 1e-10)):
        qc = np.ascontiguousarray(q.T.astype(dt))
        ld = nv * nv + 7
        out = np.full((B, ld), 7.0, dtype=dt)
        rc = L.rbd_compute_minverse(gpu_ctx.handle, model.handle, ctypes.c_int64(B), code,
                                    qc.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(desc.nq),
                                    out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(ld),
                                    ctypes.c_uint32(pa.RBD_HOST_PTRS))
        assert rc == 0, L.rbd_last_error()
        qq = qc.T.astype(np.float64)
        _, _, rmi, _ = oracle.aba_derivatives(desc, qq, v, tau)
        Mi = out[:, :nv * nv].astype(np.float64).reshape(B, nv, nv).transpose(0, 2, 1)
        assert rel_err_mats(Mi, rmi).max() < bar
        assert (out[:, nv * nv:] == 7.0).all()


def test_aba_derivatives_unstaged_factor_path(gpu_ctx):
    """Above 32768 configurations the FP32 tree factor of the ABA-derivative pipeline reads
    M lazily instead of staging it in LDS (k_chol instead of k_chol_staged): 40000
    quadruped configurations, every 7th checked against the oracle."""
    desc, model = gpu_model(gpu_ctx, "quadruped18")
    eng = pa.Engine(model, "batch")
    B = 40000
    q, v, _, tau = f32_round(*inputs(desc, B, k=72))
    Xq, Xv, Mi, ddq = eng.aba_derivatives(to_dev(q, "batch"), to_dev(v, "batch"), to_dev(tau, "batch"))
    idx = np.arange(0, B, 7)
    rq, rv, rmi, _ = oracle.aba_derivatives(desc, q[:, idx], v[:, idx], tau[:, idx])
    for n, o, r in (("ddq_dq", Xq, rq), ("ddq_dv", Xv, rv), ("Minv", Mi, rmi)):
        x = pa.Engine.mats_to_numpy(o[:, idx].contiguous(), "batch", desc.nv)
        e = rel_err_mats(x, r)
        print(f"quadruped18 B={B}: {n} max {e.max():.3e}")
        assert e.max() < 1e-4, n

