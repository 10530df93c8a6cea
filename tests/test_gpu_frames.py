"""GPU parity of frame placements and frame Jacobians (updateFramePlacements,
frames.hxx:15-35; getFrameJacobian, frames.hxx:156-175) against the FP64 oracle, through
the C ABI, in the three reference frames and both layouts.  Frames sit on random joints
(one on the universe) with random placements.
Bars (max-abs relative to the oracle's max-abs, per configuration): FP64 1e-12, FP32 1e-5
(oracle on the FP32-rounded inputs).
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round
from test_oracle_frames import random_frames

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "f64": 1e-12}
_models = {}


def gpu_model(ctx, name):
    if name not in _models:
        desc = getattr(M, name)()
        _models[name] = (desc, pa.Model(ctx, desc))
    return _models[name]


def rel(x, ref):
    d = np.abs(x - ref).reshape(len(ref), -1).max(axis=1)
    return float((d / np.maximum(np.abs(ref).reshape(len(ref), -1).max(axis=1), 1e-30)).max())


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("rf", ["world", "local", "local_world_aligned"])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18"])
def test_frames_parity(gpu_ctx, name, rf, layout, prec):
    desc, model = gpu_model(gpu_ctx, name)
    B, nf = 301, 5
    q, _, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 880)
    par, pl = random_frames(desc, nf, 3)
    if prec == "f32":
        (q,) = f32_round(q)
        pl = np.asarray(pl, dtype=np.float32).astype(np.float64)
    dt = torch.float32 if prec == "f32" else torch.float64
    eng = pa.Engine(model, layout)
    Q = torch.tensor(np.ascontiguousarray(q.T if layout == "col" else q), dtype=dt, device="cuda")
    oMf, J = eng.frames(Q, par, pl, rf)
    torch.cuda.synchronize()
    oMf = oMf.double().cpu().numpy()
    J = J.double().cpu().numpy()
    if layout == "batch":
        oMf, J = oMf.T, J.T
    oMf = oMf.reshape(B, nf, 12)
    J = J.reshape(B, nf, desc.nv, 6).transpose(0, 1, 3, 2)
    ref_o, ref_J = oracle.frames(desc, q, par, pl, rf)
    eo, ej = rel(oMf, ref_o), rel(J, ref_J)
    print(f"{name} {rf} {layout} {prec}: oMf {eo:.2e} J {ej:.2e}")
    assert eo < TOL[prec] and ej < TOL[prec]


def test_frames_arguments(gpu_ctx):
    desc, model = gpu_model(gpu_ctx, "arm7")
    eng = pa.Engine(model, "batch")
    Q = torch.zeros((desc.nq, 64), device="cuda")
    pl = np.zeros((1, 12))
    with pytest.raises(ValueError, match="outside the bounds"):
        eng.frames(Q, [desc.njoints], pl)
    oMf, J = eng.frames(Q, [0], np.concatenate([np.eye(3).ravel(), [1.0, 2.0, 3.0]])[None])
    torch.cuda.synchronize()
    o = oMf.cpu().numpy()
    assert np.allclose(o[:9].T, np.eye(3).ravel()) and np.allclose(o[9:].T, [1.0, 2.0, 3.0])
    assert float(J.abs().max()) == 0.0  # the universe supports no dof


def test_frames_host_pointers(gpu_ctx):
    """The raw C ABI with host arrays (RBD_HOST_PTRS), Eigen column-major layout with a
    leading dimension larger than the rows: the gaps stay as the caller left them."""
    import ctypes
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    B, nf = 130, 3
    q, _, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 881)
    par, pl = random_frames(desc, nf, 4)
    ldo, ldj = 12 * nf + 5, 6 * desc.nv * nf + 3
    oMf = np.full((B, ldo), 7.0)
    J = np.full((B, ldj), 7.0)
    qc = np.ascontiguousarray(q.T)
    L = pa.lib()
    rc = L.rbd_frames(gpu_ctx.handle, model.handle, ctypes.c_int64(B), pa.RBD_F64,
                      qc.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(desc.nq), nf,
                      np.ascontiguousarray(par, dtype=np.int32).ctypes.data_as(ctypes.c_void_p),
                      np.ascontiguousarray(pl).ctypes.data_as(ctypes.c_void_p), 2,
                      oMf.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(ldo),
                      J.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(ldj), ctypes.c_uint32(pa.RBD_HOST_PTRS))
    assert rc == 0, L.rbd_last_error()
    ref_o, ref_J = oracle.frames(desc, q, par, pl, "local_world_aligned")
    assert rel(oMf[:, :12 * nf].reshape(B, nf, 12), ref_o) < 1e-12
    Jg = J[:, :6 * desc.nv * nf].reshape(B, nf, desc.nv, 6).transpose(0, 1, 3, 2)
    assert rel(Jg, ref_J) < 1e-12
    assert (oMf[:, 12 * nf:] == 7.0).all() and (J[:, 6 * desc.nv * nf:] == 7.0).all()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("rf", ["world", "local", "local_world_aligned"])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "arm7"])
def test_frame_motions_parity(gpu_ctx, name, rf, layout, prec):
    """getFrameVelocity / getFrameAcceleration / getFrameClassicalAcceleration
    (frames.hxx:71-149) after forwardKinematics(q, v, a) against the oracle, same bars."""
    desc, model = gpu_model(gpu_ctx, name)
    B, nf = 257, 5
    q, v, a, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 881)
    par, pl = random_frames(desc, nf, 7)
    if prec == "f32":
        q, v, a = f32_round(q, v, a)
        pl = np.asarray(pl, dtype=np.float32).astype(np.float64)
    dt = torch.float32 if prec == "f32" else torch.float64
    eng = pa.Engine(model, layout)
    dev = lambda x: torch.tensor(np.ascontiguousarray(x.T if layout == "col" else x), dtype=dt, device="cuda")
    outs = eng.frame_motions(dev(q), dev(v), par, pl, a=dev(a), reference_frame=rf)
    refs = oracle.frame_motions(desc, q, v, par, pl, a=a, reference_frame=rf)
    for what, o, r in zip(("v", "a", "a_classical"), outs, refs):
        o = o.double().cpu().numpy()
        o = (o.T if layout == "batch" else o).reshape(B, nf, 6)
        e = rel(o, r)
        print(f"{name} {rf} {layout} {prec} {what}: {e:.2e}")
        assert e < TOL[prec], what
    # velocity only (no accelerations requested)
    vf, af, acf = eng.frame_motions(dev(q), dev(v), par, pl, reference_frame=rf)
    assert af is None and acf is None
    o = vf.double().cpu().numpy()
    assert rel((o.T if layout == "batch" else o).reshape(B, nf, 6), refs[0]) < TOL[prec]


def test_frame_motions_arguments(gpu_ctx):
    desc, model = gpu_model(gpu_ctx, "arm7")
    eng = pa.Engine(model, "batch")
    Q = torch.zeros((desc.nq, 64), device="cuda")
    V = torch.ones((desc.nv, 64), device="cuda")
    pl = np.concatenate([np.eye(3).ravel(), [0.1, 0.2, 0.3]])[None]
    with pytest.raises(ValueError, match="outside the bounds"):
        eng.frame_motions(Q, V, [desc.njoints], pl)
    with pytest.raises(ValueError, match="Bad reference frame"):
        _check_raw_rf(eng, Q, V, pl)
    vf, _, _ = eng.frame_motions(Q, V, [0], pl)  # on the universe
    assert float(vf.abs().max()) == 0.0


def _check_raw_rf(eng, Q, V, pl):
    import ctypes
    m = eng.model
    fj = np.zeros(1, np.int32)
    fp = np.ascontiguousarray(pl, dtype=np.float64)
    out = torch.empty((6, Q.shape[1]), device="cuda")
    pa._check(pa.lib().rbd_frame_motions(m.ctx.handle, m.handle, ctypes.c_int64(Q.shape[1]), 0,
                                         ctypes.c_void_p(Q.data_ptr()), ctypes.c_int64(Q.shape[1]),
                                         ctypes.c_void_p(V.data_ptr()), ctypes.c_int64(V.shape[1]), None,
                                         ctypes.c_int64(0), 1, fj.ctypes.data_as(ctypes.c_void_p),
                                         fp.ctypes.data_as(ctypes.c_void_p), 7, ctypes.c_void_p(out.data_ptr()), None,
                                         None, ctypes.c_int64(Q.shape[1]), ctypes.c_uint32(pa.RBD_LAYOUT_BATCH)))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("rf", ["world", "local", "local_world_aligned"])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18"])
def test_frame_jacobian_time_variation_parity(gpu_ctx, name, rf, layout, prec):
    """getFrameJacobianTimeVariation (frames.hxx:293-368) against the oracle (pinned by
    central differences of the frame Jacobian along integrate(q, v t)), same bars."""
    desc, model = gpu_model(gpu_ctx, name)
    B, nf = 201, 4
    q, v, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 882)
    par, pl = random_frames(desc, nf, 9)
    if prec == "f32":
        q, v = f32_round(q, v)
        pl = np.asarray(pl, dtype=np.float32).astype(np.float64)
    dt = torch.float32 if prec == "f32" else torch.float64
    eng = pa.Engine(model, layout)
    dev = lambda x: torch.tensor(np.ascontiguousarray(x.T if layout == "col" else x), dtype=dt, device="cuda")
    dJ = eng.frame_jacobians_time_variation(dev(q), dev(v), par, pl, rf).double().cpu().numpy()
    dJ = (dJ.T if layout == "batch" else dJ).reshape(B, nf, desc.nv, 6).transpose(0, 1, 3, 2)
    ref = oracle.frame_jacobians_time_variation(desc, q, v, par, pl, rf)
    e = rel(dJ, ref)
    print(f"{name} {rf} {layout} {prec}: dJ {e:.2e}")
    assert e < TOL[prec]
