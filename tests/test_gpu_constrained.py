"""GPU parity of constrained forward dynamics (forwardDynamics with J, gamma, inv_damping;
contact-dynamics.hxx:90-167) against the FP64 oracle, through the C ABI.

Constraints are contact-like: the WORLD Jacobian rows of end joints (both feet of the
humanoid, 6 rows each; the linear rows of the quadruped's four feet; the arm's end joint),
from the oracle's computeJointJacobians on the same q, plus random dense J.
Bars (per configuration, max-abs relative to the oracle's max-abs):
  FP32 inputs (CRBA and RNEA in FP32, factor and KKT in FP64)   ddq, lambda 2e-4
  FP64 inputs                                                   ddq, lambda 1e-9
The oracle runs in FP64 on the FP32-rounded inputs.
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, rel_err_cols

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-4, "f64": 1e-9}
FEET = {"humanoid36": ([7, 13], "6d"), "quadruped18": ([4, 7, 10, 13], "lin"), "arm7": ([7], "6d")}

_models = {}


def gpu_model(ctx, name):
    if name not in _models:
        desc = getattr(M, name)()
        _models[name] = (desc, pa.Model(ctx, desc))
    return _models[name]


def contact_jacobian(desc, q, joints, kind):
    """(B, nc, nv) stacked WORLD Jacobians of `joints` (columns outside each joint's
    support zeroed, as getJointJacobian returns them)."""
    Jall = oracle.joint_jacobians(desc, q)
    rows = []
    for j in joints:
        mask = np.zeros(desc.nv)
        i = j
        while i > 0:
            mask[desc.idx_v[i]:desc.idx_v[i] + desc.nvs[i]] = 1.0
            i = desc.parents[i]
        Jj = Jall * mask[None, None, :]
        rows.append(Jj[:, :3] if kind == "lin" else Jj)
    return np.concatenate(rows, axis=1)


def dev_rows(x, layout, dtype):
    """(rows, B) host array -> device tensor in the engine layout."""
    a = x.T if layout == "col" else x
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda").contiguous()


def host_rows(t, layout):
    a = t.double().cpu().numpy()
    return a.T if layout == "col" else a


def j_rows(J):
    """(B, nc, nv) -> (nc * nv, B) with element (i, k) at row i + k nc."""
    B = J.shape[0]
    return np.ascontiguousarray(J.transpose(0, 2, 1).reshape(B, -1).T)


def run(ctx, name, B, layout, prec, J=None, gamma_scale=0.1, mu=0.0, seed=0):
    desc, model = gpu_model(ctx, name)
    q, v, _, tau = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 700 + seed)
    if J is None:
        joints, kind = FEET[name]
        J = contact_jacobian(desc, q, joints, kind)
    nc = J.shape[1]
    gamma = gamma_scale * np.random.default_rng(seed).standard_normal((nc, B))
    dtype = torch.float32 if prec == "f32" else torch.float64
    if prec == "f32":
        q, v, tau, J, gamma = f32_round(q, v, tau, J, gamma)
    eng = pa.Engine(model, layout)
    a, lam = eng.constrained_forward_dynamics(dev_rows(q, layout, dtype), dev_rows(v, layout, dtype),
                                              dev_rows(tau, layout, dtype), dev_rows(j_rows(J), layout, dtype),
                                              dev_rows(gamma, layout, dtype), inv_damping=mu)
    torch.cuda.synchronize()
    a, lam = host_rows(a, layout), host_rows(lam, layout)
    a_ref, lam_ref = oracle.constrained_forward_dynamics(desc, q, v, tau, J, gamma, inv_damping=mu)
    return a, lam, a_ref, lam_ref, J, gamma


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "arm7"])
def test_contact_parity(gpu_ctx, name, layout, prec):
    B = 777  # ragged against both the 32- and the 64-configuration tiles
    a, lam, a_ref, lam_ref, J, gamma = run(gpu_ctx, name, B, layout, prec)
    ea, el = rel_err_cols(a, a_ref), rel_err_cols(lam, lam_ref)
    print(f"{name} {layout} {prec}: ddq max {ea.max():.2e} p50 {np.median(ea):.2e}; "
          f"lambda max {el.max():.2e} p50 {np.median(el):.2e}")
    assert ea.max() < TOL[prec] and el.max() < TOL[prec]
    res = np.einsum("bik,kb->ib", J, a) + gamma
    scale = np.abs(np.einsum("bik,kb->ib", np.abs(J), np.abs(a))).max(axis=0) + np.abs(gamma).max(axis=0)
    assert (np.abs(res).max(axis=0) / scale).max() < (1e-4 if prec == "f32" else 1e-10)


@pytest.mark.parametrize("nc", [1, 5, 12])
def test_random_constraints(gpu_ctx, nc):
    desc, _ = gpu_model(gpu_ctx, "humanoid36")
    B = 300
    J = np.random.default_rng(nc).standard_normal((B, nc, desc.nv))
    a, lam, a_ref, lam_ref, _, _ = run(gpu_ctx, "humanoid36", B, "batch", "f64", J=J, gamma_scale=1.0, seed=nc)
    assert rel_err_cols(a, a_ref).max() < 1e-9 and rel_err_cols(lam, lam_ref).max() < 1e-9


def test_damped_redundant_contacts(gpu_ctx):
    """Both feet given twice (rank-deficient J) with damping, as the reference suggests."""
    desc, _ = gpu_model(gpu_ctx, "humanoid36")
    B = 200
    q, _, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 700 + 9)
    J1 = contact_jacobian(desc, q, [7], "6d")
    J = np.concatenate([J1, J1], axis=1)
    a, lam, a_ref, lam_ref, _, _ = run(gpu_ctx, "humanoid36", B, "batch", "f64", J=J, mu=1e-6, seed=9)
    assert rel_err_cols(a, a_ref).max() < 1e-7 and rel_err_cols(lam, lam_ref).max() < 1e-7


def test_large_batch(gpu_ctx):
    """65,536 humanoid configurations, both feet: a sample of 512 against the oracle and
    the constraint residual of every configuration."""
    desc, model = gpu_model(gpu_ctx, "humanoid36")
    B = 65536
    q, v, _, tau = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 777)
    J = contact_jacobian(desc, q, [7, 13], "6d")
    gamma = np.zeros((12, B))
    q, v, tau, J = f32_round(q, v, tau, J)
    eng = pa.Engine(model, "batch")
    dt = torch.float32
    a, lam = eng.constrained_forward_dynamics(dev_rows(q, "batch", dt), dev_rows(v, "batch", dt),
                                              dev_rows(tau, "batch", dt), dev_rows(j_rows(J), "batch", dt),
                                              dev_rows(gamma, "batch", dt))
    torch.cuda.synchronize()
    a, lam = a.double().cpu().numpy(), lam.double().cpu().numpy()
    assert np.isfinite(a).all() and np.isfinite(lam).all()
    res = np.einsum("bik,kb->ib", J, a)
    scale = np.einsum("bik,kb->ib", np.abs(J), np.abs(a)).max(axis=0)
    assert (np.abs(res).max(axis=0) / scale).max() < 1e-4
    idx = np.random.default_rng(5).choice(B, 512, replace=False)
    a_ref, lam_ref = oracle.constrained_forward_dynamics(desc, q[:, idx], v[:, idx], tau[:, idx], J[idx],
                                                         gamma[:, idx])
    assert rel_err_cols(a[:, idx], a_ref).max() < 1e-3 and rel_err_cols(lam[:, idx], lam_ref).max() < 1e-3


def test_argument_errors(gpu_ctx):
    desc, model = gpu_model(gpu_ctx, "arm7")
    eng = pa.Engine(model, "batch")
    B = 64
    z = lambda r: torch.zeros((r, B), device="cuda")
    with pytest.raises(ValueError, match="constraint count"):
        eng.constrained_forward_dynamics(z(desc.nq), z(desc.nv), z(desc.nv), z(13 * desc.nv), z(13))
    with pytest.raises(ValueError):
        eng.constrained_forward_dynamics(z(desc.nq), z(desc.nv), z(desc.nv), z(6 * desc.nv + 1), z(6))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("e", [0.0, 0.5])
@pytest.mark.parametrize("name", ["humanoid36", "quadruped18"])
def test_impulse_parity(gpu_ctx, name, e, layout, prec):
    """impulseDynamics (contact-dynamics.hxx:264-331) on foot contacts against the oracle."""
    desc, model = gpu_model(gpu_ctx, name)
    B = 555
    q, v, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 720)
    joints, kind = FEET[name]
    J = contact_jacobian(desc, q, joints, kind)
    dtype = torch.float32 if prec == "f32" else torch.float64
    if prec == "f32":
        q, v, J = f32_round(q, v, J)
    eng = pa.Engine(model, layout)
    vp, imp = eng.impulse_dynamics(dev_rows(q, layout, dtype), dev_rows(v, layout, dtype),
                                   dev_rows(j_rows(J), layout, dtype), r_coeff=e)
    torch.cuda.synchronize()
    vp, imp = host_rows(vp, layout), host_rows(imp, layout)
    vr, ir = oracle.impulse_dynamics(desc, q, v, J, r_coeff=e)
    ev, ei = rel_err_cols(vp, vr), rel_err_cols(imp, ir)
    print(f"{name} e={e} {layout} {prec}: dq_after max {ev.max():.2e}; impulse max {ei.max():.2e}")
    assert ev.max() < TOL[prec] and ei.max() < TOL[prec]


def test_impulse_host_pointers(gpu_ctx):
    """rbd_impulse_dynamics through the raw C ABI on host arrays (RBD_HOST_PTRS), FP64,
    Eigen column-major layout."""
    import ctypes
    desc, model = gpu_model(gpu_ctx, "quadruped18")
    B = 97
    q, v, _, _ = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 730)
    J = contact_jacobian(desc, q, *FEET["quadruped18"])
    nc, nv = J.shape[1], desc.nv
    qc, vc = np.ascontiguousarray(q.T), np.ascontiguousarray(v.T)
    Jc = np.ascontiguousarray(J.transpose(0, 2, 1).reshape(B, -1))  # per configuration column major
    vp = np.zeros((B, nv))
    imp = np.zeros((B, nc))
    L = pa.lib()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = L.rbd_impulse_dynamics(gpu_ctx.handle, model.handle, ctypes.c_int64(B), pa.RBD_F64, p(qc),
                                ctypes.c_int64(desc.nq), p(vc), ctypes.c_int64(nv), nc, p(Jc), ctypes.c_int64(nc * nv),
                                ctypes.c_double(0.3), ctypes.c_double(0.0), p(vp), ctypes.c_int64(nv), p(imp),
                                ctypes.c_int64(nc), ctypes.c_uint32(pa.RBD_HOST_PTRS))
    assert rc == 0, L.rbd_last_error()
    vr, ir = oracle.impulse_dynamics(desc, q, v, J, r_coeff=0.3)
    assert rel_err_cols(vp.T, vr).max() < 1e-9 and rel_err_cols(imp.T, ir).max() < 1e-9
