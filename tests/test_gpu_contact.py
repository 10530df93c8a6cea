"""GPU parity of the Delassus matrix (computeDelassusMatrix, delassus.hxx:118-300) and of the
batched frictional contact solvers (PGSContactSolverTpl, pgs-solver.hxx:19-141;
ADMMContactSolverTpl, admm-solver.hxx:22-335) against the FP64 oracle, through the C ABI.

Bars:
  Delassus, FP64 inputs                          1e-10 of the matrix's max-abs
  Delassus, FP32 inputs (CRBA in FP32)           2e-4  (the oracle on the FP32-rounded q)
  PGS (same sweep order, FP64)                   1e-9 of the solution's max-abs; iteration
                                                 counts equal on >= 99 % of the problems
  ADMM (FP64; sums across lanes in another order, so a near-tie rho decision can take
  another path)                                  1e-7 on converged problems; iteration
                                                 counts equal on >= 95 %
Ragged batches (777 problems: a partial last workgroup), one to four cones, both layouts
of the C ABI, host pointers, argument errors, and a whole contact step: the humanoid's
Delassus matrix of four foot points and its free velocity, solved by ADMM on the GPU, meets
the Coulomb conditions.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, inputs

pytestmark = pytest.mark.gpu

B = 777


def dominant(rng, n):
    A = rng.normal(size=(n, n)) * 0.15
    return A @ A.T + np.diag(rng.uniform(1.0, 2.0, n))


def problems(nc, seed, coupled=False):
    rng = np.random.default_rng(seed)
    n = 3 * nc
    if coupled:
        Gs = []
        for _ in range(B):
            Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
            Gs.append((Q * np.geomspace(1.0, 20.0, n)) @ Q.T)
        G = np.stack(Gs)
    else:
        G = np.stack([dominant(rng, n) for _ in range(B)])
    g = rng.normal(size=(B, n))
    g[:, 2::3] -= 1.0
    mu = rng.uniform(0.2, 1.0, size=(B, nc))
    return G, g, mu


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


# ---------------------------------------------------------------------------------------
# Delassus
# ---------------------------------------------------------------------------------------
FOOT = np.concatenate([np.eye(3).ravel(), [0.03, -0.01, -0.08]])
SETS = {
    "points_local": [dict(type=0, joint1=7, reference_frame=1, placement1=FOOT),
                     dict(type=0, joint1=13, reference_frame=1, placement1=FOOT)],
    "frames_local": [dict(type=1, joint1=7, reference_frame=1, placement1=FOOT),
                     dict(type=1, joint1=13, reference_frame=1, placement1=FOOT)],
    "points_lwa_hands": [dict(type=0, joint1=7, reference_frame=2, placement1=FOOT),
                         dict(type=0, joint1=22, joint2=29, reference_frame=1)],
}


def to_models(cons):
    return [pa.RigidConstraintModel(c["type"], c["joint1"], c.get("placement1", pa._IDENTITY), c.get("joint2", 0),
                                    c.get("placement2", pa._IDENTITY), c["reference_frame"]) for c in cons]


@pytest.mark.parametrize("mu", [0.0, 0.5])
@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", list(SETS))
def test_delassus_matches_oracle(gpu_ctx, name, layout, mu):
    desc = M.humanoid36()
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, layout)
    q = inputs(desc, B, k=90)[0]
    cons = SETS[name]
    nr = sum(6 if c["type"] == 1 else 3 for c in cons)
    for dtype, tol in ((torch.float64, 1e-10), (torch.float32, 2e-4)):
        qq = f32_round(q)[0] if dtype == torch.float32 else q
        want = oracle.delassus(desc, qq, cons, mu=mu)  # (B, nr, nr)
        Q = torch.tensor(np.ascontiguousarray(qq.T if layout == "col" else qq), dtype=dtype, device="cuda")
        G = eng.delassus(Q, to_models(cons), mu=mu).double().cpu().numpy()
        G = G.reshape(B, nr * nr) if layout == "col" else G.T
        got = G.reshape(B, nr, nr).transpose(0, 2, 1)  # element (i, j) at i + j nr
        err = np.abs(got - want).max(axis=(1, 2)) / np.abs(want).max(axis=(1, 2))
        print(f"{name} {layout} mu={mu} {dtype}: max rel err {err.max():.2e}")
        assert err.max() < tol


def test_delassus_errors(gpu_ctx):
    desc = M.humanoid36()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), "batch")
    Q = torch.tensor(inputs(desc, 8, k=91)[0], dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        eng.delassus(Q, to_models(SETS["points_local"]), mu=-1.0)
    with pytest.raises(ValueError):
        eng.delassus(Q, to_models(SETS["frames_local"] * 2))  # 24 rows


@pytest.mark.parametrize("scaled", [False, True])
@pytest.mark.parametrize("layout", ["col", "batch"])
def test_damped_delassus_inverse(gpu_ctx, layout, scaled):
    """computeDampedDelassusMatrixInverse against numpy's inverse of the pinned Delassus"""
    desc = M.humanoid36()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), layout)
    q = inputs(desc, B, k=92)[0]
    cons = SETS["points_lwa_hands"]
    nr, mu = 6, 0.05
    for dtype, tol in ((torch.float64, 1e-10), (torch.float32, 2e-4)):
        qq = f32_round(q)[0] if dtype == torch.float32 else q
        D = oracle.delassus(desc, qq, cons)
        want = np.linalg.inv(D + mu * np.eye(nr)) * (mu if scaled else 1.0)
        Q = torch.tensor(np.ascontiguousarray(qq.T if layout == "col" else qq), dtype=dtype, device="cuda")
        G = eng.damped_delassus_inverse(Q, to_models(cons), mu, scaled=scaled).double().cpu().numpy()
        G = G.reshape(B, nr * nr) if layout == "col" else G.T
        got = G.reshape(B, nr, nr).transpose(0, 2, 1)
        err = np.abs(got - want).max(axis=(1, 2)) / np.abs(want).max(axis=(1, 2))
        print(f"damped inverse {layout} scaled={scaled} {dtype}: {err.max():.2e}")
        assert err.max() < tol
    with pytest.raises(ValueError):
        eng.damped_delassus_inverse(Q, to_models(cons), 0.0)


@pytest.mark.parametrize("mu", [0.0, 0.01])
@pytest.mark.parametrize("layout", ["col", "batch"])
def test_contact_cholesky_solve(gpu_ctx, layout, mu):
    """x = K^-1 rhs, K = [[-mu I, J], [J^T, M]], against the oracle's dense elimination of K"""
    desc = M.humanoid36()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), layout)
    q = inputs(desc, B, k=93)[0]
    cons = SETS["points_local"] + [dict(type=1, joint1=22, reference_frame=2)]
    n = 12 + desc.nv
    rhs = np.random.default_rng(5).normal(size=(n, B))
    for dtype, tol in ((torch.float64, 1e-9), (torch.float32, 2e-4)):
        qq, rr = f32_round(q, rhs) if dtype == torch.float32 else (q, rhs)
        want = oracle.contact_cholesky_solve(desc, qq, cons, mu, rr)
        dev_ = lambda x: torch.tensor(np.ascontiguousarray(x.T if layout == "col" else x), dtype=dtype, device="cuda")
        x = eng.contact_cholesky_solve(dev_(qq), to_models(cons), mu, dev_(rr)).double().cpu().numpy()
        x = x.T if layout == "col" else x
        err = np.abs(x - want).max(axis=0) / np.abs(want).max(axis=0)
        print(f"KKT solve {layout} mu={mu} {dtype}: {err.max():.2e}")
        assert err.max() < tol


# ---------------------------------------------------------------------------------------
# PGS
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("nc", [1, 2, 3, 4])
def test_pgs_matches_oracle(gpu_ctx, nc):
    G, g, mu = problems(nc, 100 + nc)
    x0 = np.zeros((B, 3 * nc))
    x0[::7, 2::3] = 0.5  # some problems start from a guess
    kw = dict(max_it=400, absolute_precision=1e-10, relative_precision=1e-12)
    want = oracle.pgs_solve(G, g, mu, x0=x0, over_relax=1.0, **kw)
    s = pa.PGSContactSolver(gpu_ctx, 3 * nc)
    s.setMaxIterations(kw["max_it"])
    s.setAbsolutePrecision(kw["absolute_precision"])
    s.setRelativePrecision(kw["relative_precision"])
    x = dev(x0)
    conv = s.solve(dev(G), dev(g), dev(mu), x).cpu().numpy()
    xs = x.cpu().numpy()
    err = np.abs(xs - want["x"]).max(axis=1) / np.maximum(np.abs(want["x"]).max(axis=1), 1e-300)
    it = s.getIterationCount().cpu().numpy()
    same = (it == want["it"]).mean()
    print(f"PGS nc={nc}: max rel err {err.max():.2e}, iterations equal on {same:.4f}, "
          f"converged {conv.mean():.3f} (oracle {want['converged'].mean():.3f})")
    assert err.max() < 1e-9
    assert same >= 0.99
    assert (conv == want["converged"]).mean() >= 0.99
    np.testing.assert_allclose(s.getAbsoluteConvergenceResidual().cpu().numpy()[it == want["it"]],
                               want["absolute_residual"][it == want["it"]], rtol=1e-6, atol=1e-12)


def test_pgs_over_relaxation_and_batch_layout(gpu_ctx):
    """over_relax 1.3 through the C ABI with batch-fastest arrays and host pointers"""
    nc, n = 3, 9
    G, g, mu = problems(nc, 120)
    want = oracle.pgs_solve(G, g, mu, over_relax=1.3, max_it=300, absolute_precision=1e-10, relative_precision=1e-12)
    L = pa.lib()
    Gb = np.ascontiguousarray(G.transpose(0, 2, 1).reshape(B, n * n).T)  # (n^2, B), element (i, j) at i + j n
    gb, mb = np.ascontiguousarray(g.T), np.ascontiguousarray(mu.T)
    xb = np.zeros((n, B))
    st = np.zeros((pa.RBD_NCP_STATS, B))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    settings = pa._PgsSettingsC(300, 1e-10, 1e-12, 1.3)
    rc = L.rbd_pgs_solve(gpu_ctx.handle, ctypes.c_int64(B), nc, p(Gb), ctypes.c_int64(B), p(gb), ctypes.c_int64(B),
                         p(mb), ctypes.c_int64(B), p(xb), ctypes.c_int64(B), ctypes.byref(settings), p(st),
                         ctypes.c_int64(B), ctypes.c_uint32(pa.RBD_HOST_PTRS | pa.RBD_LAYOUT_BATCH))
    assert rc == 0, L.rbd_last_error()
    err = np.abs(xb.T - want["x"]).max() / np.abs(want["x"]).max()
    assert err < 1e-9
    assert (st[0] == want["it"]).mean() >= 0.99
    bad = pa._PgsSettingsC(300, 1e-10, 1e-12, 2.0)
    rc = L.rbd_pgs_solve(gpu_ctx.handle, ctypes.c_int64(B), nc, p(Gb), ctypes.c_int64(B), p(gb), ctypes.c_int64(B),
                         p(mb), ctypes.c_int64(B), p(xb), ctypes.c_int64(B), ctypes.byref(bad), p(st),
                         ctypes.c_int64(B), ctypes.c_uint32(pa.RBD_HOST_PTRS | pa.RBD_LAYOUT_BATCH))
    assert rc != 0 and b"over_relax" in L.rbd_last_error()


# ---------------------------------------------------------------------------------------
# ADMM
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("coupled", [False, True])
@pytest.mark.parametrize("nc", [1, 2, 4])
def test_admm_matches_oracle(gpu_ctx, nc, coupled):
    n = 3 * nc
    G, g, mu = problems(nc, 200 + nc + 10 * coupled, coupled)
    R = np.full((B, n), 1e-6)
    R[::5] = 1e-3
    kw = dict(max_it=2000, absolute_precision=1e-10, relative_precision=1e-12)
    want = oracle.admm_solve(G, g, mu, R=R, **kw)
    s = pa.ADMMContactSolver(gpu_ctx, n)
    s.setMaxIterations(kw["max_it"])
    s.setAbsolutePrecision(kw["absolute_precision"])
    s.setRelativePrecision(kw["relative_precision"])
    conv = s.solve(dev(G), dev(g), dev(mu), R=dev(R)).cpu().numpy()
    y, z = s.getPrimalSolution().cpu().numpy(), s.getDualSolution().cpu().numpy()
    ok = conv & want["converged"]
    scale = np.maximum(np.abs(want["y"]).max(axis=1), 1e-3)
    ey = (np.abs(y - want["y"]).max(axis=1) / scale)[ok]
    ez = (np.abs(z - want["z"]).max(axis=1) / np.maximum(np.abs(want["z"]).max(axis=1), 1e-3))[ok]
    it = s.getIterationCount().cpu().numpy()
    same = (it == want["it"]).mean()
    L_err = np.abs(s.getLargestEigenValue().cpu().numpy() - want["largest_eigen_value"]).max()
    print(f"ADMM nc={nc} coupled={coupled}: y {ey.max():.2e} z {ez.max():.2e}, iterations equal on {same:.4f}, "
          f"converged {conv.mean():.3f} (oracle {want['converged'].mean():.3f}), eigenvalue {L_err:.1e}")
    assert ok.mean() >= 0.95
    assert ey.max() < 1e-7 and ez.max() < 1e-7
    assert same >= 0.95
    assert L_err < 1e-10 * np.abs(want["largest_eigen_value"]).max()


def test_admm_guesses_and_errors(gpu_ctx):
    nc, n = 2, 6
    G, g, mu = problems(nc, 230)
    pg = np.random.default_rng(1).normal(size=(B, n)) * 0.1
    dg = np.abs(np.random.default_rng(2).normal(size=(B, n)))
    kw = dict(max_it=1500, absolute_precision=1e-10, relative_precision=1e-12)
    want = oracle.admm_solve(G, g, mu, primal_guess=pg, dual_guess=dg, **kw)
    s = pa.ADMMContactSolver(gpu_ctx, n)
    s.setMaxIterations(kw["max_it"])
    s.setAbsolutePrecision(kw["absolute_precision"])
    conv = s.solve(dev(G), dev(g), dev(mu), primal_guess=dev(pg), dual_guess=dev(dg)).cpu().numpy()
    ok = conv & want["converged"]
    assert ok.mean() >= 0.95
    assert np.abs(s.getPrimalSolution().cpu().numpy() - want["y"])[ok].max() < 1e-7
    with pytest.raises(ValueError):
        pa.ADMMContactSolver(gpu_ctx, n, tau=1.5).solve(dev(G), dev(g), dev(mu))
    with pytest.raises(ValueError):
        pa.ADMMContactSolver(gpu_ctx, 15)


# ---------------------------------------------------------------------------------------
# a whole contact step on the GPU
# ---------------------------------------------------------------------------------------
def test_contact_step_meets_coulomb_conditions(gpu_ctx):
    """Four foot points of the humanoid (two per foot, LOCAL frames at the soles): the
    Delassus matrix and the free contact velocity from the engine, ADMM on the GPU, and the
    Coulomb conditions checked on the result."""
    desc = M.humanoid36()
    model = pa.Model(gpu_ctx, desc)
    eng = pa.Engine(model, "col")
    Bc = 257
    q, v, _, tau = inputs(desc, Bc, k=95)
    corners = [np.concatenate([np.eye(3).ravel(), [x, 0.0, -0.08]]) for x in (0.06, -0.04)]
    cons = [dict(type=0, joint1=j, reference_frame=1, placement1=c) for j in (7, 13) for c in corners]
    Q, V, T = (torch.tensor(np.ascontiguousarray(x.T), dtype=torch.float64, device="cuda") for x in (q, v, tau))
    G = eng.delassus(Q, to_models(cons))  # (B, 144), element (i, j) at i + 12 j
    # free contact velocity after one step of the unconstrained dynamics: g = J (v + dt ddq_free)
    dt = 1e-3
    ddq = eng.aba(Q, V, T)
    Gnp = G.cpu().numpy().reshape(Bc, 12, 12).transpose(0, 2, 1)
    # the contact points' velocities J v_free (test input, formed on the CPU)
    vf = v + dt * ddq.cpu().numpy().T
    Jv = np.stack([oracle_contact_velocity(desc, q[:, b], vf[:, b], cons) for b in range(Bc)])
    g = Jv
    mu = np.full((Bc, 4), 0.7)
    s = pa.ADMMContactSolver(gpu_ctx, 12)
    s.setAbsolutePrecision(1e-10)
    s.setMaxIterations(3000)
    conv = s.solve(dev(Gnp), dev(g), dev(mu)).cpu().numpy()
    lam = s.getPrimalSolution().cpu().numpy()
    assert conv.mean() > 0.98
    for b in np.where(conv)[0]:
        sig = Gnp[b] @ lam[b] + g[b]
        for c in range(4):
            l, sg = lam[b, 3 * c:3 * c + 3], sig[3 * c:3 * c + 3].copy()
            sg[2] += 0.7 * np.linalg.norm(sg[:2])
            assert np.linalg.norm(l[:2]) <= 0.7 * l[2] + 1e-8
            assert 0.7 * np.linalg.norm(sg[:2]) <= sg[2] + 1e-7
            assert abs(l @ sg) < 1e-7


def oracle_contact_velocity(desc, q, v, cons):
    """J v of the constraints (the contact points' LOCAL velocities) from the oracle's
    forward kinematics: the point velocity of joint j's frame at offset p."""
    _, ov, _ = oracle.forward_kinematics(desc, q[:, None], v[:, None])
    out = []
    for c in cons:
        j = c["joint1"]
        lv = ov[0, j - 1]  # local spatial velocity (linear, angular) of the joint frame
        p = c["placement1"][9:]
        Rc = c["placement1"][:9].reshape(3, 3)
        vp = lv[:3] + np.cross(lv[3:], p)
        out.append(Rc.T @ vp)
    return np.concatenate(out)
