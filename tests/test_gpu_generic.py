"""GPU parity of the run-time-topology kernels (trees that are not prebuilt) against the
FP64 oracle, through the C ABI.

The reference accepts any tree its caller builds with ModelTpl::addJoint
(multibody/model.hxx:61-170) and batches it with rneaInParallel / abaInParallel
(algorithm/parallel/rnea.hpp:39-83).  Here: random trees that are not in the prebuilt
list -- a compact free-flyer tree, a non-compact one (insertion order is not DFS order),
a fixed-base tree with several revolute roots, a 40-link chain (its per-lane state does
not fit LDS, so the device-memory variants run), a 20-DoF floating torso read from MJCF
(bodies with several hinges, flattened into chains through massless links) -- and the
36-DoF humanoid forced onto the run-time-topology kernels, compared with its prebuilt
kernels.

Bars (per configuration, inf-norm relative, oracle in FP64 on the rounded inputs):
  FP32: RNEA, CRBA, RNEA derivatives 1e-5; ABA 1e-4; ABA derivatives and Minv 1e-4, or,
        where M^-1 is worse conditioned (the 40-link chain and the MJCF torso, whose
        light hands give cond(M) ~ 3e4), four times the error of the reference algorithm
        run in FP32 (the oracle in float) or, for the ABA derivatives and Minv, four times
        the error of the exact inverse of the FP32 M the pipeline factors, if larger
  FP64: RNEA, ABA, CRBA 1e-12 (ABA: or the FP32 floor above scaled by 2^-29, the ratio
        of the two unit roundoffs: 3.4e-12 on the MJCF torso); RNEA / ABA derivatives,
        Minv 1e-10
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import mjcf
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols, rel_err_mats, two_link_arm, two_link_closed_form

pytestmark = pytest.mark.gpu

TREES = {
    "tree_ff24": lambda: M.random_tree(24, M.MODEL_SEED_BASE + 101, free_flyer=True, compact=True),
    "tree_noncompact30": lambda: M.random_tree(30, M.MODEL_SEED_BASE + 102, free_flyer=True, compact=False),
    "tree_fixed15": lambda: M.random_tree(15, M.MODEL_SEED_BASE + 103, free_flyer=False, compact=True),
    "chain40": lambda: M.serial_chain(40, M.MODEL_SEED_BASE + 104),
    "mjcf_torso20": lambda: mjcf.parse(TORSO_MJCF, root="freeflyer"),
}
COMPACT_TREES = ["tree_ff24", "tree_fixed15", "chain40", "mjcf_torso20"]

# a floating torso with two arms and a head; shoulders, wrists and neck hold several
# hinges in one body (the reference's composite joints), flattened by the MJCF reader
_ARM = """<body name="{s}shoulder" pos="0 {y} 0.4" euler="10 0 {yaw}">
  <joint type="hinge" axis="1 0 0" pos="0 0 0.02" armature="0.05"/><joint type="hinge" axis="0 1 0" pos="0.01 0 0"/>
  <joint type="hinge" axis="0 0 1"/>
  <inertial mass="2.1" pos="0 0 -0.15" diaginertia="0.02 0.021 0.004"/>
  <body name="{s}elbow" pos="0 0 -0.3"><joint type="hinge" axis="0 1 0" armature="0.02"/>
    <inertial mass="1.3" pos="0.01 0 -0.12" quat="0.98 0.1 0.05 0" diaginertia="0.01 0.012 0.002"/>
    <body name="{s}wrist" pos="0 0 -0.25" quat="0.7071 0 0 0.7071">
      <joint type="hinge" axis="0 0 1" pos="0 0 -0.01"/><joint type="hinge" axis="1 0 0" pos="0 0.005 0"/>
      <inertial mass="0.5" pos="0 0 -0.05" fullinertia="0.001 0.0012 0.0006 1e-5 -2e-5 3e-5"/>
      <body name="{s}palm" pos="0 0 -0.08"><inertial mass="0.2" pos="0 0 -0.02" diaginertia="2e-4 2e-4 1e-4"/></body>
    </body></body></body>"""
TORSO_MJCF = ('<mujoco model="torso20"><compiler angle="degree"/><worldbody><body name="torso" pos="0 0 1">'
              '<inertial mass="12" pos="0 0 0.2" diaginertia="0.4 0.35 0.2"/>'
              + _ARM.format(s="l", y=0.2, yaw=15) + _ARM.format(s="r", y=-0.2, yaw=-15)
              + '<body name="head" pos="0 0 0.55"><joint type="hinge" axis="0 0 1"/>'
              '<joint type="hinge" axis="0 1 0" pos="0 0 0.03"/>'
              '<inertial mass="3" pos="0 0 0.1" diaginertia="0.02 0.02 0.015"/></body>'
              '</body></worldbody></mujoco>')

_models = {}


def gpu_model(ctx, name):
    if name not in _models:
        if name == "humanoid36_rt":
            desc = M.humanoid36()
            _models[name] = (desc, pa.Model(ctx, desc, runtime_topology=True))
        else:
            desc = TREES[name]()
            _models[name] = (desc, pa.Model(ctx, desc))
    return _models[name]


def to_dev(x, layout, dtype=torch.float32):
    a = x.T if layout == "col" else x
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda").contiguous()


def to_host(t, layout):
    a = t.double().cpu().numpy()
    return a.T if layout == "col" else a


def fp32_bar(base, oracle_f32_err):
    """FP32 bar: `base`, or four times the error of the oracle's own FP32 run of the
    reference algorithm on the same inputs when that is larger (conditioning, not the
    kernel, sets the floor)."""
    return max(base, 4.0 * float(np.max(oracle_f32_err)))


def test_trees_are_not_prebuilt(gpu_ctx):
    """None of the test trees has prebuilt kernels; the humanoid has, unless forced."""
    for name in TREES:
        desc, model = gpu_model(gpu_ctx, name)
        assert not model.prebuilt, name
        assert model.compact == (name in COMPACT_TREES), name
    assert pa.Model(gpu_ctx, M.humanoid36()).prebuilt
    assert not gpu_model(gpu_ctx, "humanoid36_rt")[1].prebuilt


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", list(TREES) + ["humanoid36_rt"])
def test_rnea_aba(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 777
    q, v, a, tau = inputs(desc, B, k=40)
    qf, vf, af, tf = f32_round(q, v, a, tau)
    r_f = oracle.aba(desc, qf, vf, tf)
    amp = float(rel_err_cols(oracle.aba(desc, qf, vf, tf, scalar="f32"), r_f).max())  # FP32 error of the algorithm
    for dtype, bar_r, bar_a in ((torch.float32, 1e-5, 1e-4), (torch.float64, 1e-12, 1e-12)):
        qq, vv, aa, tt = (qf, vf, af, tf) if dtype == torch.float32 else (q, v, a, tau)
        Q, V, A, T = (to_dev(x, layout, dtype) for x in (qq, vv, aa, tt))
        e_r = rel_err_cols(to_host(eng.rnea(Q, V, A), layout), oracle.rnea(desc, qq, vv, aa))
        r_a = r_f if dtype == torch.float32 else oracle.aba(desc, qq, vv, tt)
        e_a = rel_err_cols(to_host(eng.aba(Q, V, T), layout), r_a)
        # conditioning floor: four times the algorithm's own FP32 error, scaled to the
        # unit roundoff of the precision under test
        bar_a = max(bar_a, 4.0 * amp * (1.0 if dtype == torch.float32 else 2.0 ** -29))
        print(f"{name} {layout} {dtype}: rnea {e_r.max():.3e} aba {e_a.max():.3e} (bar {bar_a:.1e})")
        assert e_r.max() < bar_r
        assert e_a.max() < bar_a


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", COMPACT_TREES + ["humanoid36_rt"])
def test_crba_cholesky_minverse(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 333
    q, _, _, tau = inputs(desc, B, k=41)
    qf = f32_round(q)[0]
    Mref = oracle.crba(desc, q)
    for dtype, qq, bar in ((torch.float32, qf, 1e-5), (torch.float64, q, 1e-12)):
        Mg = pa.Engine.mats_to_numpy(eng.crba(to_dev(qq, layout, dtype)), layout, desc.nv)
        e = rel_err_mats(Mg, oracle.crba(desc, qq))
        print(f"{name} {layout} {dtype}: crba {e.max():.3e}")
        assert e.max() < bar
        assert np.array_equal(Mg, Mg.transpose(0, 2, 1))
    # FP64 factor and solve of the FP64 M
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
    x = to_dev(tau, layout, torch.float64)
    eng.cholesky_solve(U, D, x)
    np.testing.assert_allclose(to_host(x, layout), x_ref, rtol=1e-9, atol=1e-10)
    # computeMinverse
    _, _, rmi, _ = oracle.aba_derivatives(desc, q, np.zeros((desc.nv, B)), tau)
    Mi64 = pa.Engine.mats_to_numpy(eng.compute_minverse(to_dev(q, layout, torch.float64)), layout, desc.nv)
    assert rel_err_mats(Mi64, rmi).max() < 1e-10
    _, _, rmi32, _ = oracle.aba_derivatives(desc, qf, np.zeros((desc.nv, B)), tau)
    Mi32 = pa.Engine.mats_to_numpy(eng.compute_minverse(to_dev(qf, layout)), layout, desc.nv)
    _, _, omi32, _ = oracle.aba_derivatives(desc, qf, np.zeros((desc.nv, B)), tau, scalar="f32")
    assert rel_err_mats(Mi32, rmi32).max() < fp32_bar(1e-4, rel_err_mats(omi32, rmi32))


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", COMPACT_TREES + ["humanoid36_rt"])
def test_derivatives(gpu_ctx, name, layout):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, layout)
    B = 200
    q, v, a, tau = inputs(desc, B, k=42)
    for dtype, bars in ((torch.float32, (1e-5, 1e-4)), (torch.float64, (1e-10, 1e-10))):
        qq, vv, aa, tt = f32_round(q, v, a, tau) if dtype == torch.float32 else (q, v, a, tau)
        Q, V, A, T = (to_dev(x, layout, dtype) for x in (qq, vv, aa, tt))
        outs = [pa.Engine.mats_to_numpy(o, layout, desc.nv) for o in eng.rnea_derivatives(Q, V, A)]
        rdq, rdv, _, _ = oracle.rnea_derivatives(desc, qq, vv, aa)
        rM = oracle.crba(desc, qq)
        for n, x, r in zip(("dtau_dq", "dtau_dv", "M"), outs, (rdq, rdv, rM)):
            e = rel_err_mats(x, r)
            print(f"{name} {layout} {dtype}: {n} {e.max():.3e}")
            assert e.max() < bars[0], n
            assert np.all(x[:, np.all(r == 0, axis=0)] == 0), n  # structural zeros
        Xq, Xv, Mi, ddq = eng.aba_derivatives(Q, V, T)
        rq, rv, rmi, rddq = oracle.aba_derivatives(desc, qq, vv, tt)
        # In FP32 an ill-conditioned tree (the 40-link chain's M^-1) sets its own floor: the
        # reference's algorithm run in FP32 (the oracle in float), and the exact inverse of
        # the FP32 world-frame M this pipeline factors (the M output above).  A backward-
        # stable FP32 factorisation perturbs M about as much as its FP32 rounding did, so
        # the pipeline's M^-1 may sit a few times above that floor.
        o32 = oracle.aba_derivatives(desc, qq, vv, tt, scalar="f32") if dtype == torch.float32 else None
        floor = rel_err_mats(np.linalg.inv(outs[2]), rmi).max() if o32 is not None else 0.0
        for k, (n, o, r) in enumerate((("ddq_dq", Xq, rq), ("ddq_dv", Xv, rv), ("Minv", Mi, rmi))):
            e = rel_err_mats(pa.Engine.mats_to_numpy(o, layout, desc.nv), r)
            bar = bars[1] if o32 is None else max(fp32_bar(bars[1], rel_err_mats(o32[k], r)), 4.0 * floor)
            print(f"{name} {layout} {dtype}: {n} {e.max():.3e} (bar {bar:.1e})")
            assert e.max() < bar, n
        bar = bars[1] if o32 is None else fp32_bar(bars[1], rel_err_cols(o32[3], rddq))
        assert rel_err_cols(to_host(ddq, layout), rddq).max() < bar


@pytest.mark.parametrize("name", list(TREES) + ["humanoid36_rt"])
def test_kinematics(gpu_ctx, name):
    desc, model = gpu_model(gpu_ctx, name)
    eng = pa.Engine(model, "batch")
    B = 129
    q, v, a, _ = inputs(desc, B, k=43)
    Q, V, A = (to_dev(x, "batch", torch.float64) for x in (q, v, a))
    oM, vo, ao = eng.forward_kinematics(Q, V, A)
    J = eng.joint_jacobians(Q)
    rM, rv, ra = oracle.forward_kinematics(desc, q, v, a)
    rJ = oracle.joint_jacobians(desc, q)
    nj = desc.njoints
    assert np.abs(oM.cpu().numpy().T.reshape(B, nj - 1, 12) - rM).max() < 1e-12 * max(1.0, np.abs(rM).max())
    gJ = J.cpu().numpy().T.reshape(B, desc.nv, 6).transpose(0, 2, 1)
    assert np.abs(gJ - rJ).max() < 1e-12 * max(1.0, np.abs(rJ).max())
    for g, r in ((vo, rv), (ao, ra)):
        g = g.cpu().numpy().T.reshape(B, nj - 1, 6)
        scale = np.abs(r).reshape(B, -1).max(1)[:, None, None]
        assert (np.abs(g - r) / np.maximum(scale, 1e-30)).max() < 1e-12


def test_noncompact_rejections(gpu_ctx):
    """CRBA and the derivatives need a DFS-compact tree, as in the reference
    (crba.hxx:573-595); RNEA and ABA do not."""
    desc, model = gpu_model(gpu_ctx, "tree_noncompact30")
    eng = pa.Engine(model, "col")
    q = torch.zeros((4, desc.nq), device="cuda")
    with pytest.raises(NotImplementedError, match="compact"):
        eng.crba(q)
    x = torch.zeros((4, desc.nv), device="cuda")
    with pytest.raises(NotImplementedError, match="compact"):
        eng.rnea_derivatives(q, x, x)


def test_runtime_matches_prebuilt(gpu_ctx):
    """The humanoid through both kernel families: same functions to rounding."""
    desc = M.humanoid36()
    pre = pa.Engine(pa.Model(gpu_ctx, desc), "batch")
    rt = pa.Engine(gpu_model(gpu_ctx, "humanoid36_rt")[1], "batch")
    q, v, a, tau = f32_round(*inputs(desc, 4096, k=44))
    Q, V, A, T = (to_dev(x, "batch") for x in (q, v, a, tau))
    for f, args, bar in ((lambda e, *x: e.rnea(*x), (Q, V, A), 1e-5), (lambda e, *x: e.aba(*x), (Q, V, T), 1e-4)):
        x1, x2 = to_host(f(pre, *args), "batch"), to_host(f(rt, *args), "batch")
        assert rel_err_cols(x1, x2).max() < bar
    m1 = pa.Engine.mats_to_numpy(pre.crba(Q), "batch", 36)
    m2 = pa.Engine.mats_to_numpy(rt.crba(Q), "batch", 36)
    assert rel_err_mats(m1, m2).max() < 1e-5


def test_mixed_batch_with_runtime_topology(gpu_ctx):
    """A tree without prebuilt kernels inside a mixed-robot batch: per-model launches in
    a captured graph, replayed after the inputs change in place."""
    d1, m1 = gpu_model(gpu_ctx, "tree_noncompact30")
    d2 = M.quadruped18()
    m2 = pa.Model(gpu_ctx, d2)
    groups, host = [], []
    for d, m, n, k in ((d1, m1, 1500, 45), (d2, m2, 700, 46)):
        q, v, a, tau = f32_round(*inputs(d, n, k=k))
        host.append((d, q, v, a, tau))
        groups.append((m, *(to_dev(x, "batch") for x in (q, v, a, tau))))
    mb = pa.MixedBatch(gpu_ctx, groups)
    for rep in range(3):
        if rep == 2:  # inputs rewritten in place: the replayed graph must see them
            for k in range(len(groups)):
                groups[k][4].mul_(0.5)
                d, q, v, a, tau = host[k]
                host[k] = (d, q, v, a, tau * 0.5)
        outs = mb.run()
        torch.cuda.synchronize()
        for (d, q, v, a, tau), (tau_o, ddq_o) in zip(host, outs):
            assert rel_err_cols(tau_o.double().cpu().numpy(), oracle.rnea(d, q, v, a)).max() < 1e-5
            assert rel_err_cols(ddq_o.double().cpu().numpy(), oracle.aba(d, q, v, tau)).max() < 1e-4


def test_two_link_closed_form_on_gpu(gpu_ctx):
    """A two-link arm with a rotated, offset placement (not prebuilt): RNEA in FP64 and FP32
    and the tip placement against the closed-form planar equations of motion."""
    d = two_link_arm()
    model = pa.Model(gpu_ctx, d)
    assert not model.prebuilt
    eng = pa.Engine(model, "batch")
    B = 300
    rng = np.random.default_rng(5)
    q, v, a = rng.uniform(-3, 3, (3, 2, B))
    tau_ref, pos_ref = two_link_closed_form(q, v, a)
    tau64 = eng.rnea(*(to_dev(x, "batch", torch.float64) for x in (q, v, a))).cpu().numpy()
    np.testing.assert_allclose(tau64, tau_ref, rtol=1e-12, atol=1e-12)
    qf, vf, af = f32_round(q, v, a)
    tau_f, _ = two_link_closed_form(qf, vf, af)
    assert rel_err_cols(to_host(eng.rnea(*(to_dev(x, "batch") for x in (qf, vf, af))), "batch"), tau_f).max() < 1e-5
    ddq = eng.aba(*(to_dev(x, "batch", torch.float64) for x in (q, v, tau_ref))).cpu().numpy()
    np.testing.assert_allclose(ddq, a, rtol=1e-10, atol=1e-10)
    oM, _, _ = eng.forward_kinematics(to_dev(q, "batch", torch.float64))
    o = oM.cpu().numpy().T.reshape(B, 2, 12)
    pos = np.stack([o[b, 1, 9:] + o[b, 1, :9].reshape(3, 3) @ np.array([0.25, 0, 0]) for b in range(B)], axis=1)
    np.testing.assert_allclose(pos, pos_ref, rtol=1e-12, atol=1e-12)
