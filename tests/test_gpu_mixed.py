"""GPU parity of the mixed-robot batch (rbd_multi_rnea_aba): several models' instances
in one launch over a cost-ordered tile table, autotuned and replayed from a HIP graph.

Every group's RNEA tau and ABA ddq are checked against the FP64 oracle with the
single-model bars (FP32 RNEA 1e-5, FP32 ABA 1e-4 relative per instance, inf-norm, the
oracle fed the FP32-rounded inputs) and against the single-model kernels.  Cases: the
four-robot mix (7-DoF arm, 18-DoF quadruped, 36-DoF humanoid, 60-DoF random tree) with
ragged group sizes, graph replays after the inputs are rewritten in place (checked
against eager launches and the oracle on every replay), an empty group, a set of models
that no fused kernel covers (per-model launches), and the autotuner's report.
"""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols

pytestmark = pytest.mark.gpu

MIX = [("arm7", 3001), ("quadruped18", 2999), ("humanoid36", 1000), ("random60", 517)]


@pytest.fixture(scope="module")
def mix_models(gpu_ctx):
    out = {}
    for name, _ in MIX + [("manipulator6", 0), ("pendulum", 0)]:
        desc = M.SAMPLE_MODELS[name]() if name in M.SAMPLE_MODELS else getattr(M, name)()
        out[name] = (desc, pa.Model(gpu_ctx, desc))
    return out


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float32, device="cuda")


def make_groups(models, spec, k0=0):
    host, groups = [], []
    for j, (name, n) in enumerate(spec):
        desc, model = models[name]
        q, v, a, tau = f32_round(*inputs(desc, n, k=k0 + j))
        host.append((desc, q, v, a, tau))
        groups.append((model, dev(q), dev(v), dev(a), dev(tau)))
    return host, groups


def check_against_oracle(host, outs, label):
    for (desc, q, v, a, tau), (tau_g, ddq_g) in zip(host, outs):
        if q.shape[1] == 0:
            continue
        e_r = rel_err_cols(tau_g.double().cpu().numpy(), oracle.rnea(desc, q, v, a)).max()
        e_a = rel_err_cols(ddq_g.double().cpu().numpy(), oracle.aba(desc, q, v, tau)).max()
        assert e_r < 1e-5, f"{label} {desc.name}: RNEA rel err {e_r:.3e}"
        assert e_a < 1e-4, f"{label} {desc.name}: ABA rel err {e_a:.3e}"


def test_mix_matches_oracle_and_single_model(gpu_ctx, mix_models):
    host, groups = make_groups(mix_models, MIX)
    mb = pa.MixedBatch(gpu_ctx, groups)
    outs = [(t.clone(), d.clone()) for t, d in mb.run()]
    check_against_oracle(host, outs, "fused")
    for (model, Q, V, A, T), (tau_g, ddq_g) in zip(groups, outs):
        eng = pa.Engine(model, "batch")
        tr, dr = eng.rnea(Q, V, A), eng.aba(Q, V, T)
        assert rel_err_cols(tau_g.double().cpu().numpy(), tr.double().cpu().numpy()).max() < 1e-6
        assert rel_err_cols(ddq_g.double().cpu().numpy(), dr.double().cpu().numpy()).max() < 1e-6
    rep = mb.tuning()
    assert rep["mix"] == "bench4", rep
    assert rep["chosen"] and len(rep["candidates"]) >= 2, rep


def test_mix_graph_replay_tracks_inputs(gpu_ctx, mix_models):
    host, groups = make_groups(mix_models, MIX, k0=10)
    mb = pa.MixedBatch(gpu_ctx, groups)
    eager = pa.MixedBatch(gpu_ctx, groups, graph=False, autotune=False)
    for rep in range(3):
        if rep:
            # new inputs written into the same device buffers the graph was captured on
            new_host, _ = make_groups(mix_models, MIX, k0=20 + 10 * rep)
            for (m, Q, V, A, T), (desc, q, v, a, tau) in zip(groups, new_host):
                for D, h in ((Q, q), (V, v), (A, a), (T, tau)):
                    D.copy_(dev(h))
            host = new_host
        outs = [(t.clone(), d.clone()) for t, d in mb.run()]
        ref = [(t.clone(), d.clone()) for t, d in eager.run()]
        check_against_oracle(host, outs, f"replay {rep}")
        for (tg, dg), (te, de) in zip(outs, ref):
            assert torch.equal(tg, te) or rel_err_cols(tg.double().cpu().numpy(), te.double().cpu().numpy()).max() < 1e-6
            assert torch.equal(dg, de) or rel_err_cols(dg.double().cpu().numpy(), de.double().cpu().numpy()).max() < 1e-6
        torch.cuda.synchronize()


def test_mix_empty_group(gpu_ctx, mix_models):
    spec = [("arm7", 130), ("quadruped18", 0), ("humanoid36", 65), ("random60", 1)]
    host, groups = make_groups(mix_models, spec, k0=40)
    mb = pa.MixedBatch(gpu_ctx, groups)
    outs = mb.run()
    check_against_oracle(host, outs, "ragged")


def test_uncovered_models_use_per_model_launches(gpu_ctx, mix_models):
    spec = [("manipulator6", 200), ("pendulum", 77)]
    host, groups = make_groups(mix_models, spec, k0=50)
    mb = pa.MixedBatch(gpu_ctx, groups)
    outs = mb.run()
    check_against_oracle(host, outs, "per-model")
    rep = mb.tuning()
    assert rep["mix"] == "none" and rep["chosen"] in ("per-model launches", "concurrent per-model launches"), rep


def test_mix_argument_errors(gpu_ctx, mix_models):
    desc, model = mix_models["arm7"]
    q = torch.zeros((desc.nq, 10), dtype=torch.float32, device="cuda")
    v = torch.zeros((desc.nv, 10), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        pa.MixedBatch(gpu_ctx, [(model, q, v, v, v[:, :5])])
    with pytest.raises(ValueError):
        pa.MixedBatch(gpu_ctx, [(model, q.double(), v, v, v)])


def test_benchmark_mix_at_full_size(gpu_ctx, mix_models):
    """The benchmark's exact mix -- arm7 x 40000, quadruped18 x 40000, humanoid36 x 15000,
    random60 x 5000, seeded as in the benchmark -- through the autotuned, graph-replayed
    plan the tuner picks at that size; per group the first and the last 2048 instances
    (so each group's last, partial chunk) against the oracle."""
    spec = [("arm7", 40000), ("quadruped18", 40000), ("humanoid36", 15000), ("random60", 5000)]
    host, groups = [], []
    for j, (name, n) in enumerate(spec):
        desc, model = mix_models[name]
        q, v, a, tau = f32_round(*M.sample_inputs(desc, n, M.INPUT_SEED_BASE + 1 + j))
        host.append((desc, q, v, a, tau))
        groups.append((model, dev(q), dev(v), dev(a), dev(tau)))
    mb = pa.MixedBatch(gpu_ctx, groups)
    mb.run()  # tunes and captures
    outs = mb.run()  # graph replay of the chosen plan
    torch.cuda.synchronize()
    print("benchmark mix:", mb.tuning().get("chosen"))
    for (desc, q, v, a, tau), (tau_g, ddq_g) in zip(host, outs):
        n = q.shape[1]
        idx = np.r_[0:2048, n - 2048:n]
        sel = lambda x: np.ascontiguousarray(x[:, idx])
        e_r = rel_err_cols(tau_g.double().cpu().numpy()[:, idx], oracle.rnea(desc, sel(q), sel(v), sel(a))).max()
        e_a = rel_err_cols(ddq_g.double().cpu().numpy()[:, idx], oracle.aba(desc, sel(q), sel(v), sel(tau))).max()
        print(f"  {desc.name} x {n}: rnea {e_r:.3e} aba {e_a:.3e}")
        assert e_r < 1e-5 and e_a < 1e-4, desc.name
