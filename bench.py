"""Benchmark: batched rigid-body dynamics on one MI355X.

Headline workload (`value`): one step = ABA (ddq, FP32) + CRBA (dense M, FP32) + FP64
tree-Cholesky factorisation of M (cholesky::decompose) for 65536 configurations of the
36-DoF humanoid (free-flyer + 2x6R legs + 2R torso + 2x7R arms + 2R head), inputs already
resident in HBM in the engine's native batch-fastest layout.  value = configurations per
second.  Alongside, on the same GPU: RNEA, ABA, CRBA, Cholesky alone, computeRNEADerivatives
and computeABADerivatives at 16384 configurations, FP64 RNEA on the 6R manipulator, and the
max relative error of every output against the FP64 CPU oracle (per configuration,
inf-norm, oracle run on the FP32-rounded inputs).

`roofline` is reported for the step's dominant kernel, timed with HIP events on the
engine's own stream around that kernel in every timed step (an untimed pass with events
around all three kernels names it and gives the others' durations; events between every
kernel would add ~15 us of stream time per step); `cpu_baseline` times the oracle (the FP64 C++ restatement of the
reference's rneaInParallel / abaInParallel batch path, OpenMP over configurations) on the
host cores.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--only aba|crba|chol|step]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
METRIC = "configs/s for ABA, RNEA, ∂RNEA/∂ABA on 36-DoF humanoid; max rel err vs FP64"


def rel_cols(x, ref):
    return float((np.abs(x - ref).max(0) / np.maximum(np.abs(ref).max(0), 1e-30)).max())


def rel_mats(x, ref):
    d = np.abs(x - ref).reshape(len(ref), -1).max(1)
    return float((d / np.maximum(np.abs(ref).reshape(len(ref), -1).max(1), 1e-30)).max())


def err_cols(x, ref):
    """per-configuration inf-norm relative errors of (n, B) arrays"""
    return np.abs(x - ref).max(0) / np.maximum(np.abs(ref).max(0), 1e-30)


def err_mats(x, ref):
    """per-configuration errors of (B, n, n) arrays"""
    d = np.abs(x - ref).reshape(len(ref), -1).max(1)
    return d / np.maximum(np.abs(ref).reshape(len(ref), -1).max(1), 1e-30)


def stats(e):
    """max, median and 99.9th percentile of per-configuration errors, over the whole batch"""
    return {"max": float(e.max()), "p50": float(np.median(e)), "p99.9": float(np.quantile(e, 0.999)),
            "configs": int(e.size)}


def host_info():
    """CPU model, sockets, cores, SMT of the host (lscpu fields), for the CPU baseline"""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)", "NUMA node(s)"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


MIX_SPEC = [("arm7", 40000), ("quadruped18", 40000), ("humanoid36", 15000), ("random60", 5000)]


def mixed_batch(torch, pa, M, ctx, dev, autotune=True):
    """The 100,000-instance four-robot mix (RNEA + ABA per instance, FP32, seeded inputs)."""
    groups = []
    for j, (name, n) in enumerate(MIX_SPEC):
        dsc = M.SAMPLE_MODELS[name]()
        mdl = pa.Model(ctx, dsc)
        xs = M.sample_inputs(dsc, n, M.INPUT_SEED_BASE + 1 + j)
        groups.append((mdl, *(torch.tensor(x.astype(np.float32), device=dev) for x in xs)))
    return pa.MixedBatch(ctx, groups, autotune=autotune), sum(n for _, n in MIX_SPEC)


class Timer:
    """HIP-event timing on the engine's stream (torch events on an external stream)."""

    def __init__(self, torch, stream_ptr):
        self.torch = torch
        self.stream = torch.cuda.ExternalStream(stream_ptr)
        self.acc = {}

    def run(self, name, fn):
        t = self.torch
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        s.record(self.stream)
        out = fn()
        e.record(self.stream)
        self.acc.setdefault(name, []).append((s, e))
        return out

    def mean_ms(self, name):
        self.torch.cuda.synchronize()
        ev = self.acc.get(name, [])
        return float(np.mean([s.elapsed_time(e) for s, e in ev])) if ev else float("nan")

    def count(self, name):
        return len(self.acc.get(name, []))

    def reset(self):
        self.acc = {}


def surrogate_bench(torch, pa, M, ctx, args):
    """Learned forward dynamics on the 18-DoF quadruped: train the 73-512-512-512-18 SiLU
    network on GPU-ABA samples (q as the headline inputs, v ~ N(0, 1), tau ~ N(0, 20^2)),
    then run surrogate + RNEA-residual gate (1e-3) + ABA fallback on 262144 held-out
    configurations next to plain ABA on the same inputs."""
    from pinocchio_amd import surrogate as S
    desc = M.quadruped18()
    model = pa.Model(ctx, desc)
    t0 = time.perf_counter()
    data = S.make_dataset(model, args.sur_samples, M.INPUT_SEED_BASE + 1000)
    t_data = time.perf_counter() - t0
    sur = S.Surrogate(model)
    sur.fit_normalization(*data)
    losses, t_train = S.train(sur, data, epochs=args.sur_epochs, batch=4096, lr=5e-4)
    del data
    n = 262144
    Q, V, T = S.sample_inputs_device(model, n, M.INPUT_SEED_BASE + 2000, tau_sigma=20.0)
    eng = pa.Engine(model, "batch")
    out = torch.empty((desc.nv, n), dtype=torch.float32, device=Q.device)

    def tm(f, reps=10):
        f()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            f()
        ctx.synchronize()
        return (time.perf_counter() - t) / reps

    t_aba = tm(lambda: eng.aba(Q, V, T, out=out, async_=True))
    t_pred = tm(lambda: sur.predict(Q, V, T, out=out, async_=True))
    t_pred_layered = tm(lambda: sur.predict(Q, V, T, out=out, async_=True, layered=True))
    t_fd = tm(lambda: sur.forward_dynamics(Q, V, T, tol=1e-3, out=out))
    _, nfb, acc, res = sur.forward_dynamics(Q, V, T, tol=1e-3, out=out, want_mask=True)
    r = res.cpu().numpy()
    fwd_flop = 2 * sum(o * i for o, i in sur.layer_shapes())
    return {
        "train_samples": args.sur_samples, "epochs": args.sur_epochs, "batch": 4096, "train_s": t_train,
        "data_s": t_data, "train_loss_per_epoch": [float(x) for x in losses],
        "infer_batch": n, "tol": 1e-3, "fallbacks": int(nfb), "accepted_frac": float(1 - nfb / n),
        "residual_p50": float(np.median(r)), "residual_p90": float(np.quantile(r, 0.9)),
        "accepted_frac_at": {str(t): float((r <= t).mean()) for t in (1e-3, 1e-2, 1e-1)},
        "aba_configs_per_s": n / t_aba, "surrogate_with_fallback_configs_per_s": n / t_fd,
        "surrogate_only_configs_per_s": n / t_pred, "speedup_vs_aba": t_aba / t_fd,
        "mlp_forward_TFLOP_s": fwd_flop * n / t_pred / 1e12, "mlp_frac_of_bf16_peak": fwd_flop * n / t_pred / 2.5e15,
        "mlp_inference_ms": {"fused": t_pred * 1e3, "layer_by_layer": t_pred_layered * 1e3},
    }


def contact_bench(torch, pa, ctx, eng, Qc, B, dev, time_call):
    """contact problems: the Delassus matrix of four foot points (two per foot, LOCAL) at
    the near-posture configurations Qc, and batched ADMM / PGS on it (FP64, 12 rows, mu 0.7)"""
    others = {}
    corners = [np.concatenate([np.eye(3).ravel(), [x, 0.0, -0.08]]) for x in (0.06, -0.04)]
    pts = [pa.RigidConstraintModel(pa.CONTACT_3D, j, c, reference_frame=pa.LOCAL) for j in (7, 13) for c in corners]
    Qc64 = Qc.double()
    Gd = torch.empty((144, B), dtype=torch.float64, device=dev)
    t = time_call(lambda: eng.delassus(Qc64, pts, out=Gd, async_=True), reps=10)
    others["delassus_humanoid36_f64"] = {"batch": B, "constraints": "four foot points (12 rows)",
                                         "configs_per_s": B / t, "ms": t * 1e3}
    Gp = Gd.T.contiguous()  # (B, 144): element (i, j) at i + 12 j
    gen = torch.Generator(device=dev).manual_seed(11)
    gv = torch.randn((B, 12), dtype=torch.float64, device=dev, generator=gen) * 0.05
    gv[:, 2::3] -= 0.2  # pressing contacts
    mus = torch.full((B, 4), 0.7, dtype=torch.float64, device=dev)
    admm = pa.ADMMContactSolver(ctx, 12)
    admm.setAbsolutePrecision(1e-8)
    admm.setMaxIterations(500)
    t = time_call(lambda: admm.solve(Gp, gv, mus, async_=True), reps=3)
    its = admm.getIterationCount().float()
    others["admm_contact_humanoid36"] = {"problems": B, "cones": 4, "ms": t * 1e3, "problems_per_s": B / t,
                                         "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
                                         "converged_frac": float(admm.stats[:, 3].mean()),
                                         "absolute_precision": 1e-8}
    # the same solver held to exactly 100 iterations (precisions 0): the cost of an
    # iteration without the spread of iteration counts within a wave
    fixed = pa.ADMMContactSolver(ctx, 12)
    fixed.setAbsolutePrecision(0.0)
    fixed.setRelativePrecision(0.0)
    fixed.setMaxIterations(100)
    t = time_call(lambda: fixed.solve(Gp, gv, mus, async_=True), reps=3)
    others["admm_contact_humanoid36"]["fixed_100_iterations_ms"] = t * 1e3
    others["admm_contact_humanoid36"]["ns_per_problem_iteration_fixed"] = t / (B * 100) * 1e9
    pgs = pa.PGSContactSolver(ctx, 12)
    pgs.setAbsolutePrecision(1e-8)
    pgs.setMaxIterations(500)
    xg = torch.zeros((B, 12), dtype=torch.float64, device=dev)
    t = time_call(lambda: (xg.zero_(), pgs.solve(Gp, gv, mus, xg, async_=True)), reps=3)
    its = pgs.getIterationCount().float()
    others["pgs_contact_humanoid36"] = {"problems": B, "cones": 4, "ms": t * 1e3, "problems_per_s": B / t,
                                        "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
                                        "converged_frac": float(pgs.stats[:, 3].mean()),
                                        "absolute_precision": 1e-8}
    return {"others": others, "Gp": Gp, "g": gv, "mu": mus, "admm": admm, "xg": xg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--deriv-batch", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-samples", type=int, default=1 << 22)
    ap.add_argument("--sur-epochs", type=int, default=8)
    ap.add_argument("--only", default=None,
                    help="one part: step|aba|crba|chol|rnea|rnead|abad|abad64|cfd|contact|mix|mlp|surrogate|copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; the batch shards with no "
                                   "exchange, so multi-GPU runs are out of scope for this build"}))
        sys.exit(2)

    import torch

    import pinocchio_amd as pa
    from pinocchio_amd import models as M

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    ctx = pa.Context(0)
    timer = Timer(torch, ctx.stream)

    if args.only == "copy":
        # plain device copy of a known size: calibrates the FETCH_SIZE / WRITE_SIZE counters
        n = 1 << 28  # 1 GiB of float32
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.warmup + args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"only": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                          "KiB_per_dispatch": 4 * n / 1024}))
        return

    if args.only == "mlp":
        # surrogate inference kernels alone (untrained weights: the kernel work is the same)
        from pinocchio_amd import surrogate as S
        dq = M.quadruped18()
        mq = pa.Model(ctx, dq)
        sur = S.Surrogate(mq)
        n = 262144
        Qm, Vm, Tm = S.sample_inputs_device(mq, n, M.INPUT_SEED_BASE + 2000, tau_sigma=20.0)
        out = torch.empty((dq.nv, n), dtype=torch.float32, device=dev)
        fwd_flop = 2 * sum(o * i for o, i in sur.layer_shapes())
        res = {"only": "mlp", "batch": n, "forward_flop_per_config": fwd_flop}
        tmr = Timer(torch, ctx.stream)
        for name, layered in (("fused", False), ("layer_by_layer", True)):
            for _ in range(args.warmup):
                sur.predict(Qm, Vm, Tm, out=out, async_=True, layered=layered)
            ctx.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                tmr.run(name, lambda: sur.predict(Qm, Vm, Tm, out=out, async_=True, layered=layered))
            ctx.synchronize()
            wall = (time.perf_counter() - t) / args.steps * 1e3
            ms = tmr.mean_ms(name)  # stream time of one call (HIP events)
            res[name] = {"ms": ms, "wall_ms": wall, "TFLOP_s": fwd_flop * n / ms / 1e9,
                         "frac_of_bf16_peak": fwd_flop * n / ms / 2.5e12}
        print(json.dumps(res))
        return

    if args.only == "surrogate":
        print(json.dumps({"only": "surrogate", **surrogate_bench(torch, pa, M, ctx, args)}))
        return

    if args.only == "mix":
        # the mixed-robot batch twice: the autotuned plan and the one-launch plan with
        # default tile widths (for kernel traces and counters of both)
        out = {}
        for tuned in (True, False):
            mb, n_mix = mixed_batch(torch, pa, M, ctx, dev, autotune=tuned)
            mb.run()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                mb.run(async_=True)
            ctx.synchronize()
            t = (time.perf_counter() - t0) / args.steps
            out["tuned" if tuned else "fused_width1"] = {"ms": t * 1e3, "chosen": mb.tuning().get("chosen")}
        print(json.dumps({"only": "mix", **out}))
        return

    desc = M.humanoid36()
    model = pa.Model(ctx, desc)
    eng = pa.Engine(model, "batch")
    B = args.batch
    q, v, a, tau = M.sample_inputs(desc, B, M.INPUT_SEED_BASE + 1)
    q32, v32, a32, tau32 = (x.astype(np.float32) for x in (q, v, a, tau))
    Q, V, A, T = (torch.tensor(x, device=dev).contiguous() for x in (q32, v32, a32, tau32))
    nv = desc.nv
    ddq = torch.empty((nv, B), device=dev)
    Mbuf = torch.empty((nv * nv, B), device=dev)
    rows, cols = model.cholesky_pattern()
    nnz = len(rows)
    Ubuf = torch.empty((nnz, B), dtype=torch.float64, device=dev)
    Dbuf = torch.empty((nv, B), dtype=torch.float64, device=dev)
    # constrained forward dynamics (forwardDynamics with J, gamma): both feet of the humanoid
    # in contact, J = their WORLD Jacobians from the engine's own kernel (support columns),
    # gamma = 0; element (i, k) of a configuration's 12 x nv J at row i + 12 k
    feet = (7, 13)
    ncon = 6 * len(feet)
    J6 = eng.joint_jacobians(Q).view(nv, 6, B)
    Jc = torch.zeros((nv, ncon, B), device=dev)
    for f, j in enumerate(feet):
        sup, i = [], j
        while i > 0:
            sup += list(range(desc.idx_v[i], desc.idx_v[i] + desc.nvs[i]))
            i = int(desc.parents[i])
        Jc[sup, 6 * f:6 * f + 6, :] = J6[sup]
    Jrows = Jc.reshape(nv * ncon, B).contiguous()
    gam0 = torch.zeros((ncon, B), device=dev)
    cddq = torch.empty((nv, B), device=dev)
    clam = torch.empty((ncon, B), device=dev)
    L = pa.lib()
    import ctypes

    def chol():
        pa._check(L.rbd_cholesky_decompose(ctx.handle, model.handle, ctypes.c_int64(B), 0,
                                           ctypes.c_void_p(Mbuf.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_void_p(Ubuf.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_void_p(Dbuf.data_ptr()), ctypes.c_int64(B),
                                           ctypes.c_uint32(pa.RBD_LAYOUT_BATCH | pa.RBD_ASYNC)))

    calls = {"aba": lambda: eng.aba(Q, V, T, out=ddq, async_=True),
             "crba": lambda: eng.crba(Q, out=Mbuf, async_=True),
             "chol": chol}

    def step(timed=()):
        # one step; the kernels named in `timed` are bracketed by HIP events
        for k in ("aba", "crba", "chol"):
            if k in timed:
                timer.run(k, calls[k])
            else:
                calls[k]()

    only = args.only
    # ---------------- headline step ----------------
    cal_ms, dom = {}, None
    if only in (None, "step", "aba", "crba", "chol"):
        if only in (None, "step"):
            for _ in range(args.warmup):
                step()
            ctx.synchronize()
            # Untimed pass with every kernel bracketed by events: per-kernel durations and
            # the dominant kernel.  A timing event between two kernels costs ~5 us of
            # stream time, so the timed steps bracket the dominant kernel only.
            timer.reset()
            for _ in range(args.steps):
                step(("aba", "crba", "chol"))
            cal_ms = {k: timer.mean_ms(k) for k in ("aba", "crba", "chol")}
            dom = max(cal_ms, key=lambda k: cal_ms[k])
            fn = lambda: step((dom,))
        else:
            fn = lambda: (calls["crba"]() if only == "chol" else None, timer.run(only, calls[only]))
            for _ in range(args.warmup):
                fn()
        torch.cuda.synchronize()
        timer.reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    if only is not None:
        extra = {}
        if only == "contact":
            def tc(f, reps=10):
                f()
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    f()
                ctx.synchronize()
                return (time.perf_counter() - t0) / reps
            qc = M.perturb_configurations(desc, q32[:, 0], B, 0.25, M.INPUT_SEED_BASE + 900).astype(np.float32)
            cres = contact_bench(torch, pa, ctx, eng, torch.tensor(qc, device=dev).contiguous(), B, dev, tc)
            print(json.dumps({"only": "contact", **cres["others"]}))
            return
        if only in ("rnea", "rnead", "abad", "abad64", "cfd"):
            Bd = args.deriv_batch
            Qd, Vd, Ad, Td = Q[:, :Bd].contiguous(), V[:, :Bd].contiguous(), A[:, :Bd].contiguous(), T[:, :Bd].contiguous()
            Qd64, Vd64, Td64 = (x.double() for x in (Qd, Vd, Td))
            f = {"rnea": lambda: eng.rnea(Q, V, A, async_=True),
                 "rnead": lambda: eng.rnea_derivatives(Qd, Vd, Ad, async_=True),
                 "abad": lambda: eng.aba_derivatives(Qd, Vd, Td, async_=True),
                 "abad64": lambda: eng.aba_derivatives(Qd64, Vd64, Td64, async_=True),
                 "cfd": lambda: eng.constrained_forward_dynamics(Q, V, T, Jrows, gam0, out=cddq, lambda_out=clam,
                                                                 async_=True)}[only]
            for _ in range(args.warmup):
                f()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                f()
            ctx.synchronize()
            extra["ms"] = (time.perf_counter() - t0) / args.steps * 1e3
        else:
            extra["ms"] = {k: timer.mean_ms(k) for k in ("aba", "crba", "chol") if timer.count(k)}
            if only == "step":
                extra["ms_per_step"] = elapsed / args.steps * 1e3
        print(json.dumps({"only": only, **extra}))
        return

    ms_step = elapsed / args.steps * 1e3
    value = B * args.steps / elapsed
    # the dominant kernel's duration comes from the timed steps, the others' from the
    # calibration pass before them
    k_ms = dict(cal_ms)
    k_ms[dom] = timer.mean_ms(dom)

    # algorithmic work per configuration (bytes: inputs read + outputs written; FLOP:
    # op-counted from the oracle's restatement of the reference algorithm)
    import oracle

    fl = {a: oracle.flops(desc, a) for a in ("rnea", "aba_world", "crba", "cholesky", "rnea_derivatives",
                                             "aba_derivatives")}
    nq = desc.nq
    work = {
        "aba": {"bytes": 4 * (nq + 3 * nv), "flop": fl["aba_world"], "bound": "mfma", "peak": FP32_PEAK_TFLOPS},
        "crba": {"bytes": 4 * (nq + nv * nv), "flop": fl["crba"], "bound": "hbm", "peak": HBM_PEAK_GBS},
        "chol": {"bytes": 4 * (nnz + nv) + 8 * (nnz + nv), "flop": fl["cholesky"], "bound": "hbm",
                 "peak": HBM_PEAK_GBS},
    }
    w = work[dom]
    if w["bound"] == "hbm":
        achieved = w["bytes"] * B / (k_ms[dom] * 1e-3) / 1e9
        unit = "GB/s"
    else:
        achieved = w["flop"] * B / (k_ms[dom] * 1e-3) / 1e12
        unit = "TFLOP/s"
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc
    # summary (FETCH_SIZE and WRITE_SIZE, each scaled by a 1 GiB device copy)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "step_humanoid36_b65536_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        t_ = tj["bytes_per_launch"].get("k_" + dom)
        if t_ and tj["batch"] == B:
            traffic = t_["read"] + t_["write"]
    roofline = {"bound": w["bound"], "achieved": achieved, "peak": w["peak"], "unit": unit,
                "frac": achieved / w["peak"], "traffic": traffic, "kernel": dom,
                "algorithmic_bytes_per_launch": w["bytes"] * B,
                "per_kernel_ms": k_ms,
                "timing": f"{dom}: HIP events around it in the timed steps; other kernels: HIP events in an "
                          f"untimed pass of the same steps",
                "per_kernel": {k: {"ms": k_ms[k], "GB_s": work[k]["bytes"] * B / (k_ms[k] * 1e-3) / 1e9,
                                   "TFLOP_s": work[k]["flop"] * B / (k_ms[k] * 1e-3) / 1e12,
                                   "bytes_per_config": work[k]["bytes"], "flop_per_config": work[k]["flop"]}
                               for k in k_ms}}
    if dom == "aba":
        roofline["note"] = ("ABA is bound by FP32 vector ALU throughput; peak is MI355X's FP32 157.3 TFLOP/s, "
                            "algorithmic FLOP counted on the reference's WORLD-convention operation sequence")

    # ---------------- other algorithms (same GPU, same stream) ----------------
    def time_call(f, reps=10):
        f()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        ctx.synchronize()
        return (time.perf_counter() - t0) / reps

    Bd = args.deriv_batch
    Qd, Vd, Ad, Td = (x[:, :Bd].contiguous() for x in (Q, V, A, T))
    others = {}
    # the headline step recorded once (rbd_capture_begin / _end) and replayed as one HIP
    # graph per step (launch gaps between the three kernels shrink)
    with ctx.capture() as cap:
        step()
    t = time_call(lambda: cap.graph.launch(async_=True), reps=args.steps)
    cap.graph.close()
    others["step_graph_replay"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3,
                                   "note": "ABA + CRBA + FP64 Cholesky captured once, replayed per step"}
    t = time_call(lambda: eng.rnea(Q, V, A, async_=True))
    others["rnea_humanoid36_f32"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3}
    t = time_call(lambda: eng.aba(Q, V, T, out=ddq, async_=True))
    others["aba_humanoid36_f32"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3}
    t = time_call(lambda: eng.rnea_derivatives(Qd, Vd, Ad, async_=True), reps=5)
    others["rnea_derivatives_humanoid36_f32"] = {"batch": Bd, "configs_per_s": Bd / t, "ms": t * 1e3,
                                                 "GB_s": (4 * (nq + 2 * nv) + 3 * 4 * nv * nv) * Bd / t / 1e9}
    t = time_call(lambda: eng.aba_derivatives(Qd, Vd, Td, async_=True), reps=5)
    others["aba_derivatives_humanoid36_f32"] = {"batch": Bd, "configs_per_s": Bd / t, "ms": t * 1e3,
                                                "GB_s": (4 * (nq + 2 * nv) + 3 * 4 * nv * nv) * Bd / t / 1e9}
    man = M.manipulator6()
    mm = pa.Model(ctx, man)
    qm, vm, am, _ = M.sample_inputs(man, 4096, M.INPUT_SEED_BASE + 0)
    em = pa.Engine(mm, "batch")
    Qm, Vm, Am = (torch.tensor(x, dtype=torch.float64, device=dev) for x in (qm, vm, am))
    t = time_call(lambda: em.rnea(Qm, Vm, Am, async_=True), reps=50)
    others["rnea_manipulator6_f64"] = {"batch": 4096, "configs_per_s": 4096 / t, "us": t * 1e6}

    # forward kinematics (placements) and WORLD joint Jacobians, FP32, HBM-bound outputs
    t = time_call(lambda: eng.forward_kinematics(Q, async_=True), reps=10)
    nbytes = 4 * (desc.nq + 12 * (desc.njoints - 1))
    others["forward_kinematics_humanoid36_f32"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3,
                                                   "GB_s": nbytes * B / t / 1e9}
    t = time_call(lambda: eng.joint_jacobians(Q, async_=True), reps=10)
    nbytes = 4 * (desc.nq + 6 * nv)
    others["joint_jacobians_humanoid36_f32"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3,
                                                "GB_s": nbytes * B / t / 1e9}

    t = time_call(lambda: eng.constrained_forward_dynamics(Q, V, T, Jrows, gam0, out=cddq, lambda_out=clam,
                                                           async_=True), reps=10)
    others["constrained_forward_dynamics_humanoid36_f32"] = {
        "batch": B, "constraints": ncon, "configs_per_s": B / t, "ms": t * 1e3,
        "note": "CRBA, RNEA(a=0) in FP32; FP64 tree factor, KKT solve and LLT; J = both feet's 6D Jacobians"}

    # constraintDynamics: both feet held in 6-D on the world frames where they are at a
    # reference posture q0, LOCAL / LOCAL_WORLD_ALIGNED, Baumgarte gains (12 rows), the
    # reference's default settings (mu 0, one iteration); configurations scattered around q0
    qc = M.perturb_configurations(desc, q32[:, 0], B, 0.25, M.INPUT_SEED_BASE + 900).astype(np.float32)
    Qc = torch.tensor(qc, device=dev).contiguous()
    oM0 = eng.forward_kinematics(Q[:, :1].contiguous())[0].double().cpu().numpy()[:, 0]
    foot = np.concatenate([np.eye(3).ravel(), [0.03, -0.01, -0.08]])

    def foot_world(j):
        R, p = oM0[12 * (j - 1):12 * (j - 1) + 9].reshape(3, 3), oM0[12 * (j - 1) + 9:12 * j]
        return np.concatenate([R.ravel(), p + R @ foot[9:]])

    rig = [pa.RigidConstraintModel(pa.CONTACT_6D, 7, foot, 0, foot_world(7), pa.LOCAL, [10] * 6, [2] * 6),
           pa.RigidConstraintModel(pa.CONTACT_6D, 13, foot, 0, foot_world(13), pa.LOCAL_WORLD_ALIGNED, [10] * 6,
                                   [2] * 6)]
    kddq = torch.empty_like(cddq)
    klam = torch.empty_like(clam)
    t = time_call(lambda: eng.constraint_dynamics(Qc, V, T, rig, out=kddq, lambda_out=klam, async_=True), reps=10)
    others["constraint_dynamics_humanoid36_f32"] = {
        "batch": B, "constraints": "two 6D feet, Baumgarte Kp 10 Kd 2", "configs_per_s": B / t, "ms": t * 1e3,
        "note": "kinematics, CRBA, RNEA(a=0) in FP32; constraint rows, FP64 tree factor and KKT in FP64"}

    cres = contact_bench(torch, pa, ctx, eng, Qc, B, dev, time_call)
    others.update(cres["others"])
    Gp, gv, mus, admm, xg = cres["Gp"], cres["g"], cres["mu"], cres["admm"], cres["xg"]

    idq = torch.empty_like(cddq)
    iimp = torch.empty_like(clam)
    t = time_call(lambda: eng.impulse_dynamics(Q, V, Jrows, r_coeff=0.0, out=idq, impulse_out=iimp, async_=True),
                  reps=10)
    others["impulse_dynamics_humanoid36_f32"] = {"batch": B, "constraints": ncon, "configs_per_s": B / t,
                                                 "ms": t * 1e3}
    # frames: both feet and both hands (end joints) with a 5 cm offset, placements and
    # LOCAL_WORLD_ALIGNED Jacobians (synchronous call: tables copied from the host)
    fj = [7, 13, 22, 29]
    fpl = np.tile(np.concatenate([np.eye(3).ravel(), [0.0, 0.0, -0.05]]), (len(fj), 1))
    t = time_call(lambda: eng.frames(Q, fj, fpl, "local_world_aligned"), reps=10)
    others["frames_humanoid36_f32"] = {"batch": B, "frames": len(fj), "configs_per_s": B / t, "ms": t * 1e3,
                                       "GB_s": 4 * (desc.nq + len(fj) * (12 + 6 * nv)) * B / t / 1e9}
    # frame velocities + spatial and classical accelerations of the same four frames (LOCAL),
    # and the WORLD joint Jacobian with its time variation (synchronous calls)
    t = time_call(lambda: eng.frame_motions(Q, V, fj, fpl, a=A), reps=10)
    others["frame_motions_humanoid36_f32"] = {"batch": B, "frames": len(fj), "configs_per_s": B / t,
                                              "ms": t * 1e3}
    t = time_call(lambda: eng.joint_jacobians_time_variation(Q, V), reps=10)
    others["jacobian_time_variation_humanoid36_f32"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3}

    # drop-in host path: numpy matrices in and out (PCIe both ways), not the device-resident figure
    pool = pa.ModelPool(desc, size=8)
    qh, vh, th = (np.asfortranarray(x) for x in (q, v, tau))
    pa.abaInParallel(8, pool, qh, vh, th)
    t0_ = time.perf_counter()
    for _ in range(3):
        pa.abaInParallel(8, pool, qh, vh, th)
    t = (time.perf_counter() - t0_) / 3
    others["aba_humanoid36_host_arrays"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3,
                                            "note": "abaInParallel on float64 numpy arrays: conversion, PCIe both ways"}

    # mixed-robot batch: RNEA + ABA per instance, one graph-replayed launch
    mb, n_mix = mixed_batch(torch, pa, M, ctx, dev)
    mb.run()
    tune = mb.tuning()
    t = time_call(lambda: mb.run(async_=True), reps=10)
    others["mixed_rnea_aba_f32"] = {"instances": n_mix, "groups": dict(MIX_SPEC), "configs_per_s": n_mix / t,
                                    "ms": t * 1e3, "chosen": tune.get("chosen"),
                                    "candidates_ms": {c["variant"]: c["ms"] for c in tune.get("candidates", [])}}
    # run-time-topology kernels (the tree read as data; what every tree without prebuilt
    # kernels runs on) on the same humanoid and inputs, against the prebuilt kernels
    ert = pa.Engine(pa.Model(ctx, desc, runtime_topology=True), "batch")
    Mrt = torch.empty_like(Mbuf)
    rt = {}
    for name, f_pre, f_rt, bb in (
            ("rnea", lambda: eng.rnea(Q, V, A, async_=True), lambda: ert.rnea(Q, V, A, async_=True), B),
            ("aba", lambda: eng.aba(Q, V, T, out=ddq, async_=True), lambda: ert.aba(Q, V, T, async_=True), B),
            ("crba", lambda: eng.crba(Q, out=Mbuf, async_=True), lambda: ert.crba(Q, out=Mrt, async_=True), B),
            ("rnea_derivatives", lambda: eng.rnea_derivatives(Qd, Vd, Ad, async_=True),
             lambda: ert.rnea_derivatives(Qd, Vd, Ad, async_=True), Bd),
            ("aba_derivatives", lambda: eng.aba_derivatives(Qd, Vd, Td, async_=True),
             lambda: ert.aba_derivatives(Qd, Vd, Td, async_=True), Bd)):
        tp, tr = time_call(f_pre, reps=5), time_call(f_rt, reps=5)
        rt[name] = {"batch": bb, "prebuilt_configs_per_s": bb / tp, "runtime_topology_configs_per_s": bb / tr,
                    "runtime_over_prebuilt": tp / tr}
    others["runtime_topology_humanoid36_f32"] = rt
    # FP64 ABA (prebuilt four-wave kernel) and ABA derivatives (FP64 RNEA derivatives and
    # solve on the run-time-topology kernels), the reference's precision
    Q64, V64, T64 = (torch.tensor(x, dtype=torch.float64, device=dev) for x in (q, v, tau))
    t = time_call(lambda: eng.aba(Q64, V64, T64, async_=True), reps=5)
    others["aba_humanoid36_f64"] = {"batch": B, "configs_per_s": B / t, "ms": t * 1e3}
    Qd64, Vd64, Td64 = (x[:, :Bd].contiguous() for x in (Q64, V64, T64))
    t = time_call(lambda: eng.aba_derivatives(Qd64, Vd64, Td64, async_=True), reps=3)
    others["aba_derivatives_humanoid36_f64"] = {"batch": Bd, "configs_per_s": Bd / t, "ms": t * 1e3}

    if not args.no_surrogate:
        others["surrogate_quadruped18"] = surrogate_bench(torch, pa, M, ctx, args)

    # ---------------- accuracy against the FP64 oracle ----------------
    errors = {}
    if not args.no_check:
        # the step's outputs over the whole 65536 batch and the derivatives over the whole
        # 16384 batch: max, median, 99.9th percentile of the per-configuration error
        qs, vs, as_, ts = (x.astype(np.float64) for x in (q32, v32, a32, tau32))
        step()  # the step's outputs of the timed buffers
        ctx.synchronize()
        errors["aba_ddq"] = stats(err_cols(ddq.double().cpu().numpy(), oracle.aba(desc, qs, vs, ts)))
        Mh = pa.Engine.mats_to_numpy(Mbuf, "batch", nv)
        errors["crba_M"] = stats(err_mats(Mh, oracle.crba(desc, qs)))
        _, Dref, _, _ = oracle.cholesky(desc, Mh)
        errors["cholesky_D_vs_fp64_of_same_M"] = stats(np.abs(Dbuf.cpu().numpy().T / Dref - 1).max(1))
        del Mh, Dref
        errors["rnea_tau"] = stats(err_cols(eng.rnea(Q, V, A).double().cpu().numpy(), oracle.rnea(desc, qs, vs, as_)))
        o = eng.rnea_derivatives(Qd, Vd, Ad)
        r = oracle.rnea_derivatives(desc, qs[:, :Bd], vs[:, :Bd], as_[:, :Bd])
        for name, x, ref in zip(("dtau_dq", "dtau_dv"), o[:2], r[:2]):
            errors[name] = stats(err_mats(pa.Engine.mats_to_numpy(x, "batch", nv), ref))
        del o, r
        o = eng.aba_derivatives(Qd, Vd, Td)
        r = oracle.aba_derivatives(desc, qs[:, :Bd], vs[:, :Bd], ts[:, :Bd])
        for name, x, ref in zip(("ddq_dq", "ddq_dv", "Minv"), o[:3], r[:3]):
            errors[name] = stats(err_mats(pa.Engine.mats_to_numpy(x, "batch", nv), ref))
        del o, r
        # FP64 ABA and derivatives against the oracle on the same (unrounded) inputs
        nf = 4096
        q64, v64, t64 = q[:, :nf], v[:, :nf], tau[:, :nf]
        g64 = eng.aba(*(torch.tensor(x, dtype=torch.float64, device=dev) for x in (q64, v64, t64)))
        errors["aba_ddq_f64"] = stats(err_cols(g64.cpu().numpy(), oracle.aba(desc, q64, v64, t64)))
        o = eng.aba_derivatives(*(torch.tensor(x, dtype=torch.float64, device=dev) for x in (q64, v64, t64)))
        r = oracle.aba_derivatives(desc, q64, v64, t64)
        for name, x, ref in zip(("ddq_dq_f64", "ddq_dv_f64", "Minv_f64"), o[:3], r[:3]):
            errors[name] = stats(err_mats(pa.Engine.mats_to_numpy(x, "batch", nv), ref))
        del o, r
        nd = 1024
        Jh = Jrows[:, :nd].double().cpu().numpy().reshape(nv, ncon, nd).transpose(2, 1, 0)
        ra, rl = oracle.constrained_forward_dynamics(desc, qs[:, :nd], vs[:, :nd], ts[:, :nd], Jh,
                                                     np.zeros((ncon, nd)))
        errors["constrained_ddq"] = rel_cols(cddq[:, :nd].double().cpu().numpy(), ra)
        errors["constrained_lambda"] = rel_cols(clam[:, :nd].double().cpu().numpy(), rl)
        rcon = [dict(type=c.type, joint1=c.joint1_id, joint2=c.joint2_id, reference_frame=c.reference_frame,
                     placement1=c.joint1_placement, placement2=c.joint2_placement, Kp=c.Kp, Kd=c.Kd) for c in rig]
        ka, kl, _ = oracle.constraint_dynamics(desc, qc[:, :nd].astype(np.float64), vs[:, :nd], ts[:, :nd], rcon)
        errors["constraint_dynamics_ddq"] = rel_cols(kddq[:, :nd].double().cpu().numpy(), ka)
        errors["constraint_dynamics_lambda"] = rel_cols(klam[:, :nd].double().cpu().numpy(), kl)
        # ADMM and PGS on the first 1,024 four-foot-point problems (solutions of the timed runs)
        Gh = Gp[:1024].cpu().numpy().reshape(1024, 12, 12).transpose(0, 2, 1)
        gh, mh = gv[:1024].cpu().numpy(), mus[:1024].cpu().numpy()
        ra = oracle.admm_solve(Gh, gh, mh, max_it=500, absolute_precision=1e-8)
        ya = admm.getPrimalSolution()[:1024].cpu().numpy()
        errors["admm_contact_y"] = float(np.abs(ya - ra["y"]).max() / np.abs(ra["y"]).max())
        errors["admm_contact_iterations_equal"] = float((admm.getIterationCount()[:1024].cpu().numpy() ==
                                                         ra["it"]).mean())
        rp = oracle.pgs_solve(Gh, gh, mh, max_it=500, absolute_precision=1e-8)
        errors["pgs_contact_x"] = float(np.abs(xg[:1024].cpu().numpy() - rp["x"]).max() / np.abs(rp["x"]).max())
        rm = oracle.rnea(man, qm[:, :1024], vm[:, :1024], am[:, :1024])
        errors["rnea_manipulator6_f64"] = rel_cols(em.rnea(Qm[:, :1024].contiguous(), Vm[:, :1024].contiguous(),
                                                           Am[:, :1024].contiguous()).cpu().numpy(), rm)

    # ---------------- CPU baseline (oracle, FP64, OpenMP over configurations) ----------------
    # The loop is timed inside the oracle library (C++ clock around an OpenMP static
    # schedule over configurations, one Data per thread, as benchmark/timings-parallel.cpp
    # drives rneaInParallel / abaInParallel), median of `reps` after one warm-up pass.
    cpu = None
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        nb = min(B, 65536)
        reps = 9
        sec = oracle.time_batch(desc, "step", q[:, :nb], v[:, :nb], tau[:, :nb], nthreads=cores, reps=reps)
        cpu = {"value": nb / sec, "unit": "configs/s", "cores": cores, "kind": "port",
               "sample": f"{nb} humanoid36 configurations x {reps} repetitions (median), "
                         f"ABA(WORLD)+CRBA+Cholesky in FP64, OpenMP static schedule over configurations",
               "host": host_info(), "threads": cores,
               "omp": {"OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND", "unset"),
                       "OMP_PLACES": os.environ.get("OMP_PLACES", "unset"), "schedule": "static"}}
        # single-thread latency per call, against the reference's published single-thread
        # figures (README.md:138-143: TALOS, the closest robot, ~4.2 / 4.9 / 8.4 us for
        # RNEA / CRBA / ABA on an i7 @ 2.4 GHz)
        n1 = 4096
        lat = {}
        for alg, x in (("rnea", a), ("crba", a), ("aba", tau)):
            s1 = oracle.time_batch(desc, alg, q[:, :n1], v[:, :n1], x[:, :n1], nthreads=1, reps=5)
            lat[alg + "_us"] = s1 / n1 * 1e6
        lat["reference_readme_talos_us"] = {"rnea": 4.2, "crba": 4.9, "aba": 8.4}
        cpu["single_thread_latency"] = lat
        cpu_alg = {}
        for alg, nbb, x in (("rnea", nb, a), ("aba", nb, tau), ("rnea_derivatives", Bd, a),
                            ("aba_derivatives", Bd, tau)):
            s_ = oracle.time_batch(desc, alg, q[:, :nbb], v[:, :nbb], x[:, :nbb], nthreads=cores, reps=5)
            cpu_alg[alg] = {"batch": nbb, "configs_per_s": nbb / s_, "cores": cores}
        gpu_key = {"rnea": "rnea_humanoid36_f32", "aba": "aba_humanoid36_f32",
                   "rnea_derivatives": "rnea_derivatives_humanoid36_f32",
                   "aba_derivatives": "aba_derivatives_humanoid36_f32"}
        nbc = 16384
        Jh = Jrows[:, :nbc].double().cpu().numpy().reshape(nv, ncon, nbc).transpose(2, 1, 0)
        s_ = oracle.time_constrained(desc, q[:, :nbc], v[:, :nbc], tau[:, :nbc], Jh, np.zeros((ncon, nbc)),
                                     nthreads=cores, reps=3)
        cpu_alg["constrained_forward_dynamics"] = {"batch": nbc, "configs_per_s": nbc / s_, "cores": cores}
        gpu_key["constrained_forward_dynamics"] = "constrained_forward_dynamics_humanoid36_f32"
        for alg, c in cpu_alg.items():
            c["gpu_over_cpu"] = others[gpu_key[alg]]["configs_per_s"] / c["configs_per_s"]
        # the contact solvers on a sample of the same four-foot-point problems
        ns = 8192
        Gh = Gp[:ns].cpu().numpy().reshape(ns, 12, 12).transpose(0, 2, 1)
        gh, mh = gv[:ns].cpu().numpy(), mus[:ns].cpu().numpy()
        for name, fn in (("admm_contact", lambda: oracle.admm_solve(Gh, gh, mh, max_it=500, absolute_precision=1e-8,
                                                                    nthreads=cores)),
                         ("pgs_contact", lambda: oracle.pgs_solve(Gh, gh, mh, max_it=500, absolute_precision=1e-8,
                                                                  nthreads=cores))):
            fn()
            t0 = time.perf_counter()
            fn()
            s_ = time.perf_counter() - t0
            cpu_alg[name] = {"batch": ns, "problems_per_s": ns / s_, "cores": cores,
                             "gpu_over_cpu": others[name + "_humanoid36"]["problems_per_s"] / (ns / s_)}
        others["cpu_fp64_oracle"] = cpu_alg
        # ABA followed by computeABADerivatives per configuration, as one pipeline
        g = 1.0 / (1.0 / others["aba_humanoid36_f32"]["configs_per_s"] +
                   1.0 / others["aba_derivatives_humanoid36_f32"]["configs_per_s"])
        c_ = 1.0 / (1.0 / cpu_alg["aba"]["configs_per_s"] + 1.0 / cpu_alg["aba_derivatives"]["configs_per_s"])
        others["aba_plus_aba_derivatives_gpu_over_cpu"] = g / c_

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Cholesky)",
        "data": "synthetic (seeded inputs on a seeded random-inertia model)",
        "config": {"workload": "humanoid36 ABA + CRBA + FP64 cholesky::decompose", "batch": B,
                   "nq": desc.nq, "nv": desc.nv, "layout": "batch-fastest, device-resident", "parallelism": "none"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "others": others,
        "max_rel_err": errors,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
