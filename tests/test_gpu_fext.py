"""GPU parity of rnea / aba with external forces (rnea.hpp:60-79, aba.hpp:63-85) against the
FP64 oracle (pinned by virtual work in test_oracle_fext.py), on a prebuilt tree (the forces
go through the run-time-topology kernels for every tree) and on one that is not prebuilt,
both layouts, FP32 (bars 1e-5 RNEA, 1e-4 ABA, relative per configuration) and FP64 (1e-12
RNEA, conditioning-scaled for ABA as in test_gpu_generic.py)."""
import numpy as np
import pytest
import torch

import oracle
import pinocchio_amd as pa
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols

pytestmark = pytest.mark.gpu

TREES = {"humanoid36": M.humanoid36,
         "tree24": lambda: M.random_tree(24, M.MODEL_SEED_BASE + 108, free_flyer=True, compact=False)}


def dev(x, layout, dtype):
    a = x.T if layout == "col" else x
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t, layout):
    a = t.double().cpu().numpy()
    return a.T if layout == "col" else a


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", list(TREES))
def test_rnea_aba_fext(gpu_ctx, name, layout):
    desc = TREES[name]()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), layout)
    B = 333
    q, v, a, tau = inputs(desc, B, k=96)
    f = np.random.default_rng(6).normal(size=(6 * (desc.njoints - 1), B))
    for dtype in (torch.float32, torch.float64):
        qq, vv, aa, tt, ff = f32_round(q, v, a, tau, f) if dtype == torch.float32 else (q, v, a, tau, f)
        rt = oracle.rnea_fext(desc, qq, vv, aa, ff)
        ra = oracle.aba_fext(desc, qq, vv, tt, ff)
        gt = host(eng.rnea_fext(dev(qq, layout, dtype), dev(vv, layout, dtype), dev(aa, layout, dtype),
                                dev(ff, layout, dtype)), layout)
        ga = host(eng.aba_fext(dev(qq, layout, dtype), dev(vv, layout, dtype), dev(tt, layout, dtype),
                               dev(ff, layout, dtype)), layout)
        et, ea = rel_err_cols(gt, rt), rel_err_cols(ga, ra)
        print(f"{name} {layout} {dtype}: rnea {et.max():.2e} aba {ea.max():.2e}")
        if dtype == torch.float32:
            assert et.max() < 1e-5 and ea.max() < 1e-4
        else:
            # ABA's FP64 bar follows the conditioning the oracle's FP32 run shows
            amp = rel_err_cols(oracle.aba_fext(desc, qq, vv, tt, ff, scalar="f32"), ra).max() / 2 ** -24
            assert et.max() < 1e-12 and ea.max() < max(1e-12, 4 * amp * 2 ** -53 * 2 ** 24)


@pytest.mark.parametrize("layout", ["col", "batch"])
@pytest.mark.parametrize("name", list(TREES))
def test_static_torque(gpu_ctx, name, layout):
    """computeStaticTorque (rnea.hpp:163) = rnea(q, 0, 0, fext): against the oracle's RNEA
    with external forces at v = a = 0, and bit-identical to rbd_rnea_fext given explicit
    zero arrays."""
    desc = TREES[name]()
    eng = pa.Engine(pa.Model(gpu_ctx, desc), layout)
    B = 257
    q, _, _, _ = inputs(desc, B, k=97)
    f = np.random.default_rng(7).normal(size=(6 * (desc.njoints - 1), B))
    z = np.zeros((desc.nv, B))
    for dtype in (torch.float32, torch.float64):
        qq, ff = f32_round(q, f) if dtype == torch.float32 else (q, f)
        Q, F, Z = dev(qq, layout, dtype), dev(ff, layout, dtype), dev(z, layout, dtype)
        g = eng.static_torque(Q, F)
        assert torch.equal(g, eng.rnea_fext(Q, Z, Z, F))
        e = rel_err_cols(host(g, layout), oracle.rnea_fext(desc, qq, z, z, ff)).max()
        assert e < (1e-5 if dtype == torch.float32 else 1e-12), e
