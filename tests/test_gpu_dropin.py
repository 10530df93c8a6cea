"""The C++ drop-in (include/pinocchio_amd/parallel.hpp) driven like a pinocchio caller:
tests/cpp/dropin_check builds a ModelPool from a flattened model, runs rneaInParallel,
abaInParallel, crbaInParallel, computeRNEADerivativesInParallel,
computeABADerivativesInParallel, computeMinverseInParallel, forwardDynamicsInParallel
(constrained) and constraintDynamicsInParallel (a point constraint with Baumgarte gains and
proximal iterations) on column-major host matrices, and checks the reference's argument
errors.
Outputs against the FP64 oracle (the calls default to FP64, like the reference): RNEA,
ABA and CRBA 1e-12, RNEA and ABA derivatives and Minv 1e-10 (also with outputs whose
leading dimension exceeds nv), constrained ddq and lambda 1e-8 with a random J; with the
FP32 option: ABA 1e-4, RNEA derivatives 1e-5, ABA derivatives 1e-4 (per configuration,
inf-norm relative; oracle on the FP32-rounded inputs for the FP32 paths).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from pinocchio_amd import models as M
from helpers import f32_round, inputs, rel_err_cols, rel_err_mats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_check")


def rd_cols(d, name, rows, B):
    return np.fromfile(os.path.join(d, f"out_{name}.f64")).reshape(B, rows).T


def blocks(x, nv, B):
    """nv x (nv * B) column-major blocks -> (B, nv, nv)."""
    return x.reshape(B, nv, nv).transpose(0, 2, 1)


@pytest.mark.parametrize("name", ["humanoid36", "manipulator6"])
def test_cpp_dropin(tmp_path, name):
    assert os.path.exists(BIN), "tests/cpp/dropin_check is not built (make -C pinocchio_amd/csrc)"
    desc = M.SAMPLE_MODELS[name]()
    B = 300
    q, v, a, tau = inputs(desc, B, k=60)
    d = str(tmp_path)
    nc = 6 if desc.nv > 12 else 3  # random J: the error grows with the conditioning of J M^-1 J^T
    J = np.random.default_rng(61).standard_normal((B, nc, desc.nv))
    gamma = np.random.default_rng(62).standard_normal((nc, B))
    J.transpose(0, 2, 1).astype(np.float64).tofile(os.path.join(d, "J.f64"))  # per configuration column major
    np.asfortranarray(gamma).T.astype(np.float64).tofile(os.path.join(d, "gamma.f64"))
    with open(os.path.join(d, "dims.txt"), "w") as f:
        f.write(f"{desc.njoints} {desc.nq} {desc.nv} {B} {nc}\n")
    desc.parents.astype(np.int32).tofile(os.path.join(d, "parents.i32"))
    desc.jtype.astype(np.int32).tofile(os.path.join(d, "jtype.i32"))
    desc.placement.astype(np.float64).tofile(os.path.join(d, "placement.f64"))
    desc.inertia.astype(np.float64).tofile(os.path.join(d, "inertia.f64"))
    for n, x in (("q", q), ("v", v), ("a", a), ("tau", tau)):
        np.asfortranarray(x).T.astype(np.float64).tofile(os.path.join(d, f"{n}.f64"))  # column-major
    r = subprocess.run([BIN, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors 1 1 1 1" in r.stdout
    nv = desc.nv

    def rd(n, shape_cols):
        return np.fromfile(os.path.join(d, f"out_{n}.f64")).reshape(shape_cols, nv).T

    tau_g = rd("tau", B)
    assert rel_err_cols(tau_g, oracle.rnea(desc, q, v, a)).max() < 1e-12
    qf, vf, af, tf = f32_round(q, v, a, tau)
    assert rel_err_cols(rd("ddq", B), oracle.aba(desc, q, v, tau)).max() < 1e-12
    assert rel_err_cols(rd("ddq32", B), oracle.aba(desc, qf, vf, tf)).max() < 1e-4
    Mg = blocks(np.fromfile(os.path.join(d, "out_M.f64")), nv, B)
    assert rel_err_mats(Mg, oracle.crba(desc, q)).max() < 1e-12

    def mats(n):
        return blocks(np.fromfile(os.path.join(d, f"out_{n}.f64")), nv, B)

    rdq, rdv, _, _ = oracle.rnea_derivatives(desc, q, v, a)
    for n, ref in (("dtau_dq", rdq), ("dtau_dv", rdv)):
        assert rel_err_mats(mats(n), ref).max() < 1e-10, n
    rq, rv, rmi, _ = oracle.aba_derivatives(desc, q, v, tau)
    for n, ref in (("ddq_dq", rq), ("ddq_dv", rv), ("minv", rmi), ("minv64", rmi), ("ddq_dv_ld", rv)):
        assert rel_err_mats(mats(n), ref).max() < 1e-10, n
    assert rel_err_mats(mats("dtau_dq32"), oracle.rnea_derivatives(desc, qf, vf, af)[0]).max() < 1e-5
    rq32, _, rmi32, _ = oracle.aba_derivatives(desc, qf, vf, tf)
    assert rel_err_mats(mats("ddq_dq32"), rq32).max() < 1e-4
    assert rel_err_mats(mats("minv32"), rmi32).max() < 1e-4
    ca, cl = oracle.constrained_forward_dynamics(desc, q, v, tau, J, gamma)
    cddq = np.fromfile(os.path.join(d, "out_cddq.f64")).reshape(B, nv).T
    clam = np.fromfile(os.path.join(d, "out_lambda.f64")).reshape(B, nc).T
    assert rel_err_cols(cddq, ca).max() < 1e-8 and rel_err_cols(clam, cl).max() < 1e-8
    # constraintDynamics: a point of the last joint held on a world point, 3-D LOCAL, gains,
    # 10 proximal iterations
    eye = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    con = [dict(type=oracle.CONTACT_3D, joint1=desc.njoints - 1, reference_frame=oracle.LOCAL, placement1=eye,
                placement2=eye[:11] + [0.5], Kp=[4.0] * 3, Kd=[1.5] * 3)]
    kd, kl, kit = oracle.constraint_dynamics(desc, q, v, tau, con, mu=1e-3, max_iter=10, absolute_accuracy=1e-10,
                                             relative_accuracy=1e-10)
    assert rel_err_cols(rd_cols(d, "kddq", nv, B), kd).max() < 1e-8
    assert rel_err_cols(rd_cols(d, "klambda", 3, B), kl).max() < 1e-8
    assert np.array_equal(np.fromfile(os.path.join(d, "out_kiters.f64")).astype(np.int32), kit)
