"""External forces in the oracle's RNEA and ABA (rnea.hpp:60-79, aba.hpp:63-85).

Pinned by the principle of virtual work: a force f_i on joint i's frame, expressed in that
frame, contributes -J_i^T f_i to the inverse dynamics, with J_i the frame's LOCAL Jacobian
(from the frames oracle, itself pinned by finite differences) -- so
rnea(q, v, a, fext) = rnea(q, v, a) - sum_i J_i^T f_i; and ABA with the same forces inverts
it.  A planted mistake (forces taken in the world frame) fails the first check.
"""
import numpy as np

import oracle
from pinocchio_amd import models as M

IDENT = np.concatenate([np.eye(3).ravel(), np.zeros(3)])


def forces(desc, B, seed):
    return np.random.default_rng(seed).normal(size=(6 * (desc.njoints - 1), B))


def generalized(desc, q, f, frame="local"):
    """sum_i J_i^T f_i with J_i the joint frames' Jacobians"""
    nj = desc.njoints
    _, J = oracle.frames(desc, q, list(range(1, nj)), np.tile(IDENT, (nj - 1, 1)), frame)
    B = q.shape[1]
    out = np.zeros((desc.nv, B))
    for b in range(B):
        for i in range(nj - 1):
            out[:, b] += J[b, i].T @ f[6 * i:6 * i + 6, b]
    return out


def test_rnea_fext_is_virtual_work():
    for desc in (M.humanoid36(), M.arm7()):
        q, v, a, _ = M.sample_inputs(desc, 5, M.INPUT_SEED_BASE + 950)
        f = forces(desc, 5, 1)
        tau0 = oracle.rnea(desc, q, v, a)
        tau = oracle.rnea_fext(desc, q, v, a, f)
        want = tau0 - generalized(desc, q, f)
        assert np.abs(tau - want).max() < 1e-10 * np.abs(want).max()
        # forces read in the world frame instead give another torque
        wrong = tau0 - generalized(desc, q, f, "world")
        assert np.abs(tau - wrong).max() > 1e-3


def test_aba_fext_inverts_rnea_fext():
    desc = M.humanoid36()
    q, v, a, _ = M.sample_inputs(desc, 5, M.INPUT_SEED_BASE + 951)
    f = forces(desc, 5, 2)
    tau = oracle.rnea_fext(desc, q, v, a, f)
    ddq = oracle.aba_fext(desc, q, v, tau, f)
    assert np.abs(ddq - a).max() < 1e-10 * np.abs(a).max()
    # zero forces: the plain algorithms
    z = np.zeros_like(f)
    np.testing.assert_allclose(oracle.rnea_fext(desc, q, v, a, z), oracle.rnea(desc, q, v, a), rtol=0, atol=1e-12)
