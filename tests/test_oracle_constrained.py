"""Pins the oracle's constrained forward dynamics (forwardDynamics with J, gamma and
inv_damping, contact-dynamics.hxx:90-167) without the reference.

  * the KKT system [[M, J^T], [J, 0]] [ddq; -lambda] = [tau - nle; -gamma], solved densely
    by numpy with M from the oracle's CRBA and nle = rnea(q, v, 0) (both pinned in
    test_oracle.py), gives the same ddq and lambda; the constraint J ddq + gamma = 0 holds;
  * with damping: lambda = -(J M^-1 J^T + mu I)^-1 (J M^-1 (tau - nle) + gamma);
  * J = rows of the identity with gamma = -a* prescribes those accelerations exactly, and
    then rnea(q, v, ddq) - tau = J^T lambda (Newton-Euler balance by RNEA alone);
  * planted mistakes (a flipped gamma sign, a dropped nle) fail the dense check;
and its impulseDynamics (contact-dynamics.hxx:264-331): the impact KKT system
[[M, J^T], [J, 0]] [v+; -lambda] = [M v-; -e J v-] solved by numpy, J v+ = -e J v-, and the
damped closed form.
"""
import numpy as np
import pytest

import oracle
from pinocchio_amd import models as M

NAMES = ["arm7", "quadruped18", "humanoid36"]


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.build()
    oracle.lib()


def mk(name):
    return getattr(M, name)()


def problem(desc, B, nc, seed):
    q, v, _, tau = M.sample_inputs(desc, B, 4321 + seed)
    rng = np.random.default_rng(seed)
    J = rng.standard_normal((B, nc, desc.nv))
    gamma = rng.standard_normal((nc, B))
    return q, v, tau, J, gamma


def dense(desc, q, v, tau, J, gamma, mu=0.0):
    """ddq, lambda from the dense KKT system per configuration."""
    Ms = oracle.crba(desc, q)
    nle = oracle.rnea(desc, q, v, np.zeros_like(v))
    B, nv = q.shape[1], desc.nv
    nc = J.shape[1]
    a = np.zeros((nv, B))
    lam = np.zeros((nc, B))
    for b in range(B):
        Mb, Jb = Ms[b], J[b]
        if mu == 0.0:
            K = np.block([[Mb, Jb.T], [Jb, np.zeros((nc, nc))]])
            sol = np.linalg.solve(K, np.concatenate([tau[:, b] - nle[:, b], -gamma[:, b]]))
            a[:, b], lam[:, b] = sol[:nv], -sol[nv:]
        else:
            r = np.linalg.solve(Mb, tau[:, b] - nle[:, b])
            G = Jb @ np.linalg.solve(Mb, Jb.T) + mu * np.eye(nc)
            lam[:, b] = np.linalg.solve(G, -Jb @ r - gamma[:, b])
            a[:, b] = r + np.linalg.solve(Mb, Jb.T @ lam[:, b])
    return a, lam


def relmax(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-30))


@pytest.mark.parametrize("name", NAMES)
def test_matches_dense_kkt(name):
    d = mk(name)
    nc = min(6, d.nv - 1)
    q, v, tau, J, gamma = problem(d, 7, nc, 1)
    a, lam = oracle.constrained_forward_dynamics(d, q, v, tau, J, gamma)
    a_ref, lam_ref = dense(d, q, v, tau, J, gamma)
    assert relmax(a, a_ref) < 1e-9 and relmax(lam, lam_ref) < 1e-9
    res = np.einsum("bik,kb->ib", J, a) + gamma
    assert np.abs(res).max() < 1e-9 * max(1.0, np.abs(gamma).max())


@pytest.mark.parametrize("name", ["quadruped18", "humanoid36"])
def test_damping_matches_dense(name):
    d = mk(name)
    q, v, tau, J, gamma = problem(d, 5, 12, 2)
    a, lam = oracle.constrained_forward_dynamics(d, q, v, tau, J, gamma, inv_damping=1e-3)
    a_ref, lam_ref = dense(d, q, v, tau, J, gamma, mu=1e-3)
    assert relmax(a, a_ref) < 1e-9 and relmax(lam, lam_ref) < 1e-9


@pytest.mark.parametrize("name", ["arm7", "humanoid36"])
def test_prescribed_accelerations_balance_rnea(name):
    d = mk(name)
    B = 4
    q, v, _, tau = M.sample_inputs(d, B, 999)
    dofs = [0, 2, d.nv - 1]
    J = np.zeros((B, len(dofs), d.nv))
    for i, k in enumerate(dofs):
        J[:, i, k] = 1.0
    astar = np.random.default_rng(3).standard_normal((len(dofs), B))
    a, lam = oracle.constrained_forward_dynamics(d, q, v, tau, J, -astar)
    assert np.abs(a[dofs] - astar).max() < 1e-10
    # Newton-Euler balance: rnea(q, v, ddq) - tau = J^T lambda
    f = oracle.rnea(d, q, v, a) - tau
    JTl = np.einsum("bik,ib->kb", J, lam)
    assert relmax(f, JTl) < 1e-9


def test_planted_mistakes_are_caught():
    d = mk("quadruped18")
    q, v, tau, J, gamma = problem(d, 3, 6, 4)
    a_ref, lam_ref = dense(d, q, v, tau, J, gamma)
    # gamma with the wrong sign
    a, lam = oracle.constrained_forward_dynamics(d, q, v, tau, J, -gamma)
    assert relmax(a, a_ref) > 1e-3
    # nle dropped: equivalent to feeding tau + nle
    nle = oracle.rnea(d, q, v, np.zeros_like(v))
    a, lam = oracle.constrained_forward_dynamics(d, q, v, tau + nle, J, gamma)
    assert relmax(a, a_ref) > 1e-3


# ---- impulseDynamics (contact-dynamics.hxx:264-331) ---------------------------------------
@pytest.mark.parametrize("name", ["arm7", "quadruped18", "humanoid36"])
@pytest.mark.parametrize("e", [0.0, 0.4])
def test_impulse_matches_dense_kkt(name, e):
    """[[M, J^T], [J, 0]] [v+; -lambda] = [M v-; -e J v-] solved by numpy; J v+ = -e J v-."""
    d = mk(name)
    nc = min(6, d.nv - 1)
    q, v, _, J, _ = problem(d, 6, nc, 7)
    vp, lam = oracle.impulse_dynamics(d, q, v, J, r_coeff=e)
    Ms = oracle.crba(d, q)
    for b in range(q.shape[1]):
        Mb, Jb = Ms[b], J[b]
        K = np.block([[Mb, Jb.T], [Jb, np.zeros((nc, nc))]])
        sol = np.linalg.solve(K, np.concatenate([Mb @ v[:, b], -e * (Jb @ v[:, b])]))
        assert relmax(vp[:, b], sol[:d.nv]) < 1e-9 and relmax(lam[:, b], -sol[d.nv:]) < 1e-9
        assert np.abs(Jb @ vp[:, b] + e * (Jb @ v[:, b])).max() < 1e-9 * max(1.0, np.abs(Jb @ v[:, b]).max())


def test_impulse_damping_matches_dense():
    d = mk("humanoid36")
    q, v, _, J, _ = problem(d, 4, 12, 8)
    mu = 1e-3
    vp, lam = oracle.impulse_dynamics(d, q, v, J, r_coeff=0.2, inv_damping=mu)
    Ms = oracle.crba(d, q)
    for b in range(q.shape[1]):
        Mb, Jb = Ms[b], J[b]
        G = Jb @ np.linalg.solve(Mb, Jb.T) + mu * np.eye(12)
        lr = np.linalg.solve(G, -1.2 * (Jb @ v[:, b]))
        assert relmax(lam[:, b], lr) < 1e-9
        assert relmax(vp[:, b], v[:, b] + np.linalg.solve(Mb, Jb.T @ lr)) < 1e-9
