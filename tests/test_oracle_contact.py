"""Pins of the oracle's frictional contact solvers (PGS, ADMM) and Delassus matrix.

The reference holds no unit test or fixture for its contact solvers, so the oracle is
pinned by what the mathematics fixes:
  * one contact with a known answer: sticking (lambda = -G^-1 g strictly inside the cone,
    zero contact velocity), sliding with a diagonal G (lambda_n = -g_n / G_nn, the tangential
    force on the cone's edge opposite the tangential velocity), separation (g_n > 0: zero);
  * several contacts: PGS and ADMM -- two different algorithms -- reach the same forces,
    and those forces satisfy the Coulomb conditions with the De Saxce shift: lambda in the
    cone, sigma + (0, 0, mu |sigma_t|) in the dual cone (sigma = G lambda + g), and their
    product zero per contact;
  * the reference's own bookkeeping: rho = sqrt(L m) (L / m)^rho_power at the start, the
    largest eigenvalue against numpy's, iteration conventions (PGS counts from 0, ADMM
    from 1), over_relax and tau range errors;
  * the Delassus matrix of a single free body against J M^-1 J^T formed from the spatial
    inertia matrix in closed form (point and frame contacts, a rotated, offset contact
    frame, mu > 0 through the definition J (M + mu J^T J)^-1 J^T);
  * planted mistakes (a flipped tangential sign in the sliding answer, a transposed
    contact rotation) fail those checks.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from pinocchio_amd import models as M


def spd(rng, n, cond=30.0):
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.geomspace(1.0, cond, n)
    return (Q * ev) @ Q.T


def ncp_residuals(G, g, mu, lam):
    """max over contacts of: distance of lambda outside its cone, of the shifted velocity
    outside the dual cone, and |lambda . shifted velocity|."""
    sig = G @ lam + g
    out = [0.0, 0.0, 0.0]
    for c in range(len(mu)):
        l, s = lam[3 * c:3 * c + 3], sig[3 * c:3 * c + 3].copy()
        s[2] += mu[c] * np.linalg.norm(s[:2])
        out[0] = max(out[0], np.linalg.norm(l[:2]) - mu[c] * l[2], -l[2])
        out[1] = max(out[1], mu[c] * np.linalg.norm(s[:2]) - s[2])
        out[2] = max(out[2], abs(l @ s))
    return out


TIGHT = dict(absolute_precision=1e-12, relative_precision=1e-14)


def test_sticking_contact_is_the_linear_solve():
    rng = np.random.default_rng(1)
    G = spd(rng, 3)
    lam = np.array([0.1, -0.05, 1.0])  # |t| = 0.11 < mu n = 0.8
    g = -G @ lam
    p = oracle.pgs_solve(G, g, [0.8], max_it=5000, **TIGHT)
    a = oracle.admm_solve(G, g, [0.8], max_it=5000, absolute_precision=1e-11, relative_precision=1e-14)
    assert p["converged"][0] and a["converged"][0]
    np.testing.assert_allclose(p["x"][0], lam, atol=1e-9)
    np.testing.assert_allclose(a["y"][0], lam, atol=1e-8)


@pytest.mark.parametrize("mu", [0.3, 0.9])
def test_sliding_contact_closed_form(mu):
    G = np.diag([2.0, 2.0, 3.0])  # isotropic tangential block: sigma_t stays along g_t
    g = np.array([5.0, -1.0, -2.0])
    ln = 2.0 / 3.0
    lt = -mu * ln * g[:2] / np.linalg.norm(g[:2])
    want = np.array([lt[0], lt[1], ln])
    p = oracle.pgs_solve(G, g, [mu], max_it=5000, **TIGHT)
    a = oracle.admm_solve(G, g, [mu], max_it=5000, absolute_precision=1e-11, relative_precision=1e-14)
    np.testing.assert_allclose(p["x"][0], want, atol=1e-9)
    np.testing.assert_allclose(a["y"][0], want, atol=1e-8)
    # a planted sign mistake in the tangential direction is caught by the same check
    bad = np.array([-lt[0], -lt[1], ln])
    assert np.abs(p["x"][0] - bad).max() > 0.1


def test_separating_contact_has_no_force():
    G = np.diag([2.0, 2.0, 3.0])
    g = np.array([0.3, 0.1, 2.0])
    assert np.abs(oracle.pgs_solve(G, g, [0.5])["x"]).max() == 0.0
    assert np.abs(oracle.admm_solve(G, g, [0.5])["y"]).max() < 1e-12


def dominant(rng, n):
    """SPD with a dominant diagonal: the regime where Gauss-Seidel over the cones converges
    (weakly coupled contacts)"""
    A = rng.normal(size=(n, n)) * 0.15
    return A @ A.T + np.diag(rng.uniform(1.0, 2.0, n))


@pytest.mark.parametrize("nc", [2, 4])
def test_pgs_and_admm_agree_and_satisfy_coulomb(nc):
    rng = np.random.default_rng(10 + nc)
    n = 3 * nc
    B = 24
    G = np.stack([dominant(rng, n) for _ in range(B)])
    g = rng.normal(size=(B, n))
    g[:, 2::3] -= 1.5  # mostly pressing contacts
    mu = rng.uniform(0.2, 1.0, size=(B, nc))
    p = oracle.pgs_solve(G, g, mu, max_it=20000, **TIGHT)
    a = oracle.admm_solve(G, g, mu, max_it=20000, absolute_precision=1e-10, relative_precision=1e-14)
    assert p["converged"].all() and a["converged"].all()
    for b in range(B):
        np.testing.assert_allclose(p["x"][b], a["y"][b], atol=2e-6, rtol=0)
        r = ncp_residuals(G[b], g[b], mu[b], a["y"][b])
        assert max(r) < 1e-6, (b, r)
        r = ncp_residuals(G[b], g[b], mu[b], p["x"][b])
        assert max(r) < 1e-6, (b, r)


def test_admm_solves_strongly_coupled_contacts():
    """condition number 20, strongly coupled: ADMM still meets the Coulomb conditions (PGS
    does not converge on some of these, as Gauss-Seidel may not)"""
    rng = np.random.default_rng(12)
    n, B = 6, 24
    G = np.stack([spd(rng, n, cond=20.0) for _ in range(B)])
    g = rng.normal(size=(B, n))
    g[:, 2::3] -= 1.5
    mu = rng.uniform(0.2, 1.0, size=(B, 2))
    a = oracle.admm_solve(G, g, mu, max_it=20000, absolute_precision=1e-10, relative_precision=1e-14)
    assert a["converged"].all()
    for b in range(B):
        assert max(ncp_residuals(G[b], g[b], mu[b], a["y"][b])) < 1e-8


def test_admm_bookkeeping():
    rng = np.random.default_rng(3)
    nc, n = 3, 9
    G = spd(rng, n, cond=50.0)
    g = rng.normal(size=n)
    R = np.full(n, 1e-4)
    a = oracle.admm_solve(G, g, [0.5] * nc, R=R, max_it=1, max_it_largest_eigen_value=200)
    L = np.linalg.eigvalsh(G)[-1]
    assert abs(a["largest_eigen_value"][0] - L) < 1e-6 * L
    m = 1e-6 + 1e-4
    Le = a["largest_eigen_value"][0]  # the power iteration's estimate, which rho is built on
    rho0 = np.sqrt(Le * m) * (Le / m) ** 0.2
    # one iteration: rho moved by at most one factor (L / m)^0.05
    f = (Le / m) ** 0.05
    assert any(abs(a["rho"][0] - r) < 1e-9 * r for r in (rho0, rho0 * f, rho0 / f))
    assert a["it"][0] == 2 and not a["converged"][0]  # it counts from 1; max_it + 1 when not converged
    p = oracle.pgs_solve(G, g, [0.5] * nc, max_it=3, absolute_precision=0, relative_precision=0)
    assert p["it"][0] == 3 and not p["converged"][0]
    with pytest.raises(RuntimeError, match="over_relax"):
        oracle.pgs_solve(G, g, [0.5] * nc, over_relax=2.0)
    with pytest.raises(RuntimeError, match="tau"):
        oracle.admm_solve(G, g, [0.5] * nc, tau=0.0)


# ---------------------------------------------------------------------------------------
# Delassus matrix of a free body in closed form
# ---------------------------------------------------------------------------------------
def skew(p):
    return np.array([[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]])


def spatial_inertia(row):
    m, c = row[0], row[1:4]
    xx, xy, yy, xz, yz, zz = row[4:10]
    Ic = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
    Mi = np.zeros((6, 6))
    Mi[:3, :3] = m * np.eye(3)
    Mi[3:, :3] = m * skew(c)
    Mi[:3, 3:] = -m * skew(c)
    Mi[3:, 3:] = Ic - m * skew(c) @ skew(c)
    return Mi


def contact_jacobian(R, p, six):
    """rows of the contact frame's LOCAL velocity from the body's (v, w)"""
    J = np.zeros((6, 6))
    J[:3, :3] = R.T
    J[:3, 3:] = -R.T @ skew(p)
    J[3:, 3:] = R.T
    return J if six else J[:3]


@pytest.mark.parametrize("six", [False, True])
@pytest.mark.parametrize("mu", [0.0, 0.7])
def test_delassus_of_a_free_body(six, mu):
    desc = M.free_body()
    Mb = spatial_inertia(desc.inertia[1]) + np.diag(desc.armature)
    R = Rotation.from_euler("xyz", [0.4, -0.3, 1.1]).as_matrix()
    p = np.array([0.12, -0.3, 0.25])
    cons = [dict(type=1 if six else 0, joint1=1, reference_frame=1, placement1=np.concatenate([R.ravel(), p]))]
    q = M.sample_inputs(desc, 3, M.INPUT_SEED_BASE + 901)[0]
    G = oracle.delassus(desc, q, cons, mu=mu)
    J = contact_jacobian(R, p, six)
    want = J @ np.linalg.inv(Mb + mu * J.T @ J) @ J.T
    for b in range(3):  # the LOCAL Delassus of one body does not depend on q
        np.testing.assert_allclose(G[b], want, rtol=1e-11, atol=1e-12)
    # a transposed contact rotation gives another matrix
    Jbad = contact_jacobian(R.T, p, six)
    assert np.abs(Jbad @ np.linalg.inv(Mb) @ Jbad.T - G[0]).max() > 1e-3 or mu > 0


def test_delassus_is_the_schur_complement_of_constraint_dynamics():
    """On the humanoid: the KKT multipliers of constraintDynamics (mu = 0) solve
    G lambda = -(J M^-1 (tau - nle) + gamma) ... checked through lambda = -G^-1 rhs' by
    the damped identity: with mu > 0, one proximal step gives (G + mu I) x = rhs, and
    G from delassus() reproduces the difference of two such steps."""
    desc = M.humanoid36()
    q, v, _, tau = M.sample_inputs(desc, 4, M.INPUT_SEED_BASE + 902)
    q[:, :] = q[:, :1]  # one posture, different velocities and torques are not needed
    cons = [dict(type=0, joint1=7, reference_frame=1), dict(type=0, joint1=13, reference_frame=1)]
    G = oracle.delassus(desc, q, cons)
    assert np.allclose(G, G.transpose(0, 2, 1), atol=1e-12)
    assert (np.linalg.eigvalsh(G) > 0).all()
    # lambda at two dampings: (G + mu I) x = rhs for both, so
    # (G + mu1 I) x1 - (G + mu2 I) x2 = 0
    _, l1, _ = oracle.constraint_dynamics(desc, q, v, tau, cons, mu=1e-2)
    _, l2, _ = oracle.constraint_dynamics(desc, q, v, tau, cons, mu=5e-2)
    for b in range(4):
        r = (G[b] + 1e-2 * np.eye(6)) @ l1[:, b] - (G[b] + 5e-2 * np.eye(6)) @ l2[:, b]
        assert np.abs(r).max() < 1e-9 * max(1.0, np.abs(l1[:, b]).max())


# ---------------------------------------------------------------------------------------
# the contact Cholesky's KKT solve
# ---------------------------------------------------------------------------------------
KKT_CONS = [dict(type=0, joint1=7, reference_frame=1, placement1=np.concatenate([np.eye(3).ravel(), [0.03, 0, -0.08]])),
            dict(type=1, joint1=13, reference_frame=2)]


def test_kkt_solve_constraint_rows_are_the_damped_delassus_solve():
    """b = [b1; 0]: x1 = -(J M^-1 J^T + mu I)^-1 b1, with the Delassus matrix pinned above"""
    desc = M.humanoid36()
    q = M.sample_inputs(desc, 3, M.INPUT_SEED_BASE + 903)[0]
    nr = 9
    b1 = np.random.default_rng(4).normal(size=(nr, 3))
    for mu in (0.0, 0.01):
        x = oracle.contact_cholesky_solve(desc, q, KKT_CONS, mu, np.vstack([b1, np.zeros((desc.nv, 3))]))
        G = oracle.delassus(desc, q, KKT_CONS)
        for b in range(3):
            want = -np.linalg.solve(G[b] + mu * np.eye(nr), b1[:, b])
            assert np.abs(x[:nr, b] - want).max() < 1e-10 * np.abs(want).max()


def test_kkt_solve_is_constraint_dynamics_at_rest():
    """at v = 0 with no correction the constraint drift vanishes, so K [-lambda; ddq] =
    [0; tau - g(q)] is constraintDynamics' own system"""
    desc = M.humanoid36()
    q, _, _, tau = M.sample_inputs(desc, 3, M.INPUT_SEED_BASE + 904)
    v0 = np.zeros((desc.nv, 3))
    ddq, lam, _ = oracle.constraint_dynamics(desc, q, v0, tau, KKT_CONS)
    g = oracle.rnea(desc, q, v0, v0)
    x = oracle.contact_cholesky_solve(desc, q, KKT_CONS, 0.0, np.vstack([np.zeros((9, 3)), tau - g]))
    np.testing.assert_allclose(x[:9], -lam, rtol=0, atol=1e-9 * np.abs(lam).max())
    np.testing.assert_allclose(x[9:], ddq, rtol=0, atol=1e-9 * np.abs(ddq).max())
