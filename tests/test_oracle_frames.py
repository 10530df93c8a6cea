"""Pins the oracle's frames (updateFramePlacements, frames.hxx:15-35; getFrameJacobian,
frames.hxx:156-175 with translateJointJacobian, jacobian.hxx:175-300) without the reference:

  * oMf = oMi[parent] * placement, composed in numpy from the oracle's forward kinematics;
  * LOCAL_WORLD_ALIGNED columns = central differences of the frame's origin and
    orientation along q (+) eps e_c (Lie-group integrate), zero outside the support;
  * WORLD columns = the parent joint's WORLD Jacobian columns on its support;
  * LOCAL = the LOCAL_WORLD_ALIGNED columns seen in the frame's axes (R^T on both halves).
"""
import numpy as np
import pytest

import oracle
from pinocchio_amd import models as M


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.build()
    oracle.lib()


def random_frames(desc, nf, seed):
    rng = np.random.default_rng(seed)
    par = rng.integers(0, desc.njoints, nf).astype(np.int32)
    par[0] = desc.njoints - 1
    par[1] = 0  # a frame on the universe
    pl = []
    for _ in range(nf):
        A = rng.standard_normal((3, 3))
        Q, _ = np.linalg.qr(A)
        Q *= np.sign(np.linalg.det(Q))
        pl.append(np.concatenate([Q.ravel(), rng.uniform(-0.3, 0.3, 3)]))
    return par, np.array(pl)


def support(desc, j):
    s = set()
    while j > 0:
        s |= set(range(desc.idx_v[j], desc.idx_v[j] + desc.nvs[j]))
        j = desc.parents[j]
    return s


@pytest.mark.parametrize("name", ["humanoid36", "quadruped18", "arm7"])
def test_placements_compose(name):
    d = getattr(M, name)()
    q, _, _, _ = M.sample_inputs(d, 4, 77)
    par, pl = random_frames(d, 6, 1)
    oMf, _ = oracle.frames(d, q, par, pl)
    oMi, _, _ = oracle.forward_kinematics(d, q)
    for b in range(4):
        for f in range(6):
            R, p = pl[f, :9].reshape(3, 3), pl[f, 9:]
            if par[f] > 0:
                Ri, pi = oMi[b, par[f] - 1, :9].reshape(3, 3), oMi[b, par[f] - 1, 9:]
                R, p = Ri @ R, Ri @ p + pi
            assert np.abs(oMf[b, f, :9] - R.ravel()).max() < 1e-13
            assert np.abs(oMf[b, f, 9:] - p).max() < 1e-13


@pytest.mark.parametrize("name", ["humanoid36", "quadruped18"])
def test_jacobians_finite_differences(name):
    d = getattr(M, name)()
    q, _, _, _ = M.sample_inputs(d, 1, 78)
    par, pl = random_frames(d, 4, 2)
    oMf, Jlwa = oracle.frames(d, q, par, pl, "local_world_aligned")
    _, Jw = oracle.frames(d, q, par, pl, "world")
    _, Jl = oracle.frames(d, q, par, pl, "local")
    Jj = oracle.joint_jacobians(d, q)[0]
    eps = 1e-6
    worst = 0.0
    for c in range(d.nv):
        e = np.zeros((d.nv, 1))
        e[c] = eps
        Mp, _ = oracle.frames(d, oracle.integrate(d, q, e), par, pl)
        Mm, _ = oracle.frames(d, oracle.integrate(d, q, -e), par, pl)
        for f in range(len(par)):
            R0 = oMf[0, f, :9].reshape(3, 3)
            dR = (Mp[0, f, :9] - Mm[0, f, :9]).reshape(3, 3) / (2 * eps)
            dp = (Mp[0, f, 9:] - Mm[0, f, 9:]) / (2 * eps)
            W = dR @ R0.T
            fd = np.concatenate([dp, [W[2, 1], W[0, 2], W[1, 0]]])
            worst = max(worst, np.abs(fd - Jlwa[0, f, :, c]).max())
    assert worst < 1e-7
    for f in range(len(par)):
        sup = support(d, par[f])
        mask = np.array([c in sup for c in range(d.nv)], dtype=float)
        assert np.abs(Jw[0, f] - Jj * mask).max() < 1e-13
        R = oMf[0, f, :9].reshape(3, 3)
        assert np.abs(Jl[0, f, :3] - R.T @ Jlwa[0, f, :3]).max() < 1e-12
        assert np.abs(Jl[0, f, 3:] - R.T @ Jlwa[0, f, 3:]).max() < 1e-12
