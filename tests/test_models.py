"""CPU checks of the host-side model builders used by the run-time-topology tests and the
pinocchio-model path of the Python drop-in: random trees (compact or not, with or without
a free-flyer root), deep chains, conversion of a pinocchio.Model-like object, and the
oracle's identities on such trees (aba(rnea(a)) = a, aba.cpp:124-184 of the reference;
M from CRBA = brute-force M from RNEA differences, crba.cpp:104-114)."""
import numpy as np
import pytest

import oracle
from pinocchio_amd import models as M
from helpers import inputs


def dfs_order(parents):
    n = len(parents)
    kids = [[] for _ in range(n)]
    for j in range(1, n):
        kids[parents[j]].append(j)
    order, stack = [], list(reversed(kids[0]))
    while stack:
        j = stack.pop()
        order.append(j)
        stack.extend(reversed(kids[j]))
    return order


@pytest.mark.parametrize("ff", [True, False])
def test_random_tree_orders(ff):
    c = M.random_tree(30, 123, free_flyer=ff, compact=True)
    n = M.random_tree(30, 123, free_flyer=ff, compact=False)
    assert dfs_order(list(c.parents)) == list(range(1, c.njoints))
    assert dfs_order(list(n.parents)) != list(range(1, n.njoints))
    for d in (c, n):
        assert np.all(d.parents[1:] < np.arange(1, d.njoints))
        assert (d.jtype[1] == M.JT_FREEFLYER) == ff
        assert np.all(d.jtype[2 if ff else 1:] >= M.JT_RX)
        assert d.nv == 30 + (6 if ff else 0)
    # a fixed-base tree has several joints attached to the universe
    assert np.count_nonzero(c.parents[1:] == 0) >= (1 if ff else 2)


def test_serial_chain():
    d = M.serial_chain(40, 5)
    assert list(d.parents) == [0] + list(range(40))
    assert d.nq == d.nv == 40


class _FakeSE3:
    def __init__(self, R, t):
        self.rotation, self.translation = R, t


class _FakeInertia:
    def __init__(self, m, c, I):
        self.mass, self.lever, self.inertia = m, c, I


class _FakeJoint:
    def __init__(self, name):
        self._n = name

    def shortname(self):
        return self._n


class _FakeMotion:
    def __init__(self, l, a):
        self.linear, self.angular = l, a


class _FakePinModel:
    """The attributes of pinocchio.Model that models.from_pinocchio reads."""

    def __init__(self, d):
        names = {M.JT_FREEFLYER: "JointModelFreeFlyer", M.JT_RX: "JointModelRX", M.JT_RY: "JointModelRY",
                 M.JT_RZ: "JointModelRZ", M.JT_UNIVERSE: "JointModelUniverse"}
        self.name = d.name
        self.njoints = d.njoints
        self.parents = [int(p) for p in d.parents]
        self.joints = [_FakeJoint(names[int(t)]) for t in d.jtype]
        self.jointPlacements = [_FakeSE3(p[:9].reshape(3, 3), p[9:]) for p in d.placement]
        self.inertias = []
        for y in d.inertia:
            I = np.array([[y[4], y[5], y[7]], [y[5], y[6], y[8]], [y[7], y[8], y[9]]])
            self.inertias.append(_FakeInertia(y[0], y[1:4], I))
        self.armature = d.armature.copy()
        self.gravity = _FakeMotion(d.gravity[:3], d.gravity[3:])


def test_from_pinocchio_round_trip():
    d = M.humanoid36()
    d.armature = np.linspace(0.0, 0.35, d.nv)
    e = M.from_pinocchio(_FakePinModel(d))
    for f in ("parents", "jtype", "placement", "inertia", "armature", "gravity"):
        np.testing.assert_array_equal(getattr(e, f), getattr(d, f))


def test_from_pinocchio_rejects_other_joints():
    fake = _FakePinModel(M.arm7())
    fake.joints[3] = _FakeJoint("JointModelPrismaticUnaligned")
    with pytest.raises(NotImplementedError, match="PrismaticUnaligned"):
        M.from_pinocchio(fake)


@pytest.mark.parametrize("ff,compact", [(True, True), (True, False), (False, True)])
def test_oracle_identities_on_random_trees(ff, compact):
    d = M.random_tree(18, 321, free_flyer=ff, compact=compact)
    q, v, a, _ = inputs(d, 16, k=60)
    tau = oracle.rnea(d, q, v, a)
    np.testing.assert_allclose(oracle.aba(d, q, v, tau), a, rtol=1e-9, atol=1e-9)
    if compact:
        Mm = oracle.crba(d, q)
        nle = oracle.rnea(d, q, v, np.zeros_like(a))
        # brute force: column k of M = rnea(q, 0, e_k) - rnea(q, 0, 0)
        g = oracle.rnea(d, q, np.zeros_like(v), np.zeros_like(a))
        for k in range(d.nv):
            e = np.zeros_like(a)
            e[k] = 1.0
            col = oracle.rnea(d, q, np.zeros_like(v), e) - g
            np.testing.assert_allclose(Mm[:, :, k], col.T, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(np.einsum("bij,jb->ib", Mm, a) + nle, tau, rtol=1e-9, atol=1e-9)
