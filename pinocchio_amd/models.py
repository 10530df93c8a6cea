"""Seeded sample models and synthetic batch inputs.

The builders play the role of the reference's sample models
(multibody/sample-models.hxx:15-54 ``addJointAndBody``, 236-308 ``humanoidRandom``), but
with the robot families and parameter distributions of the benchmark workloads:

* joint placement: translation U(-0.3, 0.3)^3, uniformly random rotation;
* body inertia: mass U(0.5, 5) kg, centre of mass U(-0.1, 0.1)^3 m, rotational inertia
  R diag(l) R^T with eigenvalues l ~ U(0.01, 0.1) kg m^2;
* armature 0, gravity (0, 0, -9.81).

Joints are added in depth-first order, so every model except the random tree is
"compact" in the sense of crba.hxx:573-595.  Randomness comes from a counter-based
splitmix64 stream so every model and input batch is reproducible from its seed alone.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

JT_UNIVERSE, JT_FREEFLYER, JT_RX, JT_RY, JT_RZ = 0, 1, 2, 3, 4
AXIS = {"X": JT_RX, "Y": JT_RY, "Z": JT_RZ}

MODEL_SEED_BASE = 0xB0D10000
INPUT_SEED_BASE = 0x5EED0000

_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of the splitmix64 generator started at `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        ctr = np.uint64(seed & _M64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = ctr
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential draws from a splitmix64 counter stream."""

    def __init__(self, seed: int):
        self.seed = seed & _M64
        self.pos = 0

    def bits(self, n: int) -> np.ndarray:
        out = splitmix64((self.seed + self.pos * 0x9E3779B97F4A7C15) & _M64, n)
        self.pos += n
        return out

    def uniform(self, lo, hi, shape) -> np.ndarray:
        n = int(np.prod(shape))
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        return (lo + (hi - lo) * u).reshape(shape)

    def normal(self, shape, sigma=1.0) -> np.ndarray:
        n = int(np.prod(shape))
        m = (n + 1) // 2
        u1 = self.uniform(0.0, 1.0, (m,))
        u2 = self.uniform(0.0, 1.0, (m,))
        r = np.sqrt(-2.0 * np.log(1.0 - u1))  # 1-u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return sigma * z.reshape(shape)

    def quaternion(self, n: int) -> np.ndarray:
        """Uniform unit quaternions [x, y, z, w] (Shoemake, as math/quaternion.hpp:115-137)."""
        u = self.uniform(0.0, 1.0, (3, n))
        s1, s2 = np.sqrt(1 - u[0]), np.sqrt(u[0])
        t1, t2 = 2 * np.pi * u[1], 2 * np.pi * u[2]
        return np.stack([s1 * np.sin(t1), s1 * np.cos(t1), s2 * np.sin(t2), s2 * np.cos(t2)], axis=0)


def quat_to_rot(qxyzw: np.ndarray) -> np.ndarray:
    x, y, z, w = qxyzw
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


class _Desc(ctypes.Structure):
    """C layout of the flattened model descriptor (matches `rbd_model_desc` in include/rbd.h
    and the oracle's ModelDesc)."""

    _fields_ = [
        ("njoints", ctypes.c_int32),
        ("parents", ctypes.POINTER(ctypes.c_int32)),
        ("jtype", ctypes.POINTER(ctypes.c_int32)),
        ("placement", ctypes.POINTER(ctypes.c_double)),
        ("inertia", ctypes.POINTER(ctypes.c_double)),
        ("armature", ctypes.POINTER(ctypes.c_double)),
        ("gravity", ctypes.POINTER(ctypes.c_double)),
    ]


@dataclass
class ModelDesc:
    """Flattened rigid-body tree: joint 0 is the universe; parents[i] < i."""

    name: str
    parents: np.ndarray  # int32 [njoints]
    jtype: np.ndarray  # int32 [njoints]
    placement: np.ndarray  # float64 [njoints, 12]: R row-major then translation
    inertia: np.ndarray  # float64 [njoints, 10]: m, c(3), I_c (xx, xy, yy, xz, yz, zz)
    armature: np.ndarray = None  # float64 [nv]
    gravity: np.ndarray = field(default_factory=lambda: np.array([0, 0, -9.81, 0, 0, 0], dtype=np.float64))

    def __post_init__(self):
        self.parents = np.ascontiguousarray(self.parents, dtype=np.int32)
        self.jtype = np.ascontiguousarray(self.jtype, dtype=np.int32)
        self.placement = np.ascontiguousarray(self.placement, dtype=np.float64)
        self.inertia = np.ascontiguousarray(self.inertia, dtype=np.float64)
        if self.armature is None:
            self.armature = np.zeros(self.nv, dtype=np.float64)
        self.armature = np.ascontiguousarray(self.armature, dtype=np.float64)
        self.gravity = np.ascontiguousarray(self.gravity, dtype=np.float64)

    @property
    def njoints(self) -> int:
        return int(len(self.parents))

    @property
    def nqs(self):
        return np.array([0] + [7 if t == JT_FREEFLYER else 1 for t in self.jtype[1:]], dtype=np.int64)

    @property
    def nvs(self):
        return np.array([0] + [6 if t == JT_FREEFLYER else 1 for t in self.jtype[1:]], dtype=np.int64)

    @property
    def nq(self) -> int:
        return int(self.nqs.sum())

    @property
    def nv(self) -> int:
        return int(self.nvs.sum())

    @property
    def idx_q(self):
        return np.concatenate([[0], np.cumsum(self.nqs)[:-1]]).astype(np.int64)

    @property
    def idx_v(self):
        return np.concatenate([[0], np.cumsum(self.nvs)[:-1]]).astype(np.int64)

    def cstruct(self) -> _Desc:
        """ctypes descriptor pointing at this model's arrays (keep `self` alive while used)."""
        d = _Desc()
        d.njoints = self.njoints
        d.parents = self.parents.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.jtype = self.jtype.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.placement = self.placement.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.inertia = self.inertia.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.armature = self.armature.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.gravity = self.gravity.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return d

    def has_freeflyer(self) -> bool:
        return bool(np.any(self.jtype == JT_FREEFLYER))


class _Builder:
    def __init__(self, name: str, seed: int):
        self.name = name
        self.rng = Stream(seed)
        self.parents = [0]
        self.jtype = [JT_UNIVERSE]
        self.placement = [np.concatenate([np.eye(3).ravel(), np.zeros(3)])]
        self.inertia = [np.zeros(10)]

    def random_placement(self):
        R = quat_to_rot(self.rng.quaternion(1)[:, 0])
        t = self.rng.uniform(-0.3, 0.3, (3,))
        return np.concatenate([R.ravel(), t])

    def random_inertia(self):
        m = self.rng.uniform(0.5, 5.0, (1,))[0]
        c = self.rng.uniform(-0.1, 0.1, (3,))
        lam = self.rng.uniform(0.01, 0.1, (3,))
        Rr = quat_to_rot(self.rng.quaternion(1)[:, 0])
        Ic = Rr @ np.diag(lam) @ Rr.T
        return np.array([m, *c, Ic[0, 0], Ic[0, 1], Ic[1, 1], Ic[0, 2], Ic[1, 2], Ic[2, 2]])

    def add(self, parent: int, jtype: int, placement=None) -> int:
        self.parents.append(parent)
        self.jtype.append(jtype)
        self.placement.append(self.random_placement() if placement is None else placement)
        self.inertia.append(self.random_inertia())
        return len(self.parents) - 1

    def chain(self, parent: int, axes: str) -> int:
        j = parent
        for a in axes:
            j = self.add(j, AXIS[a])
        return j

    def free_flyer(self) -> int:
        return self.add(0, JT_FREEFLYER, np.concatenate([np.eye(3).ravel(), np.zeros(3)]))

    def build(self) -> ModelDesc:
        return ModelDesc(self.name, np.array(self.parents), np.array(self.jtype), np.array(self.placement),
                         np.array(self.inertia))


def manipulator6(seed: int = MODEL_SEED_BASE + 0) -> ModelDesc:
    """6 revolute joints, axes alternating Z/Y (6/6)."""
    b = _Builder("manipulator6", seed)
    b.chain(0, "ZYZYZY")
    return b.build()


def arm7(seed: int = MODEL_SEED_BASE + 3) -> ModelDesc:
    """7-DoF serial arm, axes Z Y Z Y Z Y Z (7/7)."""
    b = _Builder("arm7", seed)
    b.chain(0, "ZYZYZYZ")
    return b.build()


def quadruped18(seed: int = MODEL_SEED_BASE + 4) -> ModelDesc:
    """Free-flyer base plus 4 legs of (X, Y, Y) revolute joints (19/18)."""
    b = _Builder("quadruped18", seed)
    base = b.free_flyer()
    for _ in range(4):
        b.chain(base, "XYY")
    return b.build()


def humanoid36(seed: int = MODEL_SEED_BASE + 1) -> ModelDesc:
    """Free-flyer base, two 6R legs, 2R torso carrying two 7R arms and a 2R head (37/36)."""
    b = _Builder("humanoid36", seed)
    base = b.free_flyer()
    b.chain(base, "ZXYYYX")  # left leg
    b.chain(base, "ZXYYYX")  # right leg
    torso = b.chain(base, "ZY")
    b.chain(torso, "YXZYZYX")  # left arm
    b.chain(torso, "YXZYZYX")  # right arm
    b.chain(torso, "ZY")  # head
    return b.build()


def random60(seed: int = MODEL_SEED_BASE + 5) -> ModelDesc:
    """Free-flyer base plus 54 revolute joints; the parent of joint i is uniform over the
    earlier moving joints, the axis uniform over X/Y/Z (61/60).  Not DFS-ordered."""
    b = _Builder("random60", seed)
    b.free_flyer()
    for i in range(2, 56):
        parent = 1 + int(b.rng.uniform(0.0, 1.0, (1,))[0] * (i - 1))
        parent = min(parent, i - 1)
        axis = "XYZ"[min(2, int(b.rng.uniform(0.0, 3.0, (1,))[0]))]
        b.add(parent, AXIS[axis])
    return b.build()


def pendulum(mass=1.3, length=0.7, axis="Y") -> ModelDesc:
    """Single revolute joint at the origin with a point-like bob at `length` along -z."""
    inertia = np.zeros((2, 10))
    inertia[1, 0] = mass
    inertia[1, 1:4] = [0.0, 0.0, -length]
    I = np.eye(4)[:3, :3] * 0.0
    placement = np.tile(np.concatenate([np.eye(3).ravel(), np.zeros(3)]), (2, 1))
    return ModelDesc("pendulum", np.array([0, 0]), np.array([0, AXIS[axis]]), placement, inertia)


def free_body(seed: int = 7) -> ModelDesc:
    """A single rigid body on a free-flyer joint."""
    b = _Builder("free_body", seed)
    b.free_flyer()
    return b.build()


def random_tree(n_revolute: int, seed: int, free_flyer: bool = True, compact: bool = True,
                name: str = None) -> ModelDesc:
    """Random tree of `n_revolute` revolute joints (axes uniform over X/Y/Z) on an optional
    free-flyer root, with the BASELINE parameter distributions.  The parent of each new
    joint is uniform over the joints added before it.  compact=True renumbers the joints
    in depth-first pre-order (children in increasing order), as addJoint callers do;
    compact=False keeps the insertion order, which is not DFS-ordered in general."""
    b = _Builder(name or f"tree{n_revolute}_{seed & 0xffff:x}", seed)
    parents, types = [0], [JT_UNIVERSE]
    if free_flyer:
        parents.append(0)
        types.append(JT_FREEFLYER)
    first = len(parents)
    for i in range(n_revolute):
        j = first + i
        lo = 1 if free_flyer else 0
        if j == lo:
            p = 0
        else:
            p = lo + int(b.rng.uniform(0.0, 1.0, (1,))[0] * (j - lo))
            p = min(p, j - 1)
        parents.append(p)
        types.append(AXIS["XYZ"[min(2, int(b.rng.uniform(0.0, 3.0, (1,))[0]))]])
    nj = len(parents)
    order = list(range(nj))
    if compact:
        kids = [[] for _ in range(nj)]
        for j in range(1, nj):
            kids[parents[j]].append(j)
        order, stack = [0], list(reversed(kids[0]))
        while stack:
            j = stack.pop()
            order.append(j)
            stack.extend(reversed(kids[j]))
    new_index = {old: k for k, old in enumerate(order)}
    for old in order[1:]:
        if types[old] == JT_FREEFLYER:
            b.free_flyer()
        else:
            b.add(new_index[parents[old]], types[old])
    return b.build()


def serial_chain(n: int, seed: int, axes: str = "XYZ") -> ModelDesc:
    """n revolute joints in one chain, axes cycling through `axes` (a deep tree)."""
    b = _Builder(f"chain{n}", seed)
    b.chain(0, (axes * n)[:n])
    return b.build()


def from_pinocchio(model) -> ModelDesc:
    """Flatten a pinocchio Model (the Python binding's pinocchio.Model, or any object with
    the same attributes: njoints, parents, joints[i].shortname(), jointPlacements[i]
    .rotation / .translation, inertias[i].mass / .lever / .inertia, armature, gravity
    .linear / .angular) into a ModelDesc.  Joint types FreeFlyer (at a root) and
    RX / RY / RZ only, as the engine supports."""
    kinds = {"JointModelFreeFlyer": JT_FREEFLYER, "JointModelRX": JT_RX, "JointModelRY": JT_RY,
             "JointModelRZ": JT_RZ}
    nj = int(model.njoints)
    parents = np.array([int(p) for p in list(model.parents)[:nj]], dtype=np.int32)
    jtype = np.zeros(nj, dtype=np.int32)
    placement = np.zeros((nj, 12))
    inertia = np.zeros((nj, 10))
    placement[0, :9] = np.eye(3).ravel()
    for j in range(1, nj):
        kind = model.joints[j].shortname()
        if kind not in kinds:
            raise NotImplementedError(f"joint {j}: {kind} is not supported (FreeFlyer, RX, RY, RZ)")
        jtype[j] = kinds[kind]
        M = model.jointPlacements[j]
        placement[j, :9] = np.asarray(M.rotation, dtype=np.float64).reshape(3, 3).ravel()
        placement[j, 9:] = np.asarray(M.translation, dtype=np.float64).reshape(3)
        Y = model.inertias[j]
        Ic = np.asarray(Y.inertia, dtype=np.float64).reshape(3, 3)
        inertia[j] = [float(Y.mass), *np.asarray(Y.lever, dtype=np.float64).reshape(3), Ic[0, 0], Ic[0, 1], Ic[1, 1],
                      Ic[0, 2], Ic[1, 2], Ic[2, 2]]
    g = model.gravity
    gravity = np.concatenate([np.asarray(g.linear, dtype=np.float64).reshape(3),
                              np.asarray(g.angular, dtype=np.float64).reshape(3)])
    return ModelDesc(getattr(model, "name", "") or "model", parents, jtype, placement, inertia,
                     np.asarray(model.armature, dtype=np.float64).reshape(-1), gravity)


SAMPLE_MODELS = {
    "manipulator6": manipulator6,
    "arm7": arm7,
    "quadruped18": quadruped18,
    "humanoid36": humanoid36,
    "random60": random60,
}


def sample_inputs(model: ModelDesc, batch: int, seed: int, tau_sigma: float = 1.0, base_range: float = 1.0):
    """Column-major batch inputs (q: nq x B, v, a, tau: nv x B), float64.

    Revolute q ~ U(-pi, pi); free-flyer translation U(-1, 1)^3 and a uniform unit
    quaternion; v, a ~ N(0, 1); tau ~ N(0, tau_sigma)."""
    rng = Stream(seed)
    q = np.zeros((model.nq, batch))
    iq = model.idx_q
    for j in range(1, model.njoints):
        if model.jtype[j] == JT_FREEFLYER:
            q[iq[j]:iq[j] + 3] = rng.uniform(-base_range, base_range, (3, batch))
            q[iq[j] + 3:iq[j] + 7] = rng.quaternion(batch)
        else:
            q[iq[j]] = rng.uniform(-np.pi, np.pi, (batch,))
    v = rng.normal((model.nv, batch))
    a = rng.normal((model.nv, batch))
    tau = rng.normal((model.nv, batch), sigma=tau_sigma)
    return q, v, a, tau


def perturb_configurations(model: ModelDesc, q0, batch: int, sigma: float, seed: int):
    """nq x B configurations scattered around q0: revolute angles q0 + sigma N(0, 1),
    free-flyer translation q0 + sigma N(0, 1) and orientation q0 rotated by a rotation
    vector sigma N(0, 1) (unit quaternions, [x, y, z, w]).  For workloads whose constraints
    hold near one reference posture (contacts, loop closures)."""
    rng = Stream(seed)
    q0 = np.asarray(q0, dtype=np.float64).reshape(model.nq)
    q = np.repeat(q0[:, None], batch, axis=1)
    iq = model.idx_q
    for j in range(1, model.njoints):
        if model.jtype[j] == JT_FREEFLYER:
            q[iq[j]:iq[j] + 3] += sigma * rng.normal((3, batch))
            w = sigma * rng.normal((3, batch))
            ang = np.linalg.norm(w, axis=0)
            axis = w / np.maximum(ang, 1e-300)
            dq = np.concatenate([axis * np.sin(ang / 2), np.cos(ang / 2)[None]], axis=0)  # [x, y, z, w]
            x0, y0, z0, w0 = q0[iq[j] + 3:iq[j] + 7]
            x1, y1, z1, w1 = dq
            q[iq[j] + 3:iq[j] + 7] = np.stack([w0 * x1 + x0 * w1 + y0 * z1 - z0 * y1,
                                              w0 * y1 - x0 * z1 + y0 * w1 + z0 * x1,
                                              w0 * z1 + x0 * y1 - y0 * x1 + z0 * w1,
                                              w0 * w1 - x0 * x1 - y0 * y1 - z0 * z1])
        else:
            q[iq[j]] += sigma * rng.normal((batch,))
    return q

