"""MJCF reading (pinocchio_amd/mjcf.py) against the reference's own known answers and
what the mathematics fixes (no GPU).

* tests/golden/simple_humanoid.xml is the reference's sample robot
  (models/simple_humanoid.xml): with a free-flyer root it must be the model the URDF
  reader builds from simple_humanoid.urdf (unittest/mjcf.cpp:1159-1247, compare_to_urdf),
  and nq = 29 with a fixed base (unittest/mjcf.cpp:1130-1139);
* the five orientation forms of unittest/mjcf.cpp:313-357 give one placement; the
  inertial conversions of unittest/mjcf.cpp:100-165 and 947-977; the joint kinds of
  parse_RX / parse_Free / parse_composite_RXRY (unittest/mjcf.cpp:554-742); armature with
  defaults (unittest/mjcf.cpp:979-1019, minus its ball joint, which the engine does not
  run); the default euler sequence (unittest/mjcf.cpp:1336-1365);
* a body with several joints (the reference's JointModelComposite, flattened here) moves
  as MuJoCo defines it: each hinge turns the body about its own anchor and axis, applied
  in order, checked through the oracle's forward kinematics at random angles.
"""
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from pinocchio_amd import mjcf, urdf
from pinocchio_amd.models import JT_FREEFLYER, JT_RX, JT_RY, JT_RZ

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XML = os.path.join(ROOT, "tests", "golden", "simple_humanoid.xml")
URDF = os.path.join(ROOT, "tests", "golden", "simple_humanoid.urdf")


def placement(desc, i):
    p = desc.placement[i]
    return p[:9].reshape(3, 3), p[9:]


def inertia(desc, i):
    m, cx, cy, cz, xx, xy, yy, xz, yz, zz = desc.inertia[i]
    return m, np.array([cx, cy, cz]), np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def test_simple_humanoid_matches_urdf():
    m = mjcf.parse(XML, root="freeflyer")
    u = urdf.parse(URDF, root="freeflyer")
    assert (m.nq, m.nv, m.njoints) == (u.nq, u.nv, u.njoints) == (36, 35, 31)
    assert np.array_equal(m.parents, u.parents) and np.array_equal(m.jtype, u.jtype)
    assert np.allclose(m.placement, u.placement, rtol=0, atol=1e-12)
    assert np.allclose(m.inertia, u.inertia, rtol=1e-9, atol=1e-12)
    assert np.array_equal(m.armature, u.armature)
    assert m.name == "simple_humanoid"


def test_simple_humanoid_fixed_base():
    m = mjcf.parse(XML)
    assert (m.nq, m.nv) == (29, 29)
    assert JT_FREEFLYER not in m.jtype


ORIENT = """<mujoco model="orient"><compiler angle="radian" eulerseq="xyz"/><worldbody><body name="base">
  <body name="q" pos="0.3 0.2 0.5" quat="1 -1 0 0"><joint type="hinge" axis="1 0 0"/></body>
  <body name="a" pos="0.3 0.2 0.5" axisangle="-1 0 0 1.5707963"><joint type="hinge" axis="1 0 0"/></body>
  <body name="e" pos="0.3 0.2 0.5" euler="-1.57079633 0 0"><joint type="hinge" axis="1 0 0"/></body>
  <body name="xy" pos="0.3 0.2 0.5" xyaxes="1 0 0 0 0 -1"><joint type="hinge" axis="1 0 0"/></body>
  <body name="z" pos="0.3 0.2 0.5" zaxis="0 1 0"><joint type="hinge" axis="1 0 0"/></body>
</body></worldbody></mujoco>"""


def test_orientation_forms():
    d = mjcf.parse(ORIENT)
    assert d.njoints == 6 and np.all(d.parents[1:] == 0)
    R_ref = np.array([[1.0, 0, 0], [0, 0, 1], [0, -1, 0]])
    for i in range(1, 6):
        R, t = placement(d, i)
        assert np.allclose(R, R_ref, atol=1e-7) and np.allclose(t, [0.3, 0.2, 0.5], atol=1e-12), i


def test_euler_sequences():
    """Rotation of `euler` is R_a1(e0) R_a2(e1) R_a3(e2) about the sequence's axes (moving
    axes), in degrees by default; the default sequence is xyz."""
    e = np.array([0.7, -0.4, 1.1])

    def body_R(compiler, angles):
        xml = (f'<mujoco model="e">{compiler}<worldbody><body name="b"><body name="c" euler="{angles}">'
               '<joint axis="0 1 0"/></body></body></worldbody></mujoco>')
        return placement(mjcf.parse(xml), 1)[0]

    deg = " ".join(str(x) for x in np.degrees(e))
    rad = " ".join(str(x) for x in e)
    assert np.allclose(body_R("", deg), Rotation.from_euler("XYZ", e).as_matrix(), atol=1e-12)
    assert np.allclose(body_R('<compiler angle="radian"/>', rad), Rotation.from_euler("XYZ", e).as_matrix(),
                       atol=1e-12)
    assert np.allclose(body_R('<compiler angle="radian" eulerseq="zyx"/>', rad),
                       Rotation.from_euler("ZYX", e).as_matrix(), atol=1e-12)
    # unittest/mjcf.cpp:1336-1365: euler="1.57 0 0" in radians, default sequence
    xml = ('<mujoco model="arm"><compiler angle="radian"/><worldbody><body name="base">'
           '<body name="body" pos="0 0 0.01" euler="1.57 0 0"><joint axis="1 0 0"/></body></body>'
           '</worldbody></mujoco>')
    R, t = placement(mjcf.parse(xml), 1)
    assert np.allclose(R, Rotation.from_rotvec([1.57, 0, 0]).as_matrix(), atol=1e-12)
    assert np.allclose(t, [0, 0, 0.01])
    with pytest.raises(ValueError):
        mjcf.parse('<mujoco model="e"><compiler eulerseq="xyw"/><worldbody><body/></worldbody></mujoco>')


@pytest.mark.parametrize("form", ["diaginertia", "fullinertia"])
def test_inertial(form):
    attr = ('diaginertia="0.00315 0.00388 0.004285"' if form == "diaginertia" else
            'fullinertia="0.00315 0.00388 0.004285 8.2904e-7 0.00015 8.2299e-6"')
    xml = ('<mujoco model="testJointInertia"><worldbody><body name="body1"><freejoint/>'
           f'<inertial mass="0.629769" pos="-0.041018 -0.00014 0.049974" {attr}/></body></worldbody></mujoco>')
    d = mjcf.parse(xml)
    assert d.njoints == 2 and d.jtype[1] == JT_FREEFLYER and d.nq == 7
    m, c, I = inertia(d, 1)
    I_ref = np.diag([0.00315, 0.00388, 0.004285])
    if form == "fullinertia":
        I_ref[0, 1] = I_ref[1, 0] = 8.2904e-7
        I_ref[0, 2] = I_ref[2, 0] = 0.00015
        I_ref[1, 2] = I_ref[2, 1] = 8.2299e-6
    assert m == 0.629769 and np.array_equal(c, [-0.041018, -0.00014, 0.049974])
    assert np.allclose(I, I_ref, rtol=0, atol=1e-15)
    assert np.all(d.inertia[0] == 0)


def test_inertial_frame_and_bounds():
    """A rotated inertial frame turns the principal inertia (R diag R^T); boundmass and
    boundinertia floor the mass and the diagonal."""
    xml = ('<mujoco model="b"><compiler angle="radian" boundmass="0.5" boundinertia="0.01"/><worldbody>'
           '<body name="r"><body name="c"><joint axis="0 0 1"/>'
           '<inertial mass="0.1" pos="1 2 3" quat="0.9238795 0 0 0.3826834" diaginertia="0.002 0.3 0.4"/>'
           '</body></body></worldbody></mujoco>')
    m, c, I = inertia(mjcf.parse(xml), 1)
    R = Rotation.from_quat([0, 0, 0.3826834, 0.9238795]).as_matrix()
    assert m == 0.5 and np.array_equal(c, [1, 2, 3])
    assert np.allclose(I, R @ np.diag([0.01, 0.3, 0.4]) @ R.T, atol=1e-9)


def test_joint_kinds():
    body = '<body name="link0"><body name="link1" pos="0 0 0">{}</body></body>'
    wrap = '<mujoco model="m"><worldbody>{}</worldbody></mujoco>'
    for axis, jt in (("1 0 0", JT_RX), ("0 2 0", JT_RY), ("0 0 1", JT_RZ)):
        d = mjcf.parse(wrap.format(body.format(f'<joint name="j" type="hinge" axis="{axis}"/>')))
        assert list(d.jtype[1:]) == [jt] and np.allclose(d.placement[1], np.r_[np.eye(3).ravel(), 0, 0, 0])
    d = mjcf.parse(wrap.format(body.format('<joint type="hinge"/>')))  # MuJoCo's default axis is z
    assert list(d.jtype[1:]) == [JT_RZ]
    # parse_composite_RXRY: two hinges in one body (flattened: a massless link in between)
    d = mjcf.parse(wrap.format(body.format('<joint type="hinge" axis="1 0 0"/><joint type="hinge" axis="0 1 0"/>')))
    assert list(d.jtype[1:]) == [JT_RX, JT_RY] and list(d.parents[1:]) == [0, 1] and d.nv == 2
    assert d.inertia[1][0] == 0
    # parse_Free: the fixed root link0 leaves link1's free joint as the first joint
    d = mjcf.parse(wrap.format(body.format('<freejoint name="joint1"/>')))
    assert list(d.jtype[1:]) == [JT_FREEFLYER] and d.nq == 7 and d.nv == 6
    with pytest.raises(NotImplementedError):  # a free joint below another joint
        mjcf.parse(wrap.format(body.format('<joint/><body name="x"><freejoint/></body>')))
    for bad in ('<joint type="slide" axis="1 0 0"/>', '<joint type="ball"/>', '<joint type="hinge" axis="1 1 0"/>'):
        with pytest.raises(NotImplementedError):
            mjcf.parse(wrap.format(body.format(bad)))
    with pytest.raises(NotImplementedError):  # inertia from geoms
        mjcf.parse(wrap.format(body.format('<joint/><geom type="sphere" size="0.1"/>')))


def test_armature_and_defaults():
    xml = """<mujoco model="model_RX">
      <default><joint armature="1" damping="1" limited="true"/>
        <default class="heavy"><joint armature="5"/><default class="heavier"><joint axis="0 1 0"/></default></default>
      </default>
      <worldbody><body name="link0"><body name="link1" pos="0 0 0">
        <joint name="joint1" type="hinge" axis="1 0 0" armature="1.3"/>
        <joint name="joint2" type="hinge" axis="0 1 0" armature="2.4"/>
        <joint name="joint3" type="hinge" axis="0 0 1" armature="0.4"/>
        <body pos=".2 0 0" name="body2"><joint type="hinge" axis="1 0 0"/>
          <body pos=".2 0 0" name="body3" childclass="heavy"><joint class="heavier"/>
            <body pos=".2 0 0" name="body4"><joint axis="0 0 1"/></body></body></body>
      </body></body></worldbody></mujoco>"""
    d = mjcf.parse(xml)
    assert np.array_equal(d.armature, [1.3, 2.4, 0.4, 1, 5, 5])
    assert list(d.jtype[1:]) == [JT_RX, JT_RY, JT_RZ, JT_RX, JT_RY, JT_RZ]
    assert list(d.parents[1:]) == [0, 1, 2, 3, 4, 5]


def test_jointless_bodies_merge_into_parent():
    """A body without joint is welded to its parent: its inertia joins the parent joint's
    body (parallel-axis sum) and its children hang off that joint."""
    xml = """<mujoco model="w"><compiler angle="radian"/><worldbody><body name="r"><body name="a" pos="0 0 1">
      <joint axis="1 0 0"/><inertial mass="2" pos="0 0 0" diaginertia="1 1 1"/>
      <body name="weld" pos="0 1 0" quat="0 0 0 1"><inertial mass="1" pos="0 0 0.5" diaginertia="0.1 0.2 0.3"/>
        <body name="b" pos="0 0 2"><joint axis="0 1 0"/></body></body></body></body></worldbody></mujoco>"""
    d = mjcf.parse(xml)
    assert list(d.parents) == [0, 0, 1]
    m, c, I = inertia(d, 1)
    # the welded body is turned by pi about z: its com is at (0, 1, 0.5) in a's frame
    c2 = np.array([0.0, 1.0, 0.5])
    assert m == 3 and np.allclose(c, (2 * 0 + c2) / 3)
    d1, d2 = -c, c2 - c
    pa = lambda mm, dd: mm * (dd @ dd * np.eye(3) - np.outer(dd, dd))
    assert np.allclose(I, np.eye(3) + pa(2, d1) + np.diag([0.1, 0.2, 0.3]) + pa(1, d2), atol=1e-12)
    R, t = placement(d, 2)
    assert np.allclose(R, np.diag([-1.0, -1.0, 1.0]), atol=1e-12) and np.allclose(t, [0, 1, 2])


def test_multi_joint_body_moves_as_mujoco_defines():
    """World pose of a body with two offset hinges: B T(p1) R1(q1) T(-p1) T(p2) R2(q2) T(-p2),
    with B the body pose and p_k the joint anchors (each hinge turns the body about its own
    anchor and axis, in order); compared through the oracle's forward kinematics."""
    B_q = np.array([0.9, 0.1, -0.3, 0.2])  # w x y z
    B_t = np.array([0.1, 0.2, 0.3])
    p1, p2 = np.array([0.05, 0.0, 0.02]), np.array([0.0, 0.04, -0.01])
    c, Id = np.array([0.01, 0.02, 0.03]), np.diag([0.1, 0.2, 0.3])
    xml = f"""<mujoco model="c"><worldbody><body name="base"><body name="b" pos="{' '.join(map(str, B_t))}"
      quat="{' '.join(map(str, B_q))}"><joint type="hinge" axis="1 0 0" pos="{' '.join(map(str, p1))}"/>
      <joint type="hinge" axis="0 0 1" pos="{' '.join(map(str, p2))}"/>
      <inertial mass="2" pos="{' '.join(map(str, c))}" diaginertia="0.1 0.2 0.3"/>
      <body name="tip" pos="0.3 0 0"><joint type="hinge" axis="0 1 0" pos="0.01 0 0"/></body>
    </body></body></worldbody></mujoco>"""
    d = mjcf.parse(xml)
    assert list(d.jtype[1:]) == [JT_RX, JT_RZ, JT_RY] and list(d.parents[1:]) == [0, 1, 2]
    Bq = B_q / np.linalg.norm(B_q)
    RB = Rotation.from_quat([Bq[1], Bq[2], Bq[3], Bq[0]]).as_matrix()

    def T(R=np.eye(3), t=np.zeros(3)):
        M = np.eye(4)
        M[:3, :3], M[:3, 3] = R, t
        return M

    rng = np.random.default_rng(3)
    q = rng.uniform(-np.pi, np.pi, size=(3, 5))
    oM, _, _ = oracle.forward_kinematics(d, q)
    for k in range(q.shape[1]):
        q1, q2, q3 = q[:, k]
        W = (T(RB, B_t) @ T(t=p1) @ T(Rotation.from_rotvec([q1, 0, 0]).as_matrix()) @ T(t=-p1)
             @ T(t=p2) @ T(Rotation.from_rotvec([0, 0, q2]).as_matrix()) @ T(t=-p2))
        # the body's com and rotational inertia in the world, from joint 2's frame and inertia
        M2 = T(oM[k, 1, :9].reshape(3, 3), oM[k, 1, 9:])
        m, cl, Il = inertia(d, 2)
        assert m == 2
        assert np.allclose((M2 @ np.r_[cl, 1])[:3], (W @ np.r_[c, 1])[:3], atol=1e-12)
        assert np.allclose(M2[:3, :3] @ Il @ M2[:3, :3].T, W[:3, :3] @ Id @ W[:3, :3].T, atol=1e-12)
        # the child body: anchored 0.01 along x in a frame 0.3 along x of b's body frame
        Wt = W @ T(t=[0.3, 0, 0]) @ T(t=[0.01, 0, 0]) @ T(Rotation.from_rotvec([0, q3, 0]).as_matrix())
        M3 = T(oM[k, 2, :9].reshape(3, 3), oM[k, 2, 9:])
        assert np.allclose(M3, Wt, atol=1e-12)


def test_root_handling_and_errors():
    xml = ('<mujoco model="r"><worldbody><body name="root"><inertial mass="3" pos="0 0 0" diaginertia="1 1 1"/>'
           '<body name="c" pos="1 0 0"><joint axis="0 1 0"/></body></body></worldbody></mujoco>')
    f = mjcf.parse(xml, root="freeflyer")
    assert list(f.jtype[1:]) == [JT_FREEFLYER, JT_RY] and f.inertia[1][0] == 3 and f.nq == 8
    x = mjcf.parse(xml)
    assert list(x.jtype[1:]) == [JT_RY] and x.nq == 1
    assert np.allclose(placement(x, 1)[1], [1, 0, 0])
    for bad in ('<robot name="x"/>', '<mujoco><worldbody><body/></worldbody></mujoco>',
                '<mujoco model="x"><worldbody/></mujoco>'):
        with pytest.raises(ValueError):
            mjcf.parse(bad)
    with pytest.raises(ValueError):
        mjcf.parse(xml, root="planar")
