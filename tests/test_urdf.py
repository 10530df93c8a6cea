"""URDF reading (pinocchio_amd/urdf.py) against the reference's own known answers and
closed forms (no GPU).

* tests/golden/simple_humanoid.urdf is the reference's sample robot
  (models/simple_humanoid.urdf); its unit test expects nq = 29 with a fixed base and
  nq = 36 with a free-flyer root (unittest/urdf.cpp:60, 85);
* roll-pitch-yaw against scipy's extrinsic x-y-z rotation; inertial-frame rotation and
  fixed-joint merging against hand-computed parallel-axis sums;
* the parsed tree is the prebuilt simple_humanoid topology.
"""
import os
import sys

import numpy as np
import pytest

from pinocchio_amd import urdf
from pinocchio_amd.models import JT_FREEFLYER

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
URDF = os.path.join(ROOT, "tests", "golden", "simple_humanoid.urdf")


def test_simple_humanoid_dimensions_match_reference():
    m = urdf.parse(URDF)
    assert (m.nq, m.nv, m.njoints) == (29, 29, 30)
    f = urdf.parse(URDF, root="freeflyer")
    assert (f.nq, f.nv, f.njoints) == (36, 35, 31)
    assert f.jtype[1] == JT_FREEFLYER and f.inertia[1][0] == 27.0  # the root link (WAIST_LINK0) rides the free flyer
    assert np.all(f.parents[1:] < np.arange(1, f.njoints))


def test_parsed_tree_is_the_prebuilt_topology():
    sys.path.insert(0, os.path.join(ROOT, "pinocchio_amd", "csrc"))
    import gen_topologies as g
    f = urdf.parse(URDF, root="freeflyer")
    assert list(map(int, f.parents)) == g.SIMPLE_HUMANOID_PARENTS
    assert list(map(int, f.jtype)) == g.SIMPLE_HUMANOID_TYPES


def test_rpy_matches_extrinsic_xyz():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(0)
    for _ in range(20):
        r, p, y = rng.uniform(-3, 3, 3)
        assert np.allclose(urdf.rpy_to_rot(r, p, y), Rotation.from_euler("xyz", [r, p, y]).as_matrix(), atol=1e-14)


XML = """<robot name="t">
  <link name="base"/>
  <link name="a"><inertial><origin xyz="0.1 0 0" rpy="0 0 1.5707963267948966"/><mass value="2"/>
    <inertia ixx="1" ixy="0" ixz="0" iyy="3" iyz="0" izz="5"/></inertial></link>
  <link name="b"><inertial><origin xyz="0 0 0.5"/><mass value="1"/>
    <inertia ixx="0.1" ixy="0" ixz="0" iyy="0.1" iyz="0" izz="0.1"/></inertial></link>
  <link name="c"><inertial><mass value="4"/><inertia ixx="1" ixy="0" ixz="0" iyy="1" iyz="0" izz="1"/></inertial></link>
  <joint name="j1" type="revolute"><parent link="base"/><child link="a"/><origin xyz="0 0 1"/><axis xyz="0 0 1"/>
    <limit lower="-1" upper="1" effort="1" velocity="1"/></joint>
  <joint name="j2" type="fixed"><parent link="a"/><child link="b"/><origin xyz="0 0 0.2"/></joint>
  <joint name="j3" type="revolute"><parent link="b"/><child link="c"/><origin xyz="0 1 0" rpy="0 0 0"/>
    <axis xyz="1 0 0"/><limit lower="-1" upper="1" effort="1" velocity="1"/></joint>
</robot>"""


def test_inertial_rotation_fixed_merge_and_placement():
    m = urdf.parse(XML)
    assert m.njoints == 3 and list(m.parents) == [0, 0, 1]
    # body of joint 1 = link a (rotated inertia) + link b merged through the fixed joint
    ma, ca = 2.0, np.array([0.1, 0, 0])
    Ia = np.diag([3.0, 1.0, 5.0])  # diag(1, 3, 5) turned 90 deg about z
    mb, cb, Ib = 1.0, np.array([0, 0, 0.7]), 0.1 * np.eye(3)
    mt = ma + mb
    c = (ma * ca + mb * cb) / mt

    def shift(mi, ci, Ii):
        d = ci - c
        return Ii + mi * (d @ d * np.eye(3) - np.outer(d, d))

    I = shift(ma, ca, Ia) + shift(mb, cb, Ib)
    got = m.inertia[1]
    assert got[0] == mt and np.allclose(got[1:4], c)
    assert np.allclose(got[4:], [I[0, 0], I[0, 1], I[1, 1], I[0, 2], I[1, 2], I[2, 2]], atol=1e-12)
    # joint 2 hangs off joint 1 with the placement composed through the fixed joint
    assert np.allclose(m.placement[2][9:], [0, 1, 0.2]) and np.allclose(m.placement[2][:9], np.eye(3).ravel())


@pytest.mark.parametrize("bad", [
    XML.replace('type="fixed"', 'type="prismatic"'),
    XML.replace('<axis xyz="1 0 0"/>', '<axis xyz="0 -1 0"/>'),
    XML.replace('type="revolute"><parent link="base"/>', 'type="continuous"><parent link="base"/>'),
])
def test_unsupported_joints_are_rejected(bad):
    with pytest.raises(NotImplementedError):
        urdf.parse(bad)
