"""The build's topology generator: a user's URDF / MJCF robots join the prebuilt list
(`make -C pinocchio_amd/csrc ROBOTS=...`).  Generated into a temporary folder; nothing is
compiled here."""
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "pinocchio_amd", "csrc"))
import gen_topologies as G  # noqa: E402

ARM = """<robot name="arm3">
  <link name="base"/>
  <link name="l1"><inertial><mass value="1"/><inertia ixx="0.1" iyy="0.1" izz="0.1" ixy="0" ixz="0" iyz="0"/></inertial></link>
  <link name="l2"><inertial><mass value="1"/><inertia ixx="0.1" iyy="0.1" izz="0.1" ixy="0" ixz="0" iyz="0"/></inertial></link>
  <link name="l3"><inertial><mass value="1"/><inertia ixx="0.1" iyy="0.1" izz="0.1" ixy="0" ixz="0" iyz="0"/></inertial></link>
  <joint name="j1" type="revolute"><parent link="base"/><child link="l1"/><axis xyz="0 0 1"/></joint>
  <joint name="j2" type="revolute"><parent link="l1"/><child link="l2"/><axis xyz="0 1 0"/>
    <origin xyz="0 0 0.3"/></joint>
  <joint name="j3" type="revolute"><parent link="l1"/><child link="l3"/><axis xyz="1 0 0"/>
    <origin xyz="0 0.1 0.3"/></joint>
</robot>
"""


def test_depth_first():
    assert G.depth_first([0, 0, 1, 2, 1, 4])
    assert G.depth_first([0, 0, 0, 2])
    assert not G.depth_first([0, 0, 1, 1, 2])  # joint 4's parent 2 is not on 3's path


def test_user_robots_join_the_list(tmp_path):
    arm = tmp_path / "my-arm.urdf"
    arm.write_text(ARM)
    human = os.path.join(HERE, "golden", "simple_humanoid.urdf")
    out = tmp_path / "gen"
    units = G.main(str(out), f"{arm},{arm}:freeflyer,{human}:freeflyer")
    inc = (out / "topologies.inc").read_text()
    reg = (out / "registry.inc").read_text()
    for name in ("robot_my_arm", "robot_my_arm_ff"):
        assert f"struct Topo_{name} " in inc
        assert f"RBD_REGISTER({name})" in reg
        assert f"inst_{name}_dyn" in units and f"inst_{name}_kkt" in units  # depth-first: all families
    # the sample humanoid with a free flyer is prebuilt already: no duplicate
    assert "robot_simple_humanoid_ff" not in inc
    # the listed trees are all still there
    for name, _, _ in G.TOPOLOGIES:
        assert f"struct Topo_{name} " in inc
    with pytest.raises(ValueError):
        G.main(str(out), f"{arm}:floating")
