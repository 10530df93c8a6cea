"""URDF -> flattened model (ModelDesc) for the engine.

Follows what pinocchio's URDF reader builds (src/parsers/urdf/model.cpp:67-315,
parseTree / buildModel), restricted to the joints the engine runs:

* the tree is walked depth first from the root link; a link's children are visited in
  the order of their joints' names (the URDF DOM keeps joints in a name-ordered map);
* a ``revolute`` joint about the +x, +y or +z axis of its frame becomes RX / RY / RZ
  (other axes would need pinocchio's RevoluteUnaligned joint and are rejected), with the
  joint placement = the joint origin (xyz, then rpy as R = Rz(yaw) Ry(pitch) Rx(roll));
* a ``fixed`` joint adds no degree of freedom: its link's inertia is merged into the
  parent body and its children hang off the parent joint with the composed placement;
* the body inertia of a link is (mass, lever = inertial origin, R I R^T with R the
  inertial origin's rotation) (model.cpp:30-42);
* ``root="freeflyer"`` puts a free-flyer joint carrying the root link's inertia at the
  root (model.cpp:310, addRootJoint); otherwise the root link is fixed to the universe and
  its inertia plays no role.
``continuous``, ``prismatic``, ``planar`` and ``floating`` joints are rejected
(NotImplementedError), as are mimic joints.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from .models import JT_FREEFLYER, JT_RX, JT_RY, JT_RZ, JT_UNIVERSE, ModelDesc


def rpy_to_rot(r, p, y):
    """R = Rz(y) Ry(p) Rx(r) (URDF roll-pitch-yaw)."""
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    return np.array([[cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
                     [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
                     [-sp, cp * sr, cp * cr]])


def _vec(s, default):
    return np.array([float(x) for x in s.split()]) if s is not None else np.array(default, dtype=float)


def _origin(el):
    """(R, t) of an <origin> child (identity if absent)."""
    o = el.find("origin") if el is not None else None
    if o is None:
        return np.eye(3), np.zeros(3)
    t = _vec(o.get("xyz"), [0, 0, 0])
    rpy = _vec(o.get("rpy"), [0, 0, 0])
    return rpy_to_rot(*rpy), t


def _inertia(link):
    """(m, c, I_c as a 3x3 matrix) of a link, zero if it has no <inertial>."""
    inert = link.find("inertial")
    if inert is None:
        return 0.0, np.zeros(3), np.zeros((3, 3))
    R, c = _origin(inert)
    m = float(inert.find("mass").get("value"))
    i = inert.find("inertia")
    g = {k: float(i.get(k, 0.0)) for k in ("ixx", "ixy", "ixz", "iyy", "iyz", "izz")}
    I = np.array([[g["ixx"], g["ixy"], g["ixz"]], [g["ixy"], g["iyy"], g["iyz"]], [g["ixz"], g["iyz"], g["izz"]]])
    return m, c, R @ I @ R.T


def _merge(a, b_in_a):
    """Sum of two inertias (m, c, I_c) expressed in the same frame (parallel-axis)."""
    (m1, c1, I1), (m2, c2, I2) = a, b_in_a
    m = m1 + m2
    if m <= 0:
        return 0.0, np.zeros(3), I1 + I2
    c = (m1 * c1 + m2 * c2) / m

    def shift(mi, ci, Ii):
        d = ci - c
        return Ii + mi * (np.dot(d, d) * np.eye(3) - np.outer(d, d))

    return m, c, shift(m1, c1, I1) + shift(m2, c2, I2)


def _transform(inertia, R, t):
    """An inertia (m, c, I_c) given in a child frame, expressed in the parent frame."""
    m, c, I = inertia
    return m, R @ c + t, R @ I @ R.T


def parse(path_or_xml: str, root: str = "fixed", name: str = None) -> ModelDesc:
    """Flattened model of a URDF file (or XML string).  root: "fixed" or "freeflyer"."""
    text = path_or_xml
    if not path_or_xml.lstrip().startswith("<"):
        with open(path_or_xml) as f:
            text = f.read()
    robot = ET.fromstring(text)
    links = {l.get("name"): l for l in robot.findall("link")}
    joints = robot.findall("joint")
    children = {}
    parent_of = {}
    for j in joints:
        p, c = j.find("parent").get("link"), j.find("child").get("link")
        children.setdefault(p, []).append(j)
        parent_of[c] = j
    roots = [l for l in links if l not in parent_of]
    if len(roots) != 1:
        raise ValueError(f"expected one root link, found {roots}")
    for k in children:
        children[k].sort(key=lambda j: j.get("name"))

    parents, jtype, placement, body = [0], [JT_UNIVERSE], [np.concatenate([np.eye(3).ravel(), np.zeros(3)])], [None]
    body[0] = (0.0, np.zeros(3), np.zeros((3, 3)))

    def add(parent_idx, jt, R, t, inertia):
        parents.append(parent_idx)
        jtype.append(jt)
        placement.append(np.concatenate([R.ravel(), t]))
        body.append(inertia)
        return len(parents) - 1

    root_link = roots[0]
    if root == "freeflyer":
        top = add(0, JT_FREEFLYER, np.eye(3), np.zeros(3), _inertia(links[root_link]))
    elif root == "fixed":
        top = 0
    else:
        raise ValueError("root must be 'fixed' or 'freeflyer'")

    # depth-first walk; (link, joint index its body belongs to, placement of the link
    # frame in that joint's frame) -- the placement is not the identity below fixed joints
    def walk(link_name, owner, R_in, t_in):
        for j in children.get(link_name, []):
            kind = j.get("type")
            child = j.find("child").get("link")
            if j.find("mimic") is not None:
                raise NotImplementedError(f"mimic joint {j.get('name')}")
            Rj, tj = _origin(j)
            R = R_in @ Rj
            t = R_in @ tj + t_in
            if kind == "fixed":
                # (merging into the universe, owner 0, is harmless: it has no dynamics)
                body[owner] = _merge(body[owner], _transform(_inertia(links[child]), R, t))
                walk(child, owner, R, t)
            elif kind == "revolute":
                ax = _vec(j.find("axis").get("xyz") if j.find("axis") is not None else None, [1, 0, 0])
                table = {(1, 0, 0): JT_RX, (0, 1, 0): JT_RY, (0, 0, 1): JT_RZ}
                key = tuple(int(round(x)) for x in ax)
                if key not in table or np.abs(ax - np.array(key)).max() > 1e-12:
                    raise NotImplementedError(f"revolute joint {j.get('name')} about axis {ax} (RevoluteUnaligned)")
                idx = add(owner, table[key], R, t, _inertia(links[child]))
                walk(child, idx, np.eye(3), np.zeros(3))
            else:
                raise NotImplementedError(f"joint {j.get('name')} of type {kind}")

    walk(root_link, top, np.eye(3), np.zeros(3))
    inertia = [np.concatenate([[m], c, [I[0, 0], I[0, 1], I[1, 1], I[0, 2], I[1, 2], I[2, 2]]]) for m, c, I in body]
    return ModelDesc(name or robot.get("name", "urdf"), np.array(parents), np.array(jtype), np.array(placement),
                     np.array(inertia))
