"""MJCF (MuJoCo XML) -> flattened model (ModelDesc) for the engine.

Follows what pinocchio's MJCF reader builds (src/parsers/mjcf/mjcf-graph.cpp,
include/pinocchio/parsers/mjcf/model.hxx), restricted to the joints the engine runs:

* <compiler>: ``angle`` (degrees unless "radian"), ``eulerseq`` (three of x/y/z, case
  ignored as the reference does; default xyz), ``boundmass`` / ``boundinertia``,
  ``autolimits``, ``inertiafromgeom`` (mjcf-graph.cpp:690-782);
* <default> classes nest and inherit their parents' element attributes; a joint's
  attributes are resolved as top-level default < the body's childclass < the joint's
  class < the joint element itself (MjcfJoint::fill, mjcf-graph.cpp:239-287);
* the first <body> of <worldbody> is the root link and bodies are taken depth first in
  document order (parseGraph / parseJointAndBody);
* frames (``pos`` plus one of ``quat`` [w x y z], ``axisangle``, ``euler``, ``xyaxes``,
  ``zaxis``) as convertPosition (mjcf-graph.cpp:289-358);
* <inertial>: mass (at least boundmass), lever ``pos``, ``diaginertia`` or
  ``fullinertia`` (diagonal at least boundinertia) rotated by the element's frame
  (convertInertiaFromMjcf, mjcf-graph.cpp:360-407);
* a body without joint is fixed to its parent body: its inertia is merged into the parent
  joint's body and its children hang off that joint with the composed placement; a body
  with one joint gets it with placement parent-body-frame * body pose * joint frame and
  carries its inertia at the inverse joint frame (addSoloJoint, mjcf-graph.cpp:918-971);
* a ``hinge`` about +x, +y or +z (Eigen isApprox, after normalisation) becomes RX / RY / RZ;
  ``free`` / <freejoint> a free flyer (at the root only, as the engine supports);
  ``slide``, ``ball`` and other hinge axes (RevoluteUnaligned) are rejected
  (NotImplementedError);
* several joints in one body form the reference's JointModelComposite (fillModel,
  mjcf-graph.cpp:999-1078); it is flattened here into a chain of single joints linked by
  massless bodies, with the same placements, coordinates (in the same order) and dynamics;
* ``root="freeflyer"`` adds a free flyer carrying the root link's inertia when the root
  body has no joint of its own (parseRootTree, mjcf-graph.cpp:1202-1220); with
  ``root="fixed"`` such a root is fixed to the universe;
* joint ``armature`` goes to the model's armature (default 0).
Inertias from geoms (used by the reference when a body has no <inertial>) are not
computed: such a body with geoms is rejected; a body with neither has zero inertia.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET

import numpy as np

from .models import JT_FREEFLYER, JT_RX, JT_RY, JT_RZ, JT_UNIVERSE, ModelDesc

_ELEMENTS = ("joint", "geom", "site", "mesh", "material", "inertial", "motor")


def _vec(s, n=None):
    v = np.array([float(x) for x in s.split()], dtype=float)
    if n is not None and v.size != n:
        raise ValueError(f"expected {n} numbers, got '{s}'")
    return v


def _axis_angle(axis, angle):
    """Rotation by `angle` about the unit vector `axis` (Rodrigues)."""
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    C = 1.0 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def _quat_wxyz(q):
    w, x, y, z = np.asarray(q, dtype=float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


class _Compiler:
    def __init__(self):
        self.angle = math.pi / 180.0
        self.euler_axes = np.eye(3)
        self.autolimits = True
        self.boundmass = 0.0
        self.boundinertia = 0.0
        self.inertiafromgeom = None  # True, False or None (automatic)

    def read(self, el):
        if el.get("autolimits") == "true":
            self.autolimits = True
        if el.get("angle") == "radian":
            self.angle = 1.0
        if el.get("boundmass") is not None:
            self.boundmass = float(el.get("boundmass"))
        if el.get("boundinertia") is not None:
            self.boundinertia = float(el.get("boundinertia"))
        if el.get("inertiafromgeom") in ("true", "false"):
            self.inertiafromgeom = el.get("inertiafromgeom") == "true"
        seq = el.get("eulerseq")
        if seq is not None:
            if len(seq) != 3 or any(c not in "xyzXYZ" for c in seq):
                raise ValueError("Model tried to use euler angles but euler sequence is wrong")
            for i, c in enumerate(seq.lower()):
                self.euler_axes[:, i] = np.eye(3)["xyz".index(c)]

    def frame(self, el):
        """(R, t) of an element's pos / orientation attributes (convertPosition)."""
        t = _vec(el.get("pos"), 3) if el.get("pos") is not None else np.zeros(3)
        R = np.eye(3)
        if el.get("quat") is not None:
            R = _quat_wxyz(_vec(el.get("quat"), 4))
        elif el.get("axisangle") is not None:
            a = _vec(el.get("axisangle"), 4)
            R = _axis_angle(a[:3] / np.linalg.norm(a[:3]), self.angle * a[3])
        elif el.get("euler") is not None:
            e = _vec(el.get("euler"), 3) * self.angle
            R = (_axis_angle(self.euler_axes[:, 0], e[0]) @ _axis_angle(self.euler_axes[:, 1], e[1])
                 @ _axis_angle(self.euler_axes[:, 2], e[2]))
        elif el.get("xyaxes") is not None:
            xy = _vec(el.get("xyaxes"), 6)
            x = xy[:3] / np.linalg.norm(xy[:3])
            y = xy[3:] - x * np.dot(x, xy[3:])
            y /= np.linalg.norm(y)
            R = np.stack([x, y, np.cross(x, y)], axis=1)
        elif el.get("zaxis") is not None:
            z = _vec(el.get("zaxis"), 3)
            z = z / np.linalg.norm(z)
            ez = np.array([0.0, 0.0, 1.0])
            axis = np.cross(ez, z)
            s, c = np.linalg.norm(axis), float(np.dot(ez, z))
            if s < 1e-12:
                R = np.eye(3) if c > 0 else _axis_angle(np.array([1.0, 0.0, 0.0]), math.pi)
            else:
                R = _axis_angle(axis / s, math.atan2(s, c))
        return R, t

    def inertia(self, el):
        """(m, c, I_c) of an <inertial> element."""
        m = max(float(el.get("mass")), self.boundmass)
        if m < 0:
            raise ValueError("Mass of body is not supposed to be negative")
        c = _vec(el.get("pos"), 3) if el.get("pos") is not None else np.zeros(3)
        R, _ = self.frame(el)
        I = np.eye(3)
        if el.get("diaginertia") is not None:
            I = np.diag(_vec(el.get("diaginertia"), 3))
        elif el.get("fullinertia") is not None:
            f = _vec(el.get("fullinertia"), 6)  # xx yy zz xy xz yz
            I = np.array([[f[0], f[3], f[4]], [f[3], f[1], f[5]], [f[4], f[5], f[2]]])
        for i in range(3):
            I[i, i] = max(I[i, i], self.boundinertia)
        return m, c, R @ I @ R.T


def _defaults(root):
    """Class name -> {element tag: attributes}, each class holding its ancestors'
    attributes unless it sets them itself; "mujoco_default" for the top-level block."""
    classes = {}

    def walk(el, inherited):
        own = {tag: dict(inherited.get(tag, {})) for tag in inherited}
        for child in el:
            if child.tag in _ELEMENTS:
                own.setdefault(child.tag, {}).update(child.attrib)
        name = el.get("class")
        if name is not None:
            classes[name] = own
        elif own:
            classes.setdefault("mujoco_default", own)
        for child in el:
            if child.tag == "default":
                if child.get("class") is None:
                    raise ValueError("Class does not have a name. Cannot parse model.")
                walk(child, own)

    for d in root.findall("default"):
        walk(d, {})
    return classes


def _inertia_sum(a, b):
    (m1, c1, I1), (m2, c2, I2) = a, b
    m = m1 + m2
    if m <= 0:
        return 0.0, np.zeros(3), I1 + I2
    c = (m1 * c1 + m2 * c2) / m

    def shift(mi, ci, Ii):
        d = ci - c
        return Ii + mi * (np.dot(d, d) * np.eye(3) - np.outer(d, d))

    return m, c, shift(m1, c1, I1) + shift(m2, c2, I2)


def _in_parent(inertia, R, t):
    """An inertia (m, c, I_c) of a child frame expressed in the parent frame (R, t)."""
    m, c, I = inertia
    return m, R @ c + t, R @ I @ R.T


_ZERO = (0.0, np.zeros(3), np.zeros((3, 3)))


def parse(path_or_xml: str, root: str = "fixed", name: str = None) -> ModelDesc:
    """Flattened model of an MJCF file (or XML string).  root: "fixed" or "freeflyer"."""
    text = path_or_xml
    if not path_or_xml.lstrip().startswith("<"):
        with open(path_or_xml) as f:
            text = f.read()
    mj = ET.fromstring(text)
    if mj.tag != "mujoco":
        raise ValueError("This is not a standard mujoco model. Cannot parse it.")
    if mj.get("model") is None:
        raise ValueError("Model is missing a name. Cannot parse it")
    if root not in ("fixed", "freeflyer"):
        raise ValueError("root must be 'fixed' or 'freeflyer'")
    comp = _Compiler()
    for c in mj.findall("compiler"):
        comp.read(c)
    classes = _defaults(mj)
    wb = mj.find("worldbody")
    if wb is None or wb.find("body") is None:
        raise ValueError("the model has no body")

    parents, jtype, placement, body, armature = [0], [JT_UNIVERSE], [np.concatenate([np.eye(3).ravel(), np.zeros(3)])], \
        [_ZERO], []

    def add(parent, jt, R, t, inertia, arm):
        parents.append(parent)
        jtype.append(jt)
        placement.append(np.concatenate([R.ravel(), t]))
        body.append(inertia)
        armature.extend([arm] * (6 if jt == JT_FREEFLYER else 1))
        return len(parents) - 1

    def joint_attrs(el, childclass):
        """Resolved attributes of a <joint>: default < childclass < class < element."""
        a = {"type": "hinge", "axis": "0 0 1"}
        a.update(classes.get("mujoco_default", {}).get("joint", {}))
        if childclass is not None:
            a.update(classes[childclass].get("joint", {}))
        if el.get("class") is not None:
            a.update(classes[el.get("class")].get("joint", {}))
        a.update(el.attrib)
        if not comp.autolimits and a.get("range") is not None and a.get("limited") != "true":
            raise ValueError("Range limit is specified but it was not specify to use it.")
        return a

    def body_inertia(el):
        inert = el.find("inertial")
        if comp.inertiafromgeom is False and inert is None:
            raise ValueError("Cannot get inertia from geom and no inertia was found")
        if inert is not None and comp.inertiafromgeom is not True:
            return comp.inertia(inert)
        if el.find("geom") is not None:
            raise NotImplementedError(f"body {el.get('name')}: inertia from geoms is not supported")
        return _ZERO

    def joint_type(a, jname):
        kind = a.get("type", "hinge")
        if kind == "free":
            return JT_FREEFLYER
        if kind != "hinge":
            raise NotImplementedError(f"joint {jname}: type {kind} is not supported (hinge and free only)")
        ax = _vec(a.get("axis", "0 0 1"), 3)
        n = ax / np.linalg.norm(ax)
        for k, jt in enumerate((JT_RX, JT_RY, JT_RZ)):
            if np.linalg.norm(n - np.eye(3)[k]) <= 1e-12:
                return jt
        raise NotImplementedError(f"joint {jname}: hinge about {ax} (RevoluteUnaligned) is not supported")

    def walk(el, owner, Rb, tb, childclass):
        """Body `el`; its parent body's frame is (Rb, tb) in the frame of joint `owner`."""
        childclass = el.get("childclass", childclass)
        R0, t0 = comp.frame(el)  # body pose in the parent body
        R, t = Rb @ R0, Rb @ t0 + tb  # body frame in the owner joint's frame
        js = []
        for j in el:
            if j.tag == "freejoint":
                js.append(({"type": "free"}, np.eye(3), np.zeros(3), j.get("name", "")))
            elif j.tag == "joint":
                a = joint_attrs(j, childclass)
                Rj, tj = comp.frame(j)
                js.append((a, Rj, tj, j.get("name", "")))
        inertia = body_inertia(el)
        if not js:
            body[owner] = _inertia_sum(body[owner], _in_parent(inertia, R, t))
            Rn, tn, new_owner = R, t, owner
        else:
            prev = None
            for k, (a, Rj, tj, jname) in enumerate(js):
                jt = joint_type(a, jname)
                if jt == JT_FREEFLYER and (len(js) > 1 or owner != 0):
                    raise NotImplementedError("a free joint is supported on a root body only, alone in its body")
                Rjp, tjp = R0 @ Rj, R0 @ tj + t0  # joint frame in the parent body (bodyPose * joint)
                if prev is None:
                    Rp, tp = Rb @ Rjp, Rb @ tjp + tb
                else:  # previous joint frame^-1 * this joint frame
                    Rq, tq = prev
                    Rp, tp = Rq.T @ Rjp, Rq.T @ (tjp - tq)
                last = k == len(js) - 1
                Rbj, tbj = Rj.T, -Rj.T @ tj  # body frame in this joint's frame
                owner = add(owner, jt, Rp, tp, _in_parent(inertia, Rbj, tbj) if last else _ZERO,
                            float(a.get("armature", 0.0)))
                prev = (Rjp, tjp)
            Rn, tn, new_owner = Rbj, tbj, owner
        for c in el.findall("body"):
            walk(c, new_owner, Rn, tn, childclass)

    rb = wb.find("body")
    if any(ch.tag in ("joint", "freejoint") for ch in rb):
        walk(rb, 0, np.eye(3), np.zeros(3), None)
    else:
        # a root link without joint is the root joint's body (a free flyer) or fixed to the
        # universe, its frame being that joint's frame: its own pose plays no role
        top = add(0, JT_FREEFLYER, np.eye(3), np.zeros(3), body_inertia(rb), 0.0) if root == "freeflyer" else 0
        for c in rb.findall("body"):
            walk(c, top, np.eye(3), np.zeros(3), rb.get("childclass"))
    inertia = [np.concatenate([[m], c, [I[0, 0], I[0, 1], I[1, 1], I[0, 2], I[1, 2], I[2, 2]]]) for m, c, I in body]
    return ModelDesc(name or mj.get("model"), np.array(parents), np.array(jtype), np.array(placement),
                     np.array(inertia), np.array(armature, dtype=float))
