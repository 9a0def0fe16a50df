"""ORACLE — test infrastructure only; never imported by the product (pippin_amd/).

A Python restatement of Pippin's URDF import (crates/pippin/src/urdf.rs): the
URDF document is rewritten as MJCF and loaded by the oracle's MJCF loader
(oracle/mjcf.py). Numbers are written back with repr(), which round-trips, so
the loader reads exactly the doubles computed here.
"""
from __future__ import annotations

import os
import struct
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

from . import mjcf
from .mjcf import MjcfError


def _floats(s):
    """urdf.rs:34-36"""
    try:
        return [float(t) for t in s.split()]
    except ValueError as e:
        raise MjcfError(f"bad number in URDF: {e}")


def _first(s):
    v = _floats(s)
    if not v:
        raise MjcfError("expected a number in URDF")
    return v[0]


def _attr3(node, name, default):
    """urdf.rs:38-49"""
    s = node.get(name) if node is not None else None
    if s is None:
        return list(default)
    v = _floats(s)
    if len(v) != 3:
        raise MjcfError(f"'{name}' needs 3 numbers")
    return v


def _child(n, tag):
    if n is None:
        return None
    for c in n:
        if c.tag == tag:
            return c
    return None


def _children(n, tag):
    return [c for c in n if c.tag == tag]


def _num(v):
    return repr(float(v))


def rpy_quat(rpy):
    """urdf.rs:60-64: R = Rz Ry Rx (math.rs:214-238 for the quaternion steps)."""
    qz = mjcf.from_axis_angle((0.0, 0.0, 1.0), rpy[2])
    qy = mjcf.from_axis_angle((0.0, 1.0, 0.0), rpy[1])
    qx = mjcf.from_axis_angle((1.0, 0.0, 0.0), rpy[0])
    return mjcf.qnormalized(mjcf.qmul(mjcf.qmul(qz, qy), qx))


def _origin_attrs(origin):
    """urdf.rs:66-73"""
    xyz = _attr3(origin, "xyz", (0.0, 0.0, 0.0))
    q = rpy_quat(_attr3(origin, "rpy", (0.0, 0.0, 0.0)))
    return f' pos="{_num(xyz[0])} {_num(xyz[1])} {_num(xyz[2])}" quat="{" ".join(_num(x) for x in q)}"'


def _escape(s):
    """urdf.rs:200-202"""
    return s.replace("&", "&amp;").replace('"', "&quot;").replace("<", "&lt;").replace(">", "&gt;")


def _pop(d: str) -> str:
    """PathBuf::pop: the path without its last component ('' once none is left)."""
    if not d:
        return d
    s = d.rstrip("/") or "/"
    if s == "/":
        return "/"
    k = s.rfind("/")
    if k < 0:
        return ""
    return "/" if k == 0 else s[:k]


def _join(a, b):
    if not a:
        return b
    return a + b if a.endswith("/") else a + "/" + b


class _Converter:
    def __init__(self, base, floating_base, package_dirs):
        self.base = str(base)
        self.floating_base = floating_base
        self.package_dirs = [str(d) for d in package_dirs]
        self.materials = {}
        self.meshes = []      # (name, path, scale)
        self.mesh_ids = {}
        self.warnings = []

    def resolve(self, filename):
        """urdf.rs:85-113"""
        if filename.startswith("package://"):
            rest = filename[len("package://"):]
            if "/" not in rest:
                return None
            pkg, rel = rest.split("/", 1)
            dirs = list(self.package_dirs)
            env = os.environ.get("ROS_PACKAGE_PATH")
            if env is not None:
                dirs.extend(env.split(":"))
            for up in range(4):
                d = self.base
                for _ in range(up):
                    d = _pop(d)
                dirs.append(d)
            for d in dirs:
                for cand in (_join(_join(d, pkg), rel), _join(d, rel)):
                    if os.path.exists(cand):
                        return os.path.realpath(cand)
            return None
        p = filename[len("file://"):] if filename.startswith("file://") else filename
        if not p.startswith("/"):
            p = _join(self.base, p)
        return os.path.realpath(p) if os.path.exists(p) else None

    def mesh(self, filename, scale):
        """urdf.rs:115-137"""
        path = self.resolve(filename)
        if path is None:
            self.warnings.append(f"mesh '{filename}' not found; geom skipped")
            return None
        name_part = path.rsplit("/", 1)[-1]
        ext = name_part.rsplit(".", 1)[-1].lower() if "." in name_part else ""
        if ext not in ("stl", "obj"):
            self.warnings.append(f"mesh '{filename}': .{ext} not supported (use STL/OBJ); geom skipped")
            return None
        key = (path, tuple(struct.pack("<d", float(s)) for s in scale))
        if key in self.mesh_ids:
            return self.mesh_ids[key]
        name = f"urdf_mesh{len(self.meshes)}"
        self.meshes.append((name, path, list(scale)))
        self.mesh_ids[key] = name
        return name

    def material_rgba(self, visual):
        """urdf.rs:186-193"""
        mat = _child(visual, "material")
        if mat is None:
            return None
        color = _child(mat, "color")
        if color is not None and color.get("rgba") is not None:
            try:
                v = _floats(color.get("rgba"))
            except MjcfError:
                return None
            return v if len(v) == 4 else None
        name = mat.get("name")
        return self.materials.get(name) if name is not None else None

    def geom(self, el, visual, hide):
        """urdf.rs:139-184"""
        geometry = _child(el, "geometry")
        if geometry is None or len(geometry) == 0:
            return None
        shape = geometry[0]
        t = shape.tag
        g = "<geom"
        if t == "box":
            s = _attr3(shape, "size", (0.0, 0.0, 0.0))
            g += f' type="box" size="{_num(s[0] / 2.0)} {_num(s[1] / 2.0)} {_num(s[2] / 2.0)}"'
        elif t in ("cylinder", "capsule"):
            r = _first(shape.get("radius", "0"))
            ln = _first(shape.get("length", "0"))
            g += f' type="{t}" size="{_num(r)} {_num(ln / 2.0)}"'
        elif t == "sphere":
            g += f' type="sphere" size="{_num(_first(shape.get("radius", "0")))}"'
        elif t == "mesh":
            name = self.mesh(shape.get("filename", ""), _attr3(shape, "scale", (1.0, 1.0, 1.0)))
            if name is None:
                return None
            g += f' type="mesh" mesh="{name}"'
        else:
            self.warnings.append(f"URDF geometry <{t}> not supported; skipped")
            return None
        g += _origin_attrs(_child(el, "origin"))
        if visual:
            g += ' contype="0" conaffinity="0" group="1" mass="0"'
            rgba = self.material_rgba(el)
            if rgba is not None:
                g += f' rgba="{" ".join(_num(x) for x in rgba)}"'
        elif hide:
            g += ' group="3"'
        return g + "/>"


def _attr_or(node, name, default):
    s = node.get(name) if node is not None else None
    return _first(s) if s is not None else default


def _emit_link(cv, links, kids, link, out):
    """urdf.rs:251-311"""
    l = links[link]
    inertial = _child(l, "inertial")
    if inertial is not None:
        mass = _attr_or(_child(inertial, "mass"), "value", 0.0)
        i = _child(inertial, "inertia")
        vals = [_attr_or(i, k, 0.0) for k in ("ixx", "iyy", "izz", "ixy", "ixz", "iyz")]
        if mass > 0.0:
            out.append(f'<inertial{_origin_attrs(_child(inertial, "origin"))} mass="{_num(mass)}" '
                       f'fullinertia="{" ".join(_num(x) for x in vals)}"/>')
    has_visual = len(_children(l, "visual")) > 0
    for c in _children(l, "collision"):
        g = cv.geom(c, False, has_visual)
        if g is not None:
            out.append(g)
    for v in _children(l, "visual"):
        g = cv.geom(v, True, False)
        if g is not None:
            out.append(g)
    for j in kids.get(link, []):
        name = _escape(j.get("name", "joint"))
        ch = _child(j, "child").get("link")
        out.append(f'<body name="{_escape(ch)}"{_origin_attrs(_child(j, "origin"))}>')
        axis = _attr3(_child(j, "axis"), "xyz", (1.0, 0.0, 0.0))
        damping = _attr_or(_child(j, "dynamics"), "damping", 0.0)
        limit = _child(j, "limit")
        lo, hi = _attr_or(limit, "lower", 0.0), _attr_or(limit, "upper", 0.0)
        rng = f' range="{_num(lo)} {_num(hi)}"' if lo < hi else ""
        common = f'name="{name}" axis="{_num(axis[0])} {_num(axis[1])} {_num(axis[2])}" damping="{_num(damping)}"'
        t = j.get("type", "fixed")
        if t == "revolute":
            out.append(f'<joint type="hinge" {common}{rng}/>')
        elif t == "continuous":
            out.append(f'<joint type="hinge" {common}/>')
        elif t == "prismatic":
            out.append(f'<joint type="slide" {common}{rng}/>')
        elif t == "fixed":
            pass
        elif t == "floating":
            out.append(f'<freejoint name="{name}"/>')
        else:
            cv.warnings.append(f"joint type '{t}' not supported; treated as fixed")
        _emit_link(cv, links, kids, ch, out)
        out.append("</body>")


def to_mjcf(urdf: str, base=".", floating_base=False, package_dirs=()):
    """urdf.rs:205-327: (MJCF text, warnings)."""
    robot = ET.fromstring(urdf)
    if robot.tag != "robot":
        raise MjcfError("URDF root element must be <robot>")
    cv = _Converter(base, floating_base, package_dirs)
    for m in _children(robot, "material"):
        color = _child(m, "color")
        if m.get("name") is not None and color is not None and color.get("rgba") is not None:
            v = _floats(color.get("rgba"))
            if len(v) == 4:
                cv.materials[m.get("name")] = v
    links, order = {}, []
    for l in _children(robot, "link"):
        if l.get("name") is not None:
            links[l.get("name")] = l
            order.append(l.get("name"))
    kids, is_child = {}, set()
    for j in _children(robot, "joint"):
        p = _child(j, "parent")
        c = _child(j, "child")
        p = p.get("link") if p is not None else None
        c = c.get("link") if c is not None else None
        if p is None or c is None:
            raise MjcfError("joint missing parent/child")
        if p not in links or c not in links:
            raise MjcfError(f"joint '{j.get('name', '?')}' references unknown link")
        if c in is_child:
            raise MjcfError(f"link '{c}' has two parent joints (URDF must be a tree)")
        is_child.add(c)
        kids.setdefault(p, []).append(j)
    roots = [l for l in order if l not in is_child]
    if len(roots) != 1:
        raise MjcfError(f"URDF must have exactly one root link, found {roots}")
    body = [f'<body name="{_escape(roots[0])}">']
    if floating_base:
        body.append('<freejoint name="root"/>')
    _emit_link(cv, links, kids, roots[0], body)
    body.append("</body>")
    assets = "".join(f'<mesh name="{n}" file="{_escape(p)}" scale="{" ".join(_num(x) for x in s)}"/>'
                     for n, p, s in cv.meshes)
    name = _escape(robot.get("name", "robot"))
    text = (f'<mujoco model="{name}"><compiler angle="radian" autolimits="true"/><asset>{assets}</asset>'
            f'<worldbody>{"".join(body)}</worldbody></mujoco>')
    return text, cv.warnings


def load_file(path, floating_base=False, package_dirs=()) -> dict:
    """urdf.rs:330-339"""
    path = Path(path)
    base = str(path.parent) if str(path.parent) else "."
    text, warnings = to_mjcf(path.read_text(), base, floating_base, package_dirs)
    for w in warnings:
        print(f"pippin urdf warning: {w}", file=sys.stderr)
    return mjcf.load_string(text, base)
