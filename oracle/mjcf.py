"""ORACLE — test infrastructure only; never imported by the product (pippin_amd/).

A Python restatement of Pippin's MJCF loader (crates/pippin/src/mjcf.rs) for the
subset the engine restatement supports: defaults/classes, compiler angle and
eulerseq, options, bodies with freejoint/hinge/slide/ball joints, inertial or
geom-derived mass (plane/sphere/capsule/box/cylinder), cameras, motor/position/
velocity/general actuators with joint or fixed-tendon transmission, joint
equality and <contact><exclude>. It returns the raw fields (before
`Model::compile`) as a dict of numpy arrays in the layout of ph_model_desc.

The arithmetic follows mjcf.rs operation for operation (Python floats are IEEE
doubles and nothing is fused), so the result can be compared bit for bit with
the product's own C++ loader.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

from . import mesh as mesh_mod

PI = math.pi
JNT = {"free": 0, "ball": 1, "slide": 2, "hinge": 3}
GEOM = {"plane": 0, "sphere": 1, "capsule": 2, "box": 3, "cylinder": 4, "mesh": 5}


class MjcfError(ValueError):
    pass


# ---- math.rs restated on tuples ----
def _sum(xs):
    s = -0.0  # Rust's f64 Sum folds from -0.0
    for x in xs:
        s += x
    return s


def vadd(a, b): return (a[0] + b[0], a[1] + b[1], a[2] + b[2])
def vsub(a, b): return (a[0] - b[0], a[1] - b[1], a[2] - b[2])
def vscale(a, s): return (a[0] * s, a[1] * s, a[2] * s)
def vdot(a, b): return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
def vcross(a, o): return (a[1] * o[2] - a[2] * o[1], a[2] * o[0] - a[0] * o[2], a[0] * o[1] - a[1] * o[0])
def vnorm(a): return math.sqrt(vdot(a, a))


def vnormalized(a):
    n = vnorm(a)
    return vscale(a, 1.0 / n) if n > 1e-15 else (0.0, 0.0, 0.0)


def mmul(a, o):
    return tuple(_sum(a[3 * i + k] * o[3 * k + j] for k in range(3)) for i in range(3) for j in range(3))


def mT(m): return (m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8])
def mskew(v): return (0.0, -v[2], v[1], v[2], 0.0, -v[0], -v[1], v[0], 0.0)
def mscale(m, s): return tuple(x * s for x in m)
def madd(a, b): return tuple(x + y for x, y in zip(a, b))
def mdiag(d): return (d[0], 0.0, 0.0, 0.0, d[1], 0.0, 0.0, 0.0, d[2])
def mrotate(m, r): return mmul(mmul(r, m), mT(r))
def from_cols(a, b, c): return (a[0], b[0], c[0], a[1], b[1], c[1], a[2], b[2], c[2])


def qnormalized(q):
    n = math.sqrt(_sum(v * v for v in q))
    if n < 1e-15:
        return (1.0, 0.0, 0.0, 0.0)
    return tuple(v / n for v in q)


def qmul(p, o):
    a, b, c, d = p
    e, f, g, h = o
    return (a * e - b * f - c * g - d * h, a * f + b * e + c * h - d * g,
            a * g - b * h + c * e + d * f, a * h + b * g - c * f + d * e)


def from_axis_angle(axis, angle):
    a = vnormalized(axis)
    s, c = math.sin(0.5 * angle), math.cos(0.5 * angle)
    return (c, a[0] * s, a[1] * s, a[2] * s)


def to_mat(q):
    w, x, y, z = q
    ww, xx, yy, zz = w * w, x * x, y * y, z * z
    xy, xz, yz, wx, wy, wz = x * y, x * z, y * z, w * x, w * y, w * z
    return (ww + xx - yy - zz, 2.0 * (xy - wz), 2.0 * (xz + wy), 2.0 * (xy + wz), ww - xx + yy - zz,
            2.0 * (yz - wx), 2.0 * (xz - wy), 2.0 * (yz + wx), ww - xx - yy + zz)


def from_mat(m):
    tr = m[0] + m[4] + m[8]
    if tr > 0.0:
        s = math.sqrt(tr + 1.0) * 2.0
        q = (0.25 * s, (m[7] - m[5]) / s, (m[2] - m[6]) / s, (m[3] - m[1]) / s)
    elif m[0] > m[4] and m[0] > m[8]:
        s = math.sqrt(1.0 + m[0] - m[4] - m[8]) * 2.0
        q = ((m[7] - m[5]) / s, 0.25 * s, (m[1] + m[3]) / s, (m[2] + m[6]) / s)
    elif m[4] > m[8]:
        s = math.sqrt(1.0 + m[4] - m[0] - m[8]) * 2.0
        q = ((m[2] - m[6]) / s, (m[1] + m[3]) / s, 0.25 * s, (m[5] + m[7]) / s)
    else:
        s = math.sqrt(1.0 + m[8] - m[0] - m[4]) * 2.0
        q = ((m[3] - m[1]) / s, (m[2] + m[6]) / s, (m[5] + m[7]) / s, 0.25 * s)
    q = qnormalized(q)
    return tuple(-v for v in q) if q[0] < 0.0 else q


def from_z_to(d):
    d = vnormalized(d)
    c = vdot((0.0, 0.0, 1.0), d)
    if c < -1.0 + 1e-12:
        return (0.0, 1.0, 0.0, 0.0)
    ax = vcross((0.0, 0.0, 1.0), d)
    return qnormalized((1.0 + c, ax[0], ax[1], ax[2]))


def parse_floats(s: str):
    try:
        return [float(t) for t in s.split()]
    except ValueError as e:
        raise MjcfError(f"bad number in '{s}'") from e


class Ctx:
    """mjcf.rs:35-140"""

    def __init__(self):
        self.materials = {}
        self.degrees = True
        self.eulerseq = "xyz"
        self.autolimits = True
        self.defaults = {}

    def get(self, node, cls, name):
        v = node.get(name)
        if v is not None:
            return v
        tag = node.tag
        cls = node.get("class", cls)
        dc = self.defaults.get(cls)
        if dc is None:
            return None
        v = dc.get(tag, {}).get(name)
        if v is not None:
            return v
        if tag in ("motor", "position", "velocity"):
            v = dc.get("general", {}).get(name)
            if v is not None:
                return v
        return None

    def floats(self, node, cls, name):
        s = self.get(node, cls, name)
        return None if s is None else parse_floats(s)

    def float(self, node, cls, name, default):
        v = self.floats(node, cls, name)
        return v[0] if v else default

    def vec3(self, node, cls, name, default):
        v = self.floats(node, cls, name)
        if v is None:
            return default
        if len(v) >= 3:
            return (v[0], v[1], v[2])
        raise MjcfError(f"attribute '{name}' needs 3 numbers")

    def angle(self, a):
        return a * PI / 180.0 if self.degrees else a

    def orientation(self, node, cls):
        q = self.floats(node, cls, "quat")
        if q is not None:
            if len(q) != 4:
                raise MjcfError("quat needs 4 numbers")
            return qnormalized(tuple(q))
        a = self.floats(node, cls, "axisangle")
        if a is not None:
            if len(a) != 4:
                raise MjcfError("axisangle needs 4 numbers")
            return from_axis_angle((a[0], a[1], a[2]), self.angle(a[3]))
        e = self.floats(node, cls, "euler")
        if e is not None:
            if len(e) != 3:
                raise MjcfError("euler needs 3 numbers")
            q = (1.0, 0.0, 0.0, 0.0)
            for i in range(3):
                c = self.eulerseq[i]
                axis = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}.get(c.lower())
                if axis is None:
                    raise MjcfError("bad eulerseq")
                r = from_axis_angle(axis, self.angle(e[i]))
                q = qmul(q, r) if c.islower() else qmul(r, q)
            return qnormalized(q)
        x = self.floats(node, cls, "xyaxes")
        if x is not None:
            if len(x) != 6:
                raise MjcfError("xyaxes needs 6 numbers")
            xa = vnormalized((x[0], x[1], x[2]))
            ya = (x[3], x[4], x[5])
            ya = vnormalized(vsub(ya, vscale(xa, vdot(xa, ya))))
            za = vcross(xa, ya)
            return from_mat(from_cols(xa, ya, za))
        z = self.floats(node, cls, "zaxis")
        if z is not None:
            if len(z) != 3:
                raise MjcfError("zaxis needs 3 numbers")
            return from_z_to((z[0], z[1], z[2]))
        return (1.0, 0.0, 0.0, 0.0)


def _parse_default(ctx, node, parent, is_top):
    """mjcf.rs:398-412"""
    name = "main" if is_top else node.get("class", "main")
    dc = {t: dict(a) for t, a in ctx.defaults.get(parent, {}).items()}
    for c in node:
        if c.tag == "default":
            continue
        entry = dc.setdefault(c.tag, {})
        for k, v in c.attrib.items():
            entry[k] = v
    ctx.defaults[name] = dc
    for c in node:
        if c.tag == "default":
            _parse_default(ctx, c, name, False)


def geom_volume_inertia(gt, s):
    """mjcf.rs:774-808: volume and principal inertia per unit density."""
    if gt == 1:
        r = s[0]
        v = 4.0 / 3.0 * PI * r * r * r
        i = 0.4 * v * r * r
        return v, (i, i, i)
    if gt == 3:
        a, b, c = s
        v = 8.0 * a * b * c
        return v, (v * (b * b + c * c) / 3.0, v * (a * a + c * c) / 3.0, v * (a * a + b * b) / 3.0)
    if gt == 4:
        r, h = s[0], s[1]
        v = PI * r * r * 2.0 * h
        ix = v * (3.0 * r * r + 4.0 * h * h) / 12.0
        return v, (ix, ix, 0.5 * v * r * r)
    if gt == 2:
        r, h = s[0], s[1]
        vc = PI * r * r * 2.0 * h
        vs = 4.0 / 3.0 * PI * r * r * r
        v = vc + vs
        hs = vs * 0.5
        d = h + 3.0 * r / 8.0
        hs_ix = hs * (83.0 / 320.0) * r * r
        ix = vc * (3.0 * r * r + 4.0 * h * h) / 12.0 + 2.0 * (hs_ix + hs * d * d)
        iz = 0.5 * vc * r * r + 0.4 * vs * r * r
        return v, (ix, ix, iz)
    return 0.0, (0.0, 0.0, 0.0)


class _Builder:
    def __init__(self):
        self.meshes = []  # (name, hull vertices, hull faces)
        self.nq = self.nv = 0
        self.timestep = 0.002
        self.gravity = (0.0, 0.0, -9.81)
        self.integrator = 0
        self.solver = dict(kind=0, iterations=30, tolerance=1e-8, contact_margin=1e-3,
                           solref=[0.02, 1.0], solimp=[0.9, 0.95, 0.001, 0.5, 2.0], impratio=1.0,
                           contact_hertz=0.0, contact_damping_ratio=1.0, contact_slop=5e-4)
        self.body = dict(names=[], parent=[], pos=[], quat=[], jntadr=[], jntnum=[], dofadr=[],
                         dofnum=[], mass=[], ipos=[], inertia=[])
        self.geom_mass = []
        self.explicit = []
        self.jnt = dict(names=[], type=[], body=[], qposadr=[], dofadr=[], pos=[], axis=[], limited=[],
                        range=[], stiffness=[], springref=[])
        self.dof = dict(body=[], jnt=[], damping=[], armature=[])
        self.geom = dict(type=[], body=[], pos=[], quat=[], size=[], friction=[], contype=[],
                         conaffinity=[], rgba=[], group=[], dataid=[])
        self.cam = dict(body=[], pos=[], quat=[], fovy=[])
        self.act = dict(names=[], gear=[], gain=[], bias=[], ctrllimited=[], ctrlrange=[],
                        forcelimited=[], forcerange=[], moment=[])
        self.tendons = {}
        self.eq = dict(j1=[], j2=[], poly=[], solref=[], solimp=[])
        self.exclude = []
        self.qpos0 = []

    def add_body(self, name, parent, pos, quat):
        b = self.body
        b["names"].append(name)
        b["parent"].append(parent)
        b["pos"].append(pos)
        b["quat"].append(quat)
        b["jntadr"].append(len(self.jnt["type"]))
        b["jntnum"].append(0)
        b["dofadr"].append(self.nv)
        b["dofnum"].append(0)
        b["mass"].append(0.0)
        b["ipos"].append((0.0, 0.0, 0.0))
        b["inertia"].append((0.0,) * 9)
        self.geom_mass.append([0.0, (0.0, 0.0, 0.0), (0.0,) * 9])
        self.explicit.append(False)
        return len(b["names"]) - 1


def _parse_body_contents(ctx, bld, node, body, cls):
    """mjcf.rs:522-557"""
    for c in node:
        if not isinstance(c.tag, str):
            continue
        t = c.tag
        if t == "geom":
            _parse_geom(ctx, bld, c, body, cls)
        elif t in ("joint", "freejoint"):
            if body == 0:
                raise MjcfError("joints cannot be attached to the world body")
            _parse_joint(ctx, bld, c, body, cls)
        elif t == "inertial":
            _parse_inertial(ctx, bld, c, body, cls)
        elif t == "body":
            name = c.get("name", f"body{len(bld.body['names'])}")
            pos = ctx.vec3(c, cls, "pos", (0.0, 0.0, 0.0))
            quat = ctx.orientation(c, cls)
            child_cls = c.get("childclass", cls)
            bid = bld.add_body(name, body, pos, quat)
            _parse_body_contents(ctx, bld, c, bid, child_cls)
        elif t == "camera":
            cm = bld.cam
            cm["body"].append(body)
            cm["pos"].append(ctx.vec3(c, cls, "pos", (0.0, 0.0, 0.0)))
            cm["quat"].append(ctx.orientation(c, cls))
            cm["fovy"].append(ctx.float(c, cls, "fovy", 45.0))


def _parse_joint(ctx, bld, node, body, cls):
    """mjcf.rs:559-633"""
    if node.tag == "freejoint":
        jt = 0
    else:
        tname = ctx.get(node, cls, "type") or "hinge"
        if tname not in JNT:
            raise MjcfError(f"unknown joint type '{tname}'")
        jt = JNT[tname]
    j = bld.jnt
    jid = len(j["type"])
    name = node.get("name", f"joint{jid}")
    pos = (0.0, 0.0, 0.0) if jt == 0 else ctx.vec3(node, cls, "pos", (0.0, 0.0, 0.0))
    axis = vnormalized(ctx.vec3(node, cls, "axis", (0.0, 0.0, 1.0)))
    rng = ctx.floats(node, cls, "range")
    lim = ctx.get(node, cls, "limited")
    if lim == "true":
        limited = True
    elif lim == "false":
        limited = False
    else:
        limited = ctx.autolimits and rng is not None
    limited = limited and jt in (2, 3)
    r = [rng[0], rng[1]] if rng is not None else [0.0, 0.0]
    if jt == 3:
        r = [ctx.angle(r[0]), ctx.angle(r[1])]
    refv = ctx.float(node, cls, "ref", 0.0)
    springref = ctx.float(node, cls, "springref", 0.0)
    if jt == 3:
        refv = ctx.angle(refv)
        springref = ctx.angle(springref)
    damping = ctx.float(node, cls, "damping", 0.0)
    armature = ctx.float(node, cls, "armature", 0.0)
    stiffness = ctx.float(node, cls, "stiffness", 0.0)
    j["names"].append(name)
    j["type"].append(jt)
    j["body"].append(body)
    j["qposadr"].append(bld.nq)
    j["dofadr"].append(bld.nv)
    j["pos"].append(pos)
    j["axis"].append(axis)
    j["limited"].append(1 if limited else 0)
    j["range"].append(r)
    j["stiffness"].append(stiffness)
    j["springref"].append(springref)
    if jt == 0:
        p = bld.body["pos"][body]
        q = bld.body["quat"][body]
        bld.qpos0.extend([p[0], p[1], p[2], q[0], q[1], q[2], q[3]])
    elif jt == 1:
        bld.qpos0.extend([1.0, 0.0, 0.0, 0.0])
    else:
        bld.qpos0.append(refv)
    nv = {0: 6, 1: 3}.get(jt, 1)
    nq = {0: 7, 1: 4}.get(jt, 1)
    for _ in range(nv):
        bld.dof["body"].append(body)
        bld.dof["jnt"].append(jid)
        bld.dof["damping"].append(damping)
        bld.dof["armature"].append(armature)
    bld.nq += nq
    bld.nv += nv
    b = bld.body
    b["jntnum"][body] += 1
    b["dofnum"][body] += nv
    if b["jntnum"][body] == 1:
        b["jntadr"][body] = jid
        b["dofadr"][body] = j["dofadr"][jid]


def _parse_geom(ctx, bld, node, body, cls):
    """mjcf.rs:635-738"""
    tname = ctx.get(node, cls, "type") or "sphere"
    if tname not in GEOM:
        return  # unsupported geom types are skipped with a warning
    gt = GEOM[tname]
    dataid = -1
    if gt == 5:
        mname = ctx.get(node, cls, "mesh")
        if mname is None:
            raise MjcfError("mesh geom needs a 'mesh' attribute")
        names = [mm[0] for mm in bld.meshes]
        if mname not in names:
            raise MjcfError(f"unknown mesh '{mname}'")
        dataid = names.index(mname)
    sz = ctx.floats(node, cls, "size") or []
    get = lambda i: sz[i] if i < len(sz) else 0.0
    size = (get(0), get(1), get(2))
    pos = ctx.vec3(node, cls, "pos", (0.0, 0.0, 0.0))
    quat = ctx.orientation(node, cls)
    ft = ctx.floats(node, cls, "fromto")
    if ft is not None:
        if len(ft) != 6:
            raise MjcfError("fromto needs 6 numbers")
        a = (ft[0], ft[1], ft[2])
        bb = (ft[3], ft[4], ft[5])
        d = vsub(bb, a)
        pos = vscale(vadd(a, bb), 0.5)
        quat = from_z_to(d)
        if gt in (2, 4):
            size = (get(0), vnorm(d) * 0.5, 0.0)
        elif gt == 3:
            size = (get(0), get(1), vnorm(d) * 0.5)
        else:
            raise MjcfError("fromto only valid for capsule, cylinder, box")
    fr = ctx.floats(node, cls, "friction")
    friction = fr[0] if fr else 1.0
    contype = _as_u32(ctx.float(node, cls, "contype", 1.0))
    conaffinity = _as_u32(ctx.float(node, cls, "conaffinity", 1.0))
    rgba = ctx.floats(node, cls, "rgba")
    if rgba is not None and len(rgba) == 4:
        rgba = tuple(float(np.float32(x)) for x in rgba)
    else:
        mat = ctx.get(node, cls, "material")
        rgba = ctx.materials.get(mat) if mat is not None else None
    group = _as_i32(ctx.float(node, cls, "group", 0.0))
    if gt == 5:  # mjcf.rs:686-703: the hull's solid mass properties
        _name, hv, hf = bld.meshes[dataid]
        vol, com, icom = mesh_mod.mass_properties(hv, mesh_mod.triangles(hf))
        mv = ctx.floats(node, cls, "mass")
        mass = mv[0] if mv is not None else ctx.float(node, cls, "density", 1000.0) * vol
        if mass > 0.0 and vol > 0.0:
            r = to_mat(quat)
            rc = tuple(r[3 * i] * com[0] + r[3 * i + 1] * com[1] + r[3 * i + 2] * com[2] for i in range(3))
            c = vadd(pos, rc)
            ig = mrotate(mscale(tuple(float(x) for x in np.asarray(icom).reshape(9)), mass / vol), r)
            cc = mskew(c)
            shifted = madd(ig, mscale(mmul(cc, mT(cc)), mass))
            e = bld.geom_mass[body]
            e[0] += mass
            e[1] = vadd(e[1], vscale(c, mass))
            e[2] = madd(e[2], shifted)
    elif gt != 0:
        vol, unit = geom_volume_inertia(gt, size)
        mv = ctx.floats(node, cls, "mass")
        mass = mv[0] if mv is not None else ctx.float(node, cls, "density", 1000.0) * vol
        if mass > 0.0 and vol > 0.0:
            ig = mrotate(mdiag(vscale(unit, mass / vol)), to_mat(quat))
            pp = mskew(pos)
            shifted = madd(ig, mscale(mmul(pp, mT(pp)), mass))
            e = bld.geom_mass[body]
            e[0] += mass
            e[1] = vadd(e[1], vscale(pos, mass))
            e[2] = madd(e[2], shifted)
    g = bld.geom
    g["type"].append(gt)
    g["body"].append(body)
    g["pos"].append(pos)
    g["quat"].append(quat)
    g["size"].append(size)
    g["friction"].append(friction)
    g["contype"].append(contype)
    g["conaffinity"].append(conaffinity)
    g["rgba"].append(rgba if rgba is not None else (0.5, 0.5, 0.5, 1.0))
    g["group"].append(group)
    g["dataid"].append(dataid)


def _as_u32(x):
    """Rust `f64 as u32` (saturating; NaN -> 0)."""
    if not (x > 0.0):
        return 0
    return min(int(x), 4294967295)


def _as_i32(x):
    if x != x:
        return 0
    return max(min(int(x), 2147483647), -2147483648)


def _parse_inertial(ctx, bld, node, body, cls):
    """mjcf.rs:810-828"""
    pos = ctx.vec3(node, cls, "pos", (0.0, 0.0, 0.0))
    quat = ctx.orientation(node, cls)
    mass = ctx.float(node, cls, "mass", 0.0)
    d = ctx.floats(node, cls, "diaginertia")
    if d is not None:
        inertia = mrotate(mdiag((d[0], d[1], d[2])), to_mat(quat))
    else:
        f = ctx.floats(node, cls, "fullinertia")
        if f is None:
            raise MjcfError("<inertial> requires diaginertia or fullinertia")
        inertia = mrotate((f[0], f[3], f[4], f[3], f[1], f[5], f[4], f[5], f[2]), to_mat(quat))
    bld.body["mass"][body] = mass
    bld.body["ipos"][body] = pos
    bld.body["inertia"][body] = inertia
    bld.explicit[body] = True


def _finalize_inertia(bld):
    """mjcf.rs:830-850"""
    b = bld.body
    for body in range(1, len(b["names"])):
        if bld.explicit[body]:
            continue
        mass, mc, second = bld.geom_mass[body]
        if mass <= 0.0:
            if b["jntnum"][body] > 0:
                raise MjcfError(f"moving body '{b['names'][body]}' has no mass")
            continue
        com = vscale(mc, 1.0 / mass)
        cx = mskew(com)
        icom = madd(second, mscale(mmul(cx, mT(cx)), -mass))
        b["mass"][body] = mass
        b["ipos"][body] = com
        b["inertia"][body] = icom


def _parse_actuator(ctx, bld, node):
    """mjcf.rs:852-921"""
    tag = node.tag
    if tag not in ("motor", "position", "velocity", "general"):
        raise MjcfError(f"actuator <{tag}> not supported")
    cls = "main"
    jn = node.get("joint")
    tn = node.get("tendon")
    if jn is not None:
        if jn not in bld.jnt["names"]:
            raise MjcfError(f"actuator references unknown joint '{jn}'")
        j = bld.jnt["names"].index(jn)
        if bld.jnt["type"][j] not in (2, 3):
            raise MjcfError("actuators on free/ball joints are not supported")
        moment = [(j, 1.0)]
    elif tn is not None:
        if tn not in bld.tendons:
            raise MjcfError(f"actuator references unknown tendon '{tn}'")
        moment = list(bld.tendons[tn])
    else:
        raise MjcfError("only joint and fixed-tendon transmissions are supported")
    gear = ctx.float(node, cls, "gear", 1.0)
    if tag == "motor":
        gain, bias = 1.0, (0.0, 0.0, 0.0)
    elif tag == "position":
        kp = ctx.float(node, cls, "kp", 1.0)
        kv = ctx.float(node, cls, "kv", 0.0)
        gain, bias = kp, (0.0, -kp, -kv)
    elif tag == "velocity":
        kv = ctx.float(node, cls, "kv", 1.0)
        gain, bias = kv, (0.0, 0.0, -kv)
    else:
        gp = ctx.floats(node, cls, "gainprm")
        gain = gp[0] if gp is not None else 1.0
        bp = ctx.floats(node, cls, "biasprm") or []
        if ctx.get(node, cls, "biastype") == "affine":
            bias = tuple(bp[i] if i < len(bp) else 0.0 for i in range(3))
        else:
            bias = (0.0, 0.0, 0.0)

    def limited(flag, rng):
        r = ctx.floats(node, cls, rng)
        f = ctx.get(node, cls, flag)
        lim = True if f == "true" else False if f == "false" else (ctx.autolimits and r is not None)
        return lim, ([r[0], r[1]] if r is not None else [0.0, 0.0])

    cl, cr = limited("ctrllimited", "ctrlrange")
    fl, fr = limited("forcelimited", "forcerange")
    a = bld.act
    a["names"].append(node.get("name", f"actuator{len(a['names'])}"))
    a["gear"].append(gear)
    a["gain"].append(gain)
    a["bias"].append(bias)
    a["ctrllimited"].append(1 if cl else 0)
    a["ctrlrange"].append(cr)
    a["forcelimited"].append(1 if fl else 0)
    a["forcerange"].append(fr)
    a["moment"].append(moment)


def _parse_tendon(ctx, bld, node):
    """mjcf.rs:923-945"""
    if node.tag != "fixed":
        return
    joints = []
    for jn in node:
        if jn.tag != "joint":
            continue
        name = jn.get("joint", "")
        if name not in bld.jnt["names"]:
            raise MjcfError(f"tendon references unknown joint '{name}'")
        j = bld.jnt["names"].index(name)
        if bld.jnt["type"][j] not in (2, 3):
            raise MjcfError("fixed tendons need hinge or slide joints")
        c = jn.get("coef")
        joints.append((j, parse_floats(c)[0] if c is not None else 1.0))
    if not joints:
        raise MjcfError("fixed tendon needs at least one joint")
    name = node.get("name", f"tendon{len(bld.tendons)}")
    bld.tendons[name] = joints


def _parse_equality(ctx, bld, node):
    """mjcf.rs:947-998"""
    if node.tag != "joint":
        return
    if node.get("active") == "false":
        return

    def get(name):
        v = node.get(name)
        if v is not None:
            return v
        dc = ctx.defaults.get(node.get("class", "main"))
        if dc is None:
            return None
        return dc.get("equality", {}).get(name)

    def joint(attr):
        name = node.get(attr, "")
        if name not in bld.jnt["names"]:
            raise MjcfError(f"equality references unknown joint '{name}'")
        j = bld.jnt["names"].index(name)
        if bld.jnt["type"][j] not in (2, 3):
            raise MjcfError("joint equality needs hinge or slide joints")
        return j

    j1 = joint("joint1")
    j2 = joint("joint2") if node.get("joint2") is not None else -1
    poly = [0.0, 1.0, 0.0, 0.0, 0.0]
    v = get("polycoef")
    if v is not None:
        for k, x in enumerate(parse_floats(v)[:5]):
            poly[k] = x
    solref = list(bld.solver["solref"])
    v = get("solref")
    if v is not None:
        r = parse_floats(v)
        solref = [r[0], r[1] if len(r) > 1 else 1.0]
    solimp = list(bld.solver["solimp"])
    v = get("solimp")
    if v is not None:
        for k, x in enumerate(parse_floats(v)[:5]):
            solimp[k] = x
    e = bld.eq
    e["j1"].append(j1)
    e["j2"].append(j2)
    e["poly"].append(poly)
    e["solref"].append(solref)
    e["solimp"].append(solimp)


def _parse_mesh(ctx, bld, node, meshdir):
    """mjcf.rs:740-772: vertices from an OBJ/STL file or the 'vertex' attribute,
    scaled per axis, then their convex hull (a degenerate set is an error)."""
    f = node.get("file")
    name = node.get("name") or (Path(f).stem if f else None)
    if name is None:
        raise MjcfError("mesh needs a name or file")
    scale = ctx.vec3(node, "main", "scale", (1.0, 1.0, 1.0))
    if f is not None:
        path = meshdir / f
        ext = path.suffix.lower()
        if ext == ".obj":
            pts = mesh_mod.read_obj_vertices(path.read_text(), scale)
        elif ext == ".stl":
            pts = mesh_mod.read_stl_vertices(path.read_bytes(), scale)
        else:
            raise MjcfError(f"unsupported mesh format '{ext[1:]}' (use .stl or .obj)")
    else:
        v = ctx.floats(node, "main", "vertex")
        if v is None:
            raise MjcfError(f"mesh '{name}' needs 'file' or 'vertex'")
        pts = np.array([(v[i] * scale[0], v[i + 1] * scale[1], v[i + 2] * scale[2])
                        for i in range(0, len(v) - 2, 3)], dtype=np.float64).reshape(-1, 3)
    hull = mesh_mod.convex_hull(pts)
    if hull is None:
        raise MjcfError(f"mesh '{name}' is degenerate (no volume)")
    bld.meshes.append((name, hull[0], hull[1]))


def _expand_includes(node, directory: Path, depth: int):
    """mjcf.rs:178-209: replace each <include file=.../> by the children of the
    included document's root, files relative to the including file's folder,
    at most 16 levels deep."""
    kids = []
    for k in list(node):
        if k.tag != "include":
            _expand_includes(k, directory, depth)
            kids.append(k)
            continue
        path = directory / k.get("file", "")
        try:
            text = path.read_text()
        except OSError as e:
            raise MjcfError(f'<include file="{k.get("file", "")}">: {e}')
        if depth + 1 > 16:  # the included text would be expanded one level deeper
            raise MjcfError("<include> nested too deeply (cycle?)")
        sub = ET.fromstring(text)
        _expand_includes(sub, path.parent, depth + 1)
        kids.extend(list(sub))
    for k in list(node):
        node.remove(k)
    node.extend(kids)


def load_string(xml: str, directory: str | Path = ".") -> dict:
    """mjcf.rs:211-395."""
    root = ET.fromstring(xml)
    if root.tag != "mujoco":
        raise MjcfError("root element must be <mujoco>")
    _expand_includes(root, Path(directory), 0)
    ctx = Ctx()
    meshdir = Path(directory)
    for c in root.findall("compiler"):
        d = c.get("meshdir", c.get("assetdir"))
        if d is not None:
            meshdir = Path(directory) / d
        a = c.get("angle")
        if a is not None:
            ctx.degrees = a == "degree"
        s = c.get("eulerseq")
        if s is not None:
            if len(s) != 3:
                raise MjcfError("eulerseq must have 3 characters")
            ctx.eulerseq = s
        a = c.get("autolimits")
        if a is not None:
            ctx.autolimits = a == "true"
    ctx.defaults["main"] = {}
    for d in root.findall("default"):
        _parse_default(ctx, d, "main", True)
    bld = _Builder()
    for o in root.findall("option"):
        v = o.get("timestep")
        if v is not None:
            bld.timestep = parse_floats(v)[0]
        v = o.get("gravity")
        if v is not None:
            g = parse_floats(v)
            bld.gravity = (g[0], g[1], g[2])
        v = o.get("solver")
        if v is not None:
            bld.solver["kind"] = 1 if v == "PGS" else 0
        v = o.get("impratio")
        if v is not None:
            bld.solver["impratio"] = parse_floats(v)[0]
        v = o.get("tolerance")
        if v is not None:
            bld.solver["tolerance"] = parse_floats(v)[0]
        v = o.get("o_solref")
        if v is not None:
            r = parse_floats(v)
            bld.solver["solref"] = [r[0], r[1]]
        v = o.get("iterations")
        if v is not None:
            bld.solver["iterations"] = int(parse_floats(v)[0])
        v = o.get("integrator")
        if v is not None:
            bld.integrator = 1 if v in ("implicitfast", "implicit") else 0
    for a in root.findall("asset"):
        for mn in a.findall("mesh"):
            _parse_mesh(ctx, bld, mn, meshdir)
        for mat in a.findall("material"):
            cls = mat.get("class", "main")
            name = mat.get("name")
            rgba = ctx.floats(mat, cls, "rgba")
            if name is not None and rgba is not None and len(rgba) == 4:
                ctx.materials[name] = tuple(float(np.float32(x)) for x in rgba)
    bld.add_body("world", 0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))
    worlds = root.findall("worldbody")
    if not worlds:
        raise MjcfError("missing <worldbody>")
    for w in worlds:
        _parse_body_contents(ctx, bld, w, 0, "main")
    _finalize_inertia(bld)
    for t in root.findall("tendon"):
        for c in t:
            _parse_tendon(ctx, bld, c)
    for act in root.findall("actuator"):
        for a in act:
            _parse_actuator(ctx, bld, a)
    for e in root.findall("equality"):
        for c in e:
            _parse_equality(ctx, bld, c)
    for c in root.findall("contact"):
        for x in c:
            if x.tag == "exclude":
                ids = []
                for attr in ("body1", "body2"):
                    name = x.get(attr)
                    if name is None:
                        raise MjcfError(f"<exclude> needs {attr}")
                    if name not in bld.body["names"]:
                        raise MjcfError(f"<exclude> references unknown body '{name}'")
                    ids.append(bld.body["names"].index(name))
                bld.exclude.append(tuple(ids))
    return _to_arrays(bld)


def load_file(path: str | Path) -> dict:
    path = Path(path)
    return load_string(path.read_text(), path.parent)


def _mesh_arrays(meshes) -> dict:
    """The descriptor's hull arrays: per mesh its vertices and faces, flattened."""
    vadr, vnum, fadr, fnum, verts, fvadr, fvnum, fvert = [], [], [], [], [], [], [], []
    for _name, V, faces in meshes:
        vadr.append(len(verts))
        vnum.append(len(V))
        verts.extend(np.asarray(V, dtype=np.float64).reshape(-1, 3).tolist())
        fadr.append(len(fvadr))
        fnum.append(len(faces))
        for f in faces:
            fvadr.append(len(fvert))
            fvnum.append(len(f))
            fvert.extend(int(t) for t in f)
    i32 = lambda x: np.array(x, dtype=np.int32)
    return dict(nmesh=len(meshes), nmeshvert=len(verts), nmeshface=len(fvadr), nmeshfacevert=len(fvert),
                mesh_vertadr=i32(vadr), mesh_vertnum=i32(vnum), mesh_faceadr=i32(fadr), mesh_facenum=i32(fnum),
                mesh_vert=np.array(verts, dtype=np.float64).reshape(-1, 3), face_vertadr=i32(fvadr),
                face_vertnum=i32(fvnum), face_vert=i32(fvert))


def _to_arrays(bld: _Builder) -> dict:
    b, j, dof, g, cm, a, e = bld.body, bld.jnt, bld.dof, bld.geom, bld.cam, bld.act, bld.eq
    f64 = lambda x, w: np.array(x, dtype=np.float64).reshape(-1, w) if len(x) else np.zeros((0, w))
    i32 = lambda x: np.array(x, dtype=np.int32)
    madr, mnum, mj, mc = [], [], [], []
    for mom in a["moment"]:
        madr.append(len(mj))
        mnum.append(len(mom))
        for jj, cc in mom:
            mj.append(jj)
            mc.append(cc)
    s = bld.solver
    return dict(
        nq=bld.nq, nv=bld.nv, nu=len(a["gear"]), nbody=len(b["names"]), njnt=len(j["type"]),
        ngeom=len(g["type"]), ncam=len(cm["body"]), neq=len(e["j1"]), nexclude=len(bld.exclude),
        nmoment=len(mj),
        timestep=bld.timestep, gravity=np.array(bld.gravity), integrator=bld.integrator,
        solver_kind=s["kind"], solver_iterations=s["iterations"], tolerance=s["tolerance"],
        contact_margin=s["contact_margin"], solref=np.array(s["solref"]), solimp=np.array(s["solimp"]),
        impratio=s["impratio"], contact_hertz=s["contact_hertz"],
        contact_damping_ratio=s["contact_damping_ratio"], contact_slop=s["contact_slop"],
        body_parent=i32(b["parent"]), body_pos=f64(b["pos"], 3), body_quat=f64(b["quat"], 4),
        body_jntadr=i32(b["jntadr"]), body_jntnum=i32(b["jntnum"]), body_dofadr=i32(b["dofadr"]),
        body_dofnum=i32(b["dofnum"]), body_mass=np.array(b["mass"], dtype=np.float64),
        body_ipos=f64(b["ipos"], 3), body_inertia=f64(b["inertia"], 9),
        jnt_type=i32(j["type"]), jnt_body=i32(j["body"]), jnt_qposadr=i32(j["qposadr"]),
        jnt_dofadr=i32(j["dofadr"]), jnt_pos=f64(j["pos"], 3), jnt_axis=f64(j["axis"], 3),
        jnt_limited=i32(j["limited"]), jnt_range=f64(j["range"], 2),
        jnt_stiffness=np.array(j["stiffness"], dtype=np.float64),
        jnt_springref=np.array(j["springref"], dtype=np.float64),
        dof_body=i32(dof["body"]), dof_jnt=i32(dof["jnt"]),
        dof_damping=np.array(dof["damping"], dtype=np.float64),
        dof_armature=np.array(dof["armature"], dtype=np.float64),
        geom_type=i32(g["type"]), geom_body=i32(g["body"]), geom_pos=f64(g["pos"], 3),
        geom_quat=f64(g["quat"], 4), geom_size=f64(g["size"], 3),
        geom_friction=np.array(g["friction"], dtype=np.float64),
        geom_contype=np.array(g["contype"], dtype=np.uint32),
        geom_conaffinity=np.array(g["conaffinity"], dtype=np.uint32),
        geom_rgba=np.array(g["rgba"], dtype=np.float32).reshape(-1, 4), geom_group=i32(g["group"]),
        geom_dataid=i32(g["dataid"]),
        cam_body=i32(cm["body"]), cam_pos=f64(cm["pos"], 3), cam_quat=f64(cm["quat"], 4),
        cam_fovy=np.array(cm["fovy"], dtype=np.float64),
        actuator_gear=np.array(a["gear"], dtype=np.float64),
        actuator_gain=np.array(a["gain"], dtype=np.float64), actuator_bias=f64(a["bias"], 3),
        actuator_ctrllimited=i32(a["ctrllimited"]), actuator_ctrlrange=f64(a["ctrlrange"], 2),
        actuator_forcelimited=i32(a["forcelimited"]), actuator_forcerange=f64(a["forcerange"], 2),
        actuator_moment_adr=i32(madr), actuator_moment_num=i32(mnum), moment_jnt=i32(mj),
        moment_coef=np.array(mc, dtype=np.float64),
        eq_joint1=i32(e["j1"]), eq_joint2=i32(e["j2"]), eq_polycoef=f64(e["poly"], 5),
        eq_solref=f64(e["solref"], 2), eq_solimp=f64(e["solimp"], 5),
        exclude_pairs=np.array(bld.exclude, dtype=np.int32).reshape(-1, 2),
        qpos0=np.array(bld.qpos0, dtype=np.float64), qpos_reset=np.array(bld.qpos0, dtype=np.float64),
        **_mesh_arrays(bld.meshes),
    )
