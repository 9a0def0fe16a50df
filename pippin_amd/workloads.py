"""Synthetic workloads shared by bench.py and the tests (no reference counterpart)."""
from __future__ import annotations


def random_pile_xml(seed: int, general: bool = False) -> str:
    """2 to 6 free primitives (box, sphere, capsule, cylinder) of random sizes
    and orientations dropped in a column onto a plane, after the reference's
    tabletop tumble scenes. With general=False cylinders collide with the plane
    only; with general=True they also meet every other body, through the
    general convex narrow phase."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nb = int(rng.integers(2, 7))
    bodies = []
    for i in range(nb):
        t = rng.choice(["box", "sphere", "capsule", "cylinder"])
        x, y = rng.uniform(-0.08, 0.08, 2)
        z = 0.05 + 0.07 * i
        e = rng.uniform(0, 180, 3)
        if t == "box":
            g = '<geom type="box" size="%.4f %.4f %.4f"/>' % tuple(rng.uniform(0.015, 0.035, 3))
        elif t == "sphere":
            g = '<geom type="sphere" size="%.4f"/>' % rng.uniform(0.015, 0.035)
        elif t == "capsule":
            g = '<geom type="capsule" size="%.4f %.4f"/>' % (rng.uniform(0.015, 0.03), rng.uniform(0.02, 0.05))
        else:
            mask = "" if general else ' contype="2" conaffinity="0"'
            g = ('<geom type="cylinder" size="%.4f %.4f"%s/>'
                 % (rng.uniform(0.015, 0.03), rng.uniform(0.02, 0.05), mask))
        bodies.append('<body pos="%.4f %.4f %.4f" euler="%.1f %.1f %.1f"><freejoint/>%s</body>'
                      % (x, y, z, e[0], e[1], e[2], g))
    return ('<mujoco><option timestep="0.004"/><worldbody>'
            '<geom type="plane" size="5 5 .1" contype="3" conaffinity="3"/>' + "".join(bodies) +
            "</worldbody></mujoco>")



def arm_gripper_xml(seed: int) -> str:
    """A 7-DoF serial arm (hinge joints alternating yaw and pitch, link masses
    U[0.5, 3] kg, position actuators kp = 100) with a two-finger gripper whose
    slide fingers are coupled by a fixed tendon and a joint equality, over a
    cube (half-extent U[0.02, 0.03] m, friction U[0.6, 1.0]) resting on a
    table; implicitfast. The reference ships no model file for this arm
    workload, so it is generated here."""
    import numpy as np

    rng = np.random.default_rng(seed)
    mass = rng.uniform(0.5, 3.0, 7)
    half = rng.uniform(0.02, 0.03)
    fric = rng.uniform(0.6, 1.0)
    axes = ["0 0 1", "0 1 0", "0 0 1", "0 1 0", "0 0 1", "0 1 0", "0 0 1"]
    length = [0.10, 0.16, 0.06, 0.14, 0.06, 0.10, 0.04]
    # a folded default pose that puts the hand above the cube
    links, close = "", ""
    for i in range(7):
        d = length[i]
        links += (f'<body name="l{i}" pos="0 0 {0.0 if i == 0 else length[i - 1]:.3f}">'
                  f'<joint name="j{i}" axis="{axes[i]}" range="-2.5 2.5" limited="true"/>'
                  f'<geom type="capsule" fromto="0 0 0 0 0 {d:.3f}" size="0.025" mass="{mass[i]:.4f}" '
                  f'contype="0" conaffinity="0"/>')
        close += "</body>"
    hand = ('<body name="palm" pos="0 0 0.04"><geom type="box" size="0.035 0.045 0.01" mass="0.2"/>'
            '<body name="f1" pos="0 0.03 0.04"><joint name="finger1" type="slide" axis="0 1 0" range="-0.03 0.03" '
            'limited="true"/><geom type="box" size="0.01 0.005 0.03" mass="0.05"/></body>'
            '<body name="f2" pos="0 -0.03 0.04"><joint name="finger2" type="slide" axis="0 -1 0" range="-0.03 0.03" '
            'limited="true"/><geom type="box" size="0.01 0.005 0.03" mass="0.05"/></body></body>')
    acts = "".join(f'<position joint="j{i}" kp="100" kv="5"/>' for i in range(7))
    return ('<mujoco model="arm7_gripper"><option timestep="0.004" integrator="implicitfast"/>'
            '<default><joint damping="1" armature="0.05"/><geom friction="0.8"/></default><worldbody>'
            '<geom type="plane" size="2 2 .1"/>'
            '<geom name="table" type="box" size="0.4 0.4 0.1" pos="0.35 0 0.1"/>'
            f'<body name="base" pos="0 0 0.2" euler="0 110 0">{links}{hand}{close}</body>'
            f'<body name="cube" pos="0.42 0 {0.2 + half:.4f}"><freejoint/>'
            f'<geom type="box" size="{half:.4f} {half:.4f} {half:.4f}" friction="{fric:.3f}" density="800"/></body>'
            '</worldbody>'
            '<tendon><fixed name="grip"><joint joint="finger1" coef="0.5"/><joint joint="finger2" coef="0.5"/>'
            '</fixed></tendon>'
            '<equality><joint joint1="finger2" joint2="finger1" polycoef="0 1 0 0 0"/></equality>'
            f'<actuator>{acts}<position tendon="grip" kp="50" kv="1" ctrlrange="-0.03 0.03" ctrllimited="true"/>'
            '</actuator></mujoco>')


def arm_env_params(arrays: dict, n: int, seed: int):
    """Per-env draws for the arm configuration: link masses U[0.5, 3] kg for the
    seven links (body_mass scales each link's inertia with it, params.rs:145-155),
    cube half-extent U[0.02, 0.03] m with the cube resting on the table
    (qpos0), cube friction U[0.6, 1.0]. Returns ([(param, id, values (n, width))],
    qpos0 (n, nq)); apply the params, then qpos0, then reset."""
    import numpy as np

    rng = np.random.default_rng(seed + 7919)
    nbody = int(arrays["nbody"])
    cube_geom, cube_body = int(arrays["ngeom"]) - 1, nbody - 1
    links = range(2, 9)  # l0..l6 after the world and the massless base
    mass = rng.uniform(0.5, 3.0, (n, 7))
    half = rng.uniform(0.02, 0.03, n)
    fric = rng.uniform(0.6, 1.0, n)
    params = [("body_mass", b, mass[:, i:i + 1]) for i, b in enumerate(links)]
    params.append(("geom_size", cube_geom, np.stack([half, half, half], axis=1)))
    params.append(("geom_friction", cube_geom, fric[:, None]))
    q0 = np.tile(np.asarray(arrays["qpos0"], dtype=np.float64), (n, 1))
    za = int(np.asarray(arrays["jnt_qposadr"])[int(np.asarray(arrays["body_jntadr"])[cube_body])]) + 2
    q0[:, za] = 0.2 + half  # table top at 0.2
    return params, q0


def arm_targets(seed: int, n: int, steps: int, nu: int, every: int = 50):
    """Joint-space targets along random smooth waypoints (cosine blends
    between waypoints drawn every `every` steps), per env; the last actuator
    (the gripper) opens and closes."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nw = steps // every + 2
    way = rng.uniform(-0.6, 0.6, size=(nw, n, nu))
    way[:, :, -1] = rng.uniform(-0.03, 0.03, size=(nw, n))
    t = np.arange(steps)
    k, f = t // every, (t % every) / every
    w = 0.5 - 0.5 * np.cos(np.pi * f)
    return way[k] * (1 - w)[:, None, None] + way[k + 1] * w[:, None, None]


# A wedge with its three edges bevelled: a hexagonal section in the x-z plane
# (x, z), extruded along y: 12 hull vertices, six rectangular and two
# hexagonal faces.
WEDGE_SECTION = [(-0.035, -0.02), (0.03, -0.02), (0.04, -0.012), (-0.03, 0.025), (-0.04, 0.02), (-0.04, -0.01)]
WEDGE_HALF_WIDTH = 0.025
TUMBLE_KINDS = ("box", "sphere", "cylinder", "capsule", "wedge")
TUMBLE_SLOTS = 6


def wedge_vertices():
    return [(x, y, z) for (x, z) in WEDGE_SECTION for y in (-WEDGE_HALF_WIDTH, WEDGE_HALF_WIDTH)]


def solid_mass_properties(vertices, faces):
    """Volume, centre of mass and inertia about it (per unit density) of a
    closed convex polyhedron given as outward faces, by signed tetrahedra."""
    import numpy as np

    V = np.asarray(vertices, dtype=np.float64)
    vol, first, second = 0.0, np.zeros(3), np.zeros((3, 3))
    for f in faces:
        for i in range(1, len(f) - 1):
            a, b, d = V[f[0]], V[f[i]], V[f[i + 1]]
            v6 = float(a @ np.cross(b, d))
            vol += v6 / 6.0
            first += (a + b + d) * v6 / 24.0
            s3 = a + b + d
            second += v6 / 120.0 * (np.outer(a, a) + np.outer(b, b) + np.outer(d, d) + np.outer(s3, s3))
    com = first / vol
    inertia = np.trace(second) * np.eye(3) - second
    inertia -= vol * (float(com @ com) * np.eye(3) - np.outer(com, com))
    return vol, com, inertia


class TumblePlan:
    """A tabletop tumble: one model whose TUMBLE_SLOTS free bodies each hold one
    geom per object kind, and per env which kind each slot holds, its size,
    mass, friction and drop pose, applied through per-env parameters (the
    mechanism of the reference's python/pippin/scene.py:187-368: the chosen
    candidate collides, the others do not; unused slots are parked at rest on
    the floor far away)."""

    def __init__(self, xml, params, qpos0, kinds, counts):
        self.xml = xml
        self.params = params   # [(param, id, values (n, width))]
        self.qpos0 = qpos0     # (n, nq)
        self.kinds = kinds     # (n, slots): kind index, -1 parked
        self.counts = counts   # (n,): objects in play

    def apply(self, sim, begin=0, end=None):
        """Set every per-env parameter and reset the envs to their drop poses."""
        import numpy as np

        end = self.qpos0.shape[0] if end is None else end
        order = {"body_mass": 0, "body_ipos": 1, "body_inertia": 2}
        for name, gid, values in sorted(self.params, key=lambda p: order.get(p[0], 3)):
            sim.set_param(name, gid, begin, end, np.ascontiguousarray(values[begin:end]))
        sim.set_param("qpos0", 0, begin, end, np.ascontiguousarray(self.qpos0[begin:end]))
        sim.reset(np.arange(begin, end))


def tabletop_tumble(n: int, seed: int) -> TumblePlan:
    """Per env 2 to 6 objects drawn from boxes (half-extents U[0.015, 0.035]),
    spheres (r U[0.015, 0.035]), cylinders and capsules (r U[0.015, 0.03],
    half-height U[0.02, 0.05]) and the 12-vertex hull wedge, each with density
    U[500, 1500] kg/m^3 and friction U[0.5, 1.0], dropped from non-overlapping
    poses 0.05 to 0.2 m above a plane (random orientation)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    wv = wedge_vertices()
    vert = " ".join("%.6f %.6f %.6f" % v for v in wv)
    slot_geoms = ('<geom type="box" size="0.035 0.035 0.035"/><geom type="sphere" size="0.035"/>'
                  '<geom type="cylinder" size="0.03 0.05"/><geom type="capsule" size="0.03 0.05"/>'
                  '<geom type="mesh" mesh="wedge"/>')
    bodies = "".join(f'<body name="obj{s}" pos="{100 + 2 * s} 0 1"><freejoint/>'
                     f'<inertial pos="0 0 0" mass="0.1" diaginertia="1e-4 1e-4 1e-4"/>{slot_geoms}</body>'
                     for s in range(TUMBLE_SLOTS))
    xml = ('<mujoco model="tabletop_tumble"><option timestep="0.002"/>'
           f'<asset><mesh name="wedge" vertex="{vert}"/></asset><worldbody>'
           '<geom name="floor" type="plane" size="5 5 0.1"/>' + bodies + '</worldbody></mujoco>')
    # the wedge's solid properties from its hull (every point is extreme)
    from pippin_amd.physics import load_model

    arr = load_model(xml, is_xml=True)
    hv = np.asarray(arr["mesh_vert"]).reshape(-1, 3)
    fv, fa, fn = np.asarray(arr["face_vert"]), np.asarray(arr["face_vertadr"]), np.asarray(arr["face_vertnum"])
    wvol, wcom, wI = solid_mass_properties(hv, [list(fv[a:a + k]) for a, k in zip(fa, fn)])
    w_lo, w_r = float(hv[:, 2].min()), float(np.linalg.norm(hv - wcom, axis=1).max())
    nk = len(TUMBLE_KINDS)
    ng = 1 + TUMBLE_SLOTS * nk
    nq = 7 * TUMBLE_SLOTS
    contype = np.zeros((n, ng))
    size = np.zeros((n, ng, 3))
    friction = np.ones((n, ng))
    mass = np.zeros((n, TUMBLE_SLOTS + 1))
    inertia = np.zeros((n, TUMBLE_SLOTS + 1, 3))
    ipos = np.zeros((n, TUMBLE_SLOTS + 1, 3))
    qpos0 = np.zeros((n, nq))
    kinds = np.full((n, TUMBLE_SLOTS), -1)
    counts = np.zeros(n, dtype=np.int64)
    nominal = [(0.035, 0.035, 0.035), (0.035, 0, 0), (0.03, 0.05, 0), (0.03, 0.05, 0), (0, 0, 0)]
    for e in range(n):
        k = int(rng.integers(2, TUMBLE_SLOTS + 1))
        counts[e] = k
        placed = []
        for s in range(TUMBLE_SLOTS):
            g0 = 1 + s * nk
            active = s < k
            c = int(rng.integers(nk)) if active else 0
            dens, fr = rng.uniform(500.0, 1500.0), rng.uniform(0.5, 1.0)
            com = np.zeros(3)
            if c == 0:
                a, b, h = rng.uniform(0.015, 0.035, 3)
                m = dens * 8 * a * b * h
                I = (m * (b * b + h * h) / 3, m * (a * a + h * h) / 3, m * (a * a + b * b) / 3)
                sz, radius, rest = (a, b, h), float(np.sqrt(a * a + b * b + h * h)), h
            elif c == 1:
                r = rng.uniform(0.015, 0.035)
                m = dens * 4.0 / 3.0 * np.pi * r ** 3
                I = (0.4 * m * r * r,) * 3
                sz, radius, rest = (r, 0, 0), r, r
            elif c in (2, 3):
                r, hh = rng.uniform(0.015, 0.03), rng.uniform(0.02, 0.05)
                if c == 2:
                    m = dens * np.pi * r * r * 2 * hh
                    ix = m * (3 * r * r + 4 * hh * hh) / 12
                    I = (ix, ix, m * r * r / 2)
                    radius = float(np.hypot(r, hh))
                else:
                    mc, ms = dens * np.pi * r * r * 2 * hh, dens * 4.0 / 3.0 * np.pi * r ** 3
                    m = mc + ms
                    # cylinder plus two hemispheres (centre of each 3r/8 beyond the cap)
                    ix = (mc * (3 * r * r + 4 * hh * hh) / 12
                          + ms * (0.4 * r * r + hh * hh + 0.75 * hh * r))
                    I = (ix, ix, mc * r * r / 2 + 0.4 * ms * r * r)
                    radius = r + hh
                sz, rest = (r, hh, 0), hh
            else:
                m = dens * wvol
                I = tuple(dens * np.diag(wI))
                com = wcom
                sz, radius, rest = (0, 0, 0), w_r, -w_lo
            for ci in range(nk):
                on = ci == c if active else ci == 0
                contype[e, g0 + ci] = 1.0 if on else 0.0
                size[e, g0 + ci] = sz if ci == c else nominal[ci]
                friction[e, g0 + ci] = fr
            mass[e, s + 1] = m
            inertia[e, s + 1] = I
            ipos[e, s + 1] = com
            if active:
                kinds[e, s] = c
                for _ in range(1000):
                    x, y = rng.uniform(-0.12, 0.12, 2)
                    z = radius + rng.uniform(0.05, 0.2)
                    if all((x - px) ** 2 + (y - py) ** 2 + (z - pz) ** 2 > (radius + pr + 0.005) ** 2
                           for px, py, pz, pr in placed):
                        break
                placed.append((x, y, z, radius))
                qv = rng.normal(size=4)
                qv /= np.linalg.norm(qv)
                qpos0[e, 7 * s:7 * s + 7] = (x, y, z, *qv)
            else:
                qpos0[e, 7 * s:7 * s + 7] = (100 + 2 * s + 0.01 * e, 0.0, sz[2] + 1e-3, 1, 0, 0, 0)
    params = [("geom_contype", g, contype[:, g:g + 1]) for g in range(1, ng)]
    params += [("geom_conaffinity", g, contype[:, g:g + 1]) for g in range(1, ng)]
    params += [("geom_size", g, size[:, g]) for g in range(1, ng) if (g - 1) % nk != nk - 1]
    params += [("geom_friction", g, friction[:, g:g + 1]) for g in range(1, ng)]
    params += [("body_mass", b, mass[:, b:b + 1]) for b in range(1, TUMBLE_SLOTS + 1)]
    params += [("body_inertia", b, inertia[:, b]) for b in range(1, TUMBLE_SLOTS + 1)]
    params += [("body_ipos", b, ipos[:, b]) for b in range(1, TUMBLE_SLOTS + 1)]
    return TumblePlan(xml, params, qpos0, kinds, counts)
