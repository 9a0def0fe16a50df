"""Pins of oracle functions that no reference test covers (CPU only).

Each test states a fact the mathematics or the physics fixes, independently of
how the oracle computes it, so that a plausible slip in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails:

  velocity / rne (crates/pippin/src/forward.rs:115-171)
      double pendulum Coriolis and centrifugal terms at nonzero joint
      velocity (Christoffel symbols of the closed-form mass matrix); a
      spinning, translating free body with an offset centre of mass: the
      acceleration from M and c is Euler's rigid-body equations and a
      centre of mass in free fall.
  sphere_sphere, sphere-capsule, capsule_capsule (collision.rs:59-96, 269-335)
      against closest points found by dense sampling of the segments.
  add_implicit_terms (forward.rs:407-431)
      one implicitfast step of slide joints with tendon velocity actuators
      against (M + h(D + sum kv c c^T restricted to tree-ancestor pairs))^-1,
      with M from the kinetic energy; a force-clamped actuator adds nothing.
  actuation through a fixed tendon (forward.rs:185-211)
      a motor on a tendon distributes gear * coef * ctrl over its joints.
  joint equality rows (newton.rs:184-218)
      polycoef [0, c1] at static equilibrium: the constraint transmits the
      torque with the ratio c1, and the violation stays inside the bound of
      the soft-constraint law.
  set_const (forward.rs:275-333)
      a free body's dof_invweight from the inverse of its kinetic-energy
      mass matrix; its body_invweight equals (1/m, trace(I_c^-1)/3) for any
      centre-of-mass offset; a hinge pendulum's values in closed form.
"""
import math

import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim, collide

SPHERE, CAPSULE = 1, 2


def sim_xml(xml, n=1, solver=0):
    arr = mjcf.load_string(xml)
    arr["solver_kind"] = solver
    return OracleSim(arr, n)


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ]), q


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


# ---------------------------------------------------------------- velocity terms
def test_double_pendulum_coriolis_and_centrifugal_terms():
    """c(q, qdot) = Gamma(q)[qdot, qdot] + g(q) with the Christoffel symbols of
    M(q) = [[a + 2h cos t2, d + h cos t2], [d + h cos t2, d]], h = m2 l1 c2."""
    m1, m2, l1, c1, c2, i1, i2 = 1.3, 0.7, 0.5, 0.2, 0.3, 0.011, 0.006
    xml = f"""<mujoco><worldbody>
      <body name="a"><joint axis="0 1 0"/><inertial pos="{c1} 0 0" mass="{m1}" diaginertia="0.01 {i1} 0.01"/>
        <body name="b" pos="{l1} 0 0"><joint axis="0 1 0"/><inertial pos="{c2} 0 0" mass="{m2}" diaginertia="0.01 {i2} 0.01"/></body>
      </body></worldbody></mujoco>"""
    n = 8
    s = sim_xml(xml, n)
    rng = np.random.default_rng(7)
    q = rng.uniform(-2, 2, size=(n, 2))
    qd = rng.uniform(-3, 3, size=(n, 2))
    s.set("qpos", 0, n, q)
    s.set("qvel", 0, n, qd)
    _, bias = s.mass_bias()
    g, h = 9.81, m2 * l1 * c2
    for e in range(n):
        t1, t2 = q[e]
        w1, w2 = qd[e]
        cor1 = -h * math.sin(t2) * (2 * w1 * w2 + w2 * w2)
        cor2 = h * math.sin(t2) * w1 * w1
        g1 = -g * (m1 * c1 * math.cos(t1) + m2 * (l1 * math.cos(t1) + c2 * math.cos(t1 + t2)))
        g2 = -g * m2 * c2 * math.cos(t1 + t2)
        scale = abs(cor1) + abs(g1) + 1.0
        assert abs(bias[e, 0] - (cor1 + g1)) < 1e-12 * scale, (e, bias[e], cor1 + g1)
        assert abs(bias[e, 1] - (cor2 + g2)) < 1e-12 * scale, (e, bias[e], cor2 + g2)
    # the velocity part is quadratic: c(q, 2 qdot) - g = 4 (c(q, qdot) - g)
    s.set("qvel", 0, n, 2 * qd)
    _, bias2 = s.mass_bias()
    s.set("qvel", 0, n, 0 * qd)
    _, bias0 = s.mass_bias()
    assert np.allclose(bias2 - bias0, 4 * (bias - bias0), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("gravity", [0.0, -9.81])
def test_free_body_accelerations_are_euler_equations(gravity):
    """A free body with an offset centre of mass r and full inertia I_c about
    it: qacc = -M^-1 c must give omega_dot = -I_c^-1 (w x I_c w) in the body
    frame (torque-free Euler equations: angular momentum about the centre of
    mass is conserved) and a centre-of-mass acceleration equal to gravity,
    with the free joint's velocity = (world velocity of the body origin,
    angular velocity in the body frame)."""
    r = np.array([0.07, -0.03, 0.05])
    I = np.array([[0.021, 0.002, -0.001], [0.002, 0.013, 0.003], [-0.001, 0.003, 0.017]])
    m = 1.7
    xml = f"""<mujoco><option gravity="0 0 {gravity}"/><worldbody>
      <body pos="0 0 1"><freejoint/>
        <inertial pos="{r[0]} {r[1]} {r[2]}" mass="{m}" fullinertia="{I[0,0]} {I[1,1]} {I[2,2]} {I[0,1]} {I[0,2]} {I[1,2]}"/>
      </body></worldbody></mujoco>"""
    n = 6
    s = sim_xml(xml, n)
    rng = np.random.default_rng(11)
    qpos, qvel, rots = np.zeros((n, 7)), rng.uniform(-4, 4, size=(n, 6)), []
    for e in range(n):
        R, quat = rand_rot(rng)
        qpos[e, :3] = rng.normal(size=3)
        qpos[e, 3:] = quat
        rots.append(R)
    s.set("qpos", 0, n, qpos)
    s.set("qvel", 0, n, qvel)
    qm, bias = s.mass_bias()
    g = np.array([0.0, 0.0, gravity])
    for e in range(n):
        R = rots[e]
        # mass matrix from the kinetic energy: v_com = v_o + R (w x r), w in the body frame
        Jc = np.hstack([np.eye(3), -R @ skew(r)])
        Jw = np.hstack([np.zeros((3, 3)), np.eye(3)])
        M = m * Jc.T @ Jc + Jw.T @ I @ Jw
        assert np.allclose(qm[e], M, rtol=0, atol=1e-14), e
        a = np.linalg.solve(qm[e], -bias[e])
        w = qvel[e, 3:]
        wdot = -np.linalg.solve(I, np.cross(w, I @ w))
        ao = g - R @ (np.cross(wdot, r) + np.cross(w, np.cross(w, r)))
        assert np.allclose(a[3:], wdot, rtol=0, atol=1e-12), (e, a[3:], wdot)
        assert np.allclose(a[:3], ao, rtol=0, atol=1e-12), (e, a[:3], ao)


# ---------------------------------------------------------------- sphere / capsule pairs
def _closest_on_segment(a, b, p, levels=3, k=2001):
    """Closest point of segment a-b to p by nested dense sampling."""
    lo, hi = 0.0, 1.0
    for _ in range(levels):
        t = np.linspace(lo, hi, k)
        pts = a[None] + (b - a)[None] * t[:, None]
        i = int(np.argmin(np.sum((pts - p) ** 2, axis=1)))
        step = (hi - lo) / (k - 1)
        lo, hi = max(0.0, t[i] - 2 * step), min(1.0, t[i] + 2 * step)
    return pts[i]


def _closest_segments(p0, p1, q0, q1, levels=4, k=801):
    """Closest points of two segments by nested dense sampling of (s, t)."""
    box = [0.0, 1.0, 0.0, 1.0]
    for _ in range(levels):
        s = np.linspace(box[0], box[1], k)
        t = np.linspace(box[2], box[3], k)
        P = p0[None] + (p1 - p0)[None] * s[:, None]
        Q = q0[None] + (q1 - q0)[None] * t[:, None]
        d2 = np.sum((P[:, None, :] - Q[None, :, :]) ** 2, axis=2)
        i, j = np.unravel_index(int(np.argmin(d2)), d2.shape)
        ss, st = (box[1] - box[0]) / (k - 1), (box[3] - box[2]) / (k - 1)
        box = [max(0.0, s[i] - 2 * ss), min(1.0, s[i] + 2 * ss), max(0.0, t[j] - 2 * st), min(1.0, t[j] + 2 * st)]
    return P[i], Q[j]


def _check_hit(hit, x, y, ra, rb, tol_pos=1e-7):
    """A hit between spheres of radii ra, rb centred at x (first geom) and y:
    depth = ra + rb - |y - x|, normal from x to y, position midway between the
    two surface points along the normal."""
    d = np.linalg.norm(y - x)
    n = (y - x) / d
    assert abs(hit["depth"] - (ra + rb - d)) < 1e-9, (hit["depth"], ra + rb - d)
    assert np.allclose(hit["normal"], n, atol=1e-6), (hit["normal"], n)
    mid = 0.5 * ((x + ra * n) + (y - rb * n))
    assert np.allclose(hit["pos"], mid, atol=tol_pos), (hit["pos"], mid)


def _capsule_pose(rng, c=None):
    R, _ = rand_rot(rng)
    return (rng.normal(size=3) * 0.05 if c is None else c), R


def test_sphere_sphere_against_geometry():
    rng = np.random.default_rng(3)
    eye = np.eye(3)
    hits = 0
    for _ in range(60):
        c1, c2 = rng.normal(size=3) * 0.06, rng.normal(size=3) * 0.06
        r1, r2 = rng.uniform(0.02, 0.12, size=2)
        out = collide(SPHERE, c1, eye, [r1, 0, 0], SPHERE, c2, eye, [r2, 0, 0], margin=1e-3)
        sep = np.linalg.norm(c2 - c1) - r1 - r2
        assert len(out) == (1 if sep < 1e-3 else 0), (sep, len(out))
        if len(out):
            _check_hit(out[0], c1, c2, r1, r2, tol_pos=1e-12)
            hits += 1
    assert hits > 20


@pytest.mark.parametrize("swap", [False, True])
def test_sphere_capsule_against_sampled_segment(swap):
    rng = np.random.default_rng(5 + swap)
    hits = 0
    for _ in range(40):
        cpos, R = _capsule_pose(rng)
        rc, hl = rng.uniform(0.02, 0.06), rng.uniform(0.03, 0.12)
        a, b = cpos - R[:, 2] * hl, cpos + R[:, 2] * hl
        sc = cpos + rng.normal(size=3) * 0.08
        rs = rng.uniform(0.02, 0.06)
        y = _closest_on_segment(a, b, sc)
        sep = np.linalg.norm(y - sc) - rs - rc
        if abs(sep + 1e-3) < 1e-6:
            continue  # too close to the margin for a sampled verdict
        if swap:  # capsule first: the normal points from the capsule to the sphere
            out = collide(CAPSULE, cpos, R, [rc, hl, 0], SPHERE, sc, np.eye(3), [rs, 0, 0])
        else:
            out = collide(SPHERE, sc, np.eye(3), [rs, 0, 0], CAPSULE, cpos, R, [rc, hl, 0])
        assert len(out) == (1 if sep < 1e-3 else 0), (sep, len(out))
        if len(out):
            if swap:
                _check_hit(out[0], y, sc, rc, rs)
            else:
                _check_hit(out[0], sc, y, rs, rc)
            hits += 1
    assert hits > 10


def test_capsule_capsule_against_sampled_segments():
    """The first hit is the pair of closest points; near-parallel capsules add
    endpoint hits, each from an end of the first capsule to its closest point
    on the second."""
    rng = np.random.default_rng(9)
    hits = extra = 0
    for trial in range(50):
        pa, Ra = _capsule_pose(rng)
        pb, Rb = _capsule_pose(rng)
        if trial % 3 == 0:  # near-parallel: the second axis close to the first
            Rb = Ra.copy()
            Rb[:, 2] = Ra[:, 2] + rng.normal(size=3) * 0.01
            Rb[:, 2] /= np.linalg.norm(Rb[:, 2])
            pb = pa + Ra[:, 0] * rng.uniform(0.03, 0.1) + Ra[:, 2] * rng.uniform(-0.05, 0.05)
        ra, rb = rng.uniform(0.02, 0.05, size=2)
        ha, hb = rng.uniform(0.04, 0.12, size=2)
        p0, p1 = pa - Ra[:, 2] * ha, pa + Ra[:, 2] * ha
        q0, q1 = pb - Rb[:, 2] * hb, pb + Rb[:, 2] * hb
        x, y = _closest_segments(p0, p1, q0, q1)
        sep = np.linalg.norm(y - x) - ra - rb
        if abs(sep + 1e-3) < 1e-6:
            continue
        out = collide(CAPSULE, pa, Ra, [ra, ha, 0], CAPSULE, pb, Rb, [rb, hb, 0])
        if sep >= 1e-3:
            assert len(out) == 0 or all(h["depth"] > -1e-3 for h in out)
            continue
        assert len(out) >= 1
        _check_hit(out[0], x, y, ra, rb, tol_pos=1e-6)
        hits += 1
        for h in out[1:]:
            # recover the two centres of the endpoint hit from its geometry
            n = h["normal"]
            xe = h["pos"] - n * (ra - h["depth"] * 0.5)
            end = min((p0, p1), key=lambda e: np.linalg.norm(e - xe))
            assert np.linalg.norm(end - xe) < 1e-9, "endpoint hit does not start at an end of capsule a"
            ye = _closest_on_segment(q0, q1, end)
            _check_hit(h, end, ye, ra, rb)
            extra += 1
    assert hits > 15 and extra > 0, (hits, extra)


# ---------------------------------------------------------------- implicitfast, tendons
IMPLICIT_XML = """<mujoco><option timestep="{h}" gravity="0 0 0" integrator="implicitfast"/>
  <worldbody>
    <body name="A"><joint name="j1" type="slide" axis="1 0 0" damping="{d1}" armature="0.1"/>
      <inertial pos="0 0 0" mass="{mA}" diaginertia="1 1 1"/>
      <body name="B"><joint name="j2" type="slide" axis="1 1 0" armature="0.05"/>
        <inertial pos="0 0 0" mass="{mB}" diaginertia="1 1 1"/></body>
    </body>
    <body name="C"><joint name="j3" type="slide" axis="0 0 1"/>
      <inertial pos="0 0 0" mass="{mC}" diaginertia="1 1 1"/></body>
  </worldbody>
  <tendon>
    <fixed name="t1"><joint joint="j1" coef="{c11}"/><joint joint="j2" coef="{c12}"/></fixed>
    <fixed name="t2"><joint joint="j2" coef="{c22}"/><joint joint="j3" coef="{c23}"/></fixed>
  </tendon>
  <actuator>
    <velocity tendon="t1" kv="{kv1}"/>
    <velocity tendon="t2" kv="{kv2}" {limit}/>
  </actuator></mujoco>"""


def _slide_model(clamp):
    p = dict(h=0.01, d1=0.5, mA=1.3, mB=0.8, mC=2.1, c11=1.0, c12=0.5, c22=0.7, c23=-0.4, kv1=3.0, kv2=2.0,
             limit='forcelimited="true" forcerange="-0.01 0.01"' if clamp else "")
    return p, IMPLICIT_XML.format(**p)


@pytest.mark.parametrize("clamp", [False, True])
def test_implicitfast_step_of_tendon_velocity_actuators(clamp):
    """v' = v + h (M + h K)^-1 f, where K = damping + kv c c^T per actuator,
    keeping only dof pairs on one tree path (j1-j2: ancestor; j2-j3: sibling
    branches, coupling dropped), and nothing for a force-clamped actuator."""
    p, xml = _slide_model(clamp)
    s = sim_xml(xml)
    rng = np.random.default_rng(1)
    v = rng.uniform(-1, 1, size=3)
    u = np.array([0.7, -0.4])
    s.set("qvel", 0, 1, v[None])
    s.set("ctrl", 0, 1, u[None])
    s.step(0, 1, 1)
    h = p["h"]
    e1, e2, e3 = np.array([1, 0, 0.0]), np.array([1, 1, 0.0]) / math.sqrt(2), np.array([0, 0, 1.0])
    # kinetic energy: A moves with j1; B with j1 and j2; C with j3
    M = np.zeros((3, 3))
    M[0, 0] = p["mA"] + p["mB"] + 0.1
    M[1, 1] = p["mB"] + 0.05
    M[0, 1] = M[1, 0] = p["mB"] * e1 @ e2
    M[2, 2] = p["mC"]
    c1 = np.array([p["c11"], p["c12"], 0.0])
    c2 = np.array([0.0, p["c22"], p["c23"]])
    f1 = p["kv1"] * (u[0] - c1 @ v)
    f2 = p["kv2"] * (u[1] - c2 @ v)
    if clamp:
        assert abs(f2) > 0.01
        f2 = float(np.clip(f2, -0.01, 0.01))
    f = -np.array([p["d1"], 0, 0]) * v + c1 * f1 + c2 * f2
    K = np.diag([p["d1"], 0, 0]) + p["kv1"] * np.outer(c1, c1)
    if not clamp:
        K2 = p["kv2"] * np.outer(c2, c2)
        K2[1, 2] = K2[2, 1] = 0.0  # j2 (under A) and j3 (under C) are on sibling branches
        K += K2
    v_new = v + h * np.linalg.solve(M + h * K, f)
    got = s.get("qvel")[0]
    assert np.allclose(got, v_new, rtol=0, atol=1e-13), (got, v_new)
    # the dropped sibling coupling is not negligible here: keeping it changes v'
    K3 = K + p["kv2"] * np.outer(c2, c2) * np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]])
    if not clamp:
        assert np.max(np.abs(v + h * np.linalg.solve(M + h * K3, f) - got)) > 1e-6


def test_tendon_motor_distributes_gear_coef_ctrl():
    """A motor on a fixed tendon (coefs 1.5, -0.5, gear 2) applies
    gear * coef * ctrl to each joint: one step from rest, contact-free, gives
    v = h M^-1 (gear coef ctrl)."""
    xml = """<mujoco><option timestep="0.01" gravity="0 0 0"/><worldbody>
      <body><joint name="a" type="slide" axis="1 0 0"/><inertial pos="0 0 0" mass="2" diaginertia="1 1 1"/></body>
      <body><joint name="b" type="hinge" axis="0 0 1"/><inertial pos="0.2 0 0" mass="0.5" diaginertia="0.01 0.01 0.02"/></body>
    </worldbody>
    <tendon><fixed name="t"><joint joint="a" coef="1.5"/><joint joint="b" coef="-0.5"/></fixed></tendon>
    <actuator><motor tendon="t" gear="2"/></actuator></mujoco>"""
    s = sim_xml(xml)
    s.set("ctrl", 0, 1, [[0.8]])
    s.step(0, 1, 1)
    M = np.diag([2.0, 0.02 + 0.5 * 0.2 ** 2])
    tau = 2 * np.array([1.5, -0.5]) * 0.8
    assert np.allclose(s.get("qvel")[0], 0.01 * np.linalg.solve(M, tau), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- joint equality
@pytest.mark.parametrize("ratio", [1.0, 2.0])
def test_joint_equality_transmits_torque_at_equilibrium(ratio):
    """(q1 - q1_0) = ratio (q2 - q2_0) via polycoef [0, ratio]. A spring k1 on
    joint 1 and a constant torque tau on joint 2 settle where the row's force
    lambda balances both: J = [1, -ratio], so tau = ratio k1 q1, i.e.
    q1 = tau / (ratio k1). The violation q1 - ratio q2 obeys the soft-constraint
    law at rest (a = v = 0): lambda = D k imp pos, D = imp / ((1 - imp) diag),
    so |pos| <= k1 |q1| (1 - dmin) diag / (dmin^2 k), the stiffest case."""
    k1, tau = 4.0, 0.3
    xml = f"""<mujoco><option timestep="0.002" gravity="0 0 0"/><worldbody>
      <body><joint name="j1" axis="0 0 1" stiffness="{k1}" damping="2"/>
        <inertial pos="0.1 0 0" mass="0.5" diaginertia="0.01 0.01 0.01"/></body>
      <body><joint name="j2" axis="0 0 1" damping="2"/>
        <inertial pos="0.1 0 0" mass="0.4" diaginertia="0.01 0.01 0.01"/></body>
    </worldbody>
    <equality><joint joint1="j1" joint2="j2" polycoef="0 {ratio} 0 0 0"/></equality>
    <actuator><motor joint="j2" gear="1"/></actuator></mujoco>"""
    s = sim_xml(xml)
    s.set("ctrl", 0, 1, [[tau]])
    s.step(0, 1, 20000)
    q1, q2 = s.get("qpos")[0]
    assert np.linalg.norm(s.get("qvel")[0]) < 1e-8
    assert abs(q1 - tau / (ratio * k1)) < 1e-7 * abs(q1), (q1, tau / (ratio * k1))
    dof_iw, _ = s.invweight()
    diag = dof_iw[0, 0] + dof_iw[0, 1]
    dmin, dmax, tc, dr = 0.9, 0.95, 0.02, 1.0  # MuJoCo's default equality solimp / solref
    kref = 1.0 / (dmax * dmax * tc * tc * dr * dr)
    bound = k1 * abs(q1) * (1 - dmin) * diag / (dmin * dmin * kref)
    pos = q1 - ratio * q2
    assert 0 < abs(pos) <= bound, (pos, bound)
    assert np.sign(pos) == np.sign(-k1 * q1)  # the row pulls joint 1 against its spring


def test_joint_equality_pgs_holds_the_ratio():
    xml = """<mujoco><option timestep="0.002" gravity="0 0 0"/><worldbody>
      <body><joint name="j1" axis="0 0 1" stiffness="4" damping="2"/>
        <inertial pos="0.1 0 0" mass="0.5" diaginertia="0.01 0.01 0.01"/></body>
      <body><joint name="j2" axis="0 0 1" damping="2"/>
        <inertial pos="0.1 0 0" mass="0.4" diaginertia="0.01 0.01 0.01"/></body>
    </worldbody>
    <equality><joint joint1="j1" joint2="j2" polycoef="0 2 0 0 0"/></equality>
    <actuator><motor joint="j2" gear="1"/></actuator></mujoco>"""
    s = sim_xml(xml, solver=1)
    s.set("ctrl", 0, 1, [[0.3]])
    s.step(0, 1, 6000)
    q1, q2 = s.get("qpos")[0]
    assert abs(q1 - 0.3 / 8.0) < 2e-3 * 0.3 / 8.0
    assert abs(q1 - 2 * q2) < 5e-3 * abs(q1)


# ---------------------------------------------------------------- set_const
@pytest.mark.parametrize("offset", [[0, 0, 0], [0.07, -0.03, 0.05]])
def test_free_body_invweight(offset):
    r = np.array(offset, dtype=float)
    I = np.array([[0.021, 0.002, -0.001], [0.002, 0.013, 0.003], [-0.001, 0.003, 0.017]])
    m = 1.7
    xml = f"""<mujoco><worldbody>
      <body pos="0.3 0 1" quat="0.9 0.1 -0.3 0.2"><freejoint/>
        <inertial pos="{r[0]} {r[1]} {r[2]}" mass="{m}" fullinertia="{I[0,0]} {I[1,1]} {I[2,2]} {I[0,1]} {I[0,2]} {I[1,2]}"/>
      </body></worldbody></mujoco>"""
    s = sim_xml(xml)
    dof, body = s.invweight()
    quat = np.array([0.9, 0.1, -0.3, 0.2])
    quat /= np.linalg.norm(quat)
    w, x, y, z = quat
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    Jc = np.hstack([np.eye(3), -R @ skew(r)])
    Jw = np.hstack([np.zeros((3, 3)), np.eye(3)])
    Minv = np.linalg.inv(m * Jc.T @ Jc + Jw.T @ I @ Jw)
    dg = np.diag(Minv)
    exp = np.r_[[dg[:3].mean()] * 3, [dg[3:].mean()] * 3]
    assert np.allclose(dof[0], exp, rtol=1e-12, atol=0), (dof[0], exp)
    # the mobility at the centre of mass does not depend on the parameterisation
    assert abs(body[0, 1, 0] - 1 / m) < 1e-12 / m
    assert abs(body[0, 1, 1] - np.trace(np.linalg.inv(I)) / 3) < 1e-12 * np.trace(np.linalg.inv(I))


def test_hinge_pendulum_invweight():
    """One hinge about z at the body origin, centre of mass at distance l:
    M = I_zz + m l^2 + armature; the com's mobility is |axis x l|^2 / M per
    axis direction, averaged over the three directions; rotation: 1 / (3 M)."""
    m, l, izz, arm = 0.9, 0.25, 0.013, 0.02
    xml = f"""<mujoco><worldbody><body pos="0 0 1">
      <joint axis="0 0 1" armature="{arm}"/>
      <inertial pos="{l} 0 0" mass="{m}" diaginertia="0.01 0.01 {izz}"/>
    </body></worldbody></mujoco>"""
    s = sim_xml(xml)
    dof, body = s.invweight()
    Mz = izz + m * l * l + arm
    assert abs(dof[0, 0] - 1 / Mz) < 1e-14
    assert abs(body[0, 1, 0] - l * l / Mz / 3) < 1e-14
    assert abs(body[0, 1, 1] - 1 / Mz / 3) < 1e-14
