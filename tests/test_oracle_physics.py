"""The oracle's engine against the reference's physical tests and closed forms (CPU only).

Restated from the reference:
  crates/pippin/tests/contact.rs      rest heights, box stack, incline stick/slide,
                                      joint limit, rollout_matches_step_loop, reset-state param
  crates/pippin-env/tests/params.rs   per-env friction, size and parking, mass ratio, rejection
Closed forms added: a double pendulum's mass matrix and gravity bias, a ballistic
free body under semi-implicit Euler.
"""
import math

import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleError, OracleSim

SOLVERS = {"newton": 0, "pgs": 1}


def sim_xml(xml, n=1, solver="newton"):
    arr = mjcf.load_string(xml)
    arr["solver_kind"] = SOLVERS[solver]
    return OracleSim(arr, n)


def run(s, seconds, dt):
    s.step(0, s.n, int(round(seconds / dt)))


def body_z(s, qadr):
    return s.get("qpos")[0, qadr + 2]


both = pytest.mark.parametrize("solver", ["newton", "pgs"])


# contact.rs:31-47
@both
def test_box_rests_on_plane(solver):
    s = sim_xml("""<mujoco><worldbody>
        <geom type="plane" size="5 5 .1"/>
        <body name="b" pos="0 0 0.3" euler="3 5 0"><freejoint/><geom type="box" size=".1 .1 .1"/></body>
    </worldbody></mujoco>""", solver=solver)
    run(s, 3.0, 0.002)
    q = s.get("qpos")[0]
    assert abs(q[2] - 0.1) < 2e-3, q[2]
    assert abs(q[0]) < 1e-2 and abs(q[1]) < 1e-2
    assert np.linalg.norm(s.get("qvel")[0]) < 1e-3


# contact.rs:49-62
@both
def test_sphere_and_capsule_rest(solver):
    s = sim_xml("""<mujoco><worldbody>
        <geom type="plane" size="5 5 .1"/>
        <body name="s" pos="0 0 0.5"><freejoint/><geom type="sphere" size=".05"/></body>
        <body name="c" pos="1 0 0.5" euler="90 0 0"><freejoint/><geom type="capsule" size=".04 .2"/></body>
    </worldbody></mujoco>""", solver=solver)
    run(s, 3.0, 0.002)
    q = s.get("qpos")[0]
    assert abs(q[2] - 0.05) < 2e-3 and abs(q[9] - 0.04) < 2e-3


# contact.rs:64-81
@both
def test_box_stack_is_stable(solver):
    xml = '<mujoco><option timestep="0.002"/><worldbody><geom type="plane" size="5 5 .1"/>'
    for i in range(5):
        xml += (f'<body name="b{i}" pos="0 0 {0.05 + 0.1 * i + 0.001 * i}"><freejoint/>'
                f'<geom type="box" size=".05 .05 .05"/></body>')
    xml += "</worldbody></mujoco>"
    s = sim_xml(xml, solver=solver)
    run(s, 5.0, 0.002)
    top = s.get("qpos")[0, 28:31]
    assert abs(top[0]) < 5e-3 and abs(top[1]) < 5e-3, top
    assert abs(top[2] - 0.45) < 1e-2, top


def incline_xml(deg, mu):
    return f"""<mujoco><option timestep="0.002"/><worldbody>
        <geom type="plane" size="5 5 .1" euler="0 {deg} 0" friction="{mu}"/>
        <body name="b" pos="0 0 0.1" euler="0 {deg} 0"><freejoint/><geom type="box" size=".05 .05 .05" friction="{mu}"/></body>
    </worldbody></mujoco>"""


# contact.rs:92-103
@both
def test_friction_holds_below_critical_angle(solver):
    s = sim_xml(incline_xml(20.0, 0.5), solver=solver)
    run(s, 0.5, 0.002)
    p0 = s.get("qpos")[0, :3]
    run(s, 1.0, 0.002)
    p1 = s.get("qpos")[0, :3]
    assert np.linalg.norm(p1 - p0) < 2e-3


# contact.rs:105-123
@both
def test_friction_slides_with_analytic_acceleration(solver):
    deg, mu = 35.0, 0.3
    s = sim_xml(incline_xml(deg, mu), solver=solver)
    run(s, 0.3, 0.002)
    v0 = np.linalg.norm(s.get("qvel")[0, :3])
    run(s, 0.5, 0.002)
    v1 = np.linalg.norm(s.get("qvel")[0, :3])
    t = math.radians(deg)
    expected = 9.81 * (math.sin(t) - mu * math.cos(t))
    assert abs((v1 - v0) / 0.5 - expected) / expected < 0.03


# contact.rs:125-144
@both
def test_joint_limit_is_respected(solver):
    s = sim_xml("""<mujoco><option timestep="0.002"/><worldbody>
        <body pos="0 0 1"><joint name="h" axis="0 1 0" range="-30 30"/>
          <geom type="capsule" fromto="0 0 0 0.5 0 0" size="0.03" contype="0" conaffinity="0"/></body>
    </worldbody></mujoco>""", solver=solver)
    qmax = 0.0
    for _ in range(2000):
        s.step(0, 1, 1)
        qmax = max(qmax, s.get("qpos")[0, 0])
    lim = math.radians(30)
    assert qmax < lim + 0.05
    assert abs(s.get("qpos")[0, 0] - lim) < 0.01


# contact.rs:146-163
@both
def test_sphere_capsule_box_on_box(solver):
    s = sim_xml("""<mujoco><worldbody>
        <geom type="box" size="1 1 .1" pos="0 0 -.1"/>
        <body name="s" pos="0 0 0.3"><freejoint/><geom type="sphere" size=".05"/></body>
        <body name="c" pos="0.5 0 0.3" euler="0 90 0"><freejoint/><geom type="capsule" size=".03 .1"/></body>
        <body name="bx" pos="-0.5 0 0.3" euler="0 0 30"><freejoint/><geom type="box" size=".05 .05 .05"/></body>
    </worldbody></mujoco>""", solver=solver)
    run(s, 3.0, 0.002)
    q = s.get("qpos")[0]
    assert abs(q[2] - 0.05) < 2e-3 and abs(q[9] - 0.03) < 2e-3 and abs(q[16] - 0.05) < 2e-3


# contact.rs:165-188
@pytest.mark.parametrize("name", ["cartpole.xml", "tumble.xml"])
def test_rollout_matches_step_loop(assets, name):
    arr = mjcf.load_file(assets / name)
    nu, nq, n, t = int(arr["nu"]), int(arr["nq"]), 3, 50
    ctrl = np.sin(np.arange(n * t * nu) * 0.37).reshape(n, t, nu)
    a = OracleSim(arr, n)
    out = a.rollout(ctrl)
    b = OracleSim(arr, n)
    for k in range(t):
        b.set("ctrl", 0, n, ctrl[:, k])
        b.step(0, n, 1)
    assert np.array_equal(out[:, t - 1], b.get("qpos"))


# contact.rs:190-206: the reset state param does not move the kinematic reference
def test_qpos0_param_is_reset_state_only():
    xml = """<mujoco><worldbody><body name="b"><joint name="h" axis="0 0 1" ref="30"/>
        <geom type="box" size=".1 .02 .02" pos=".1 0 0" contype="0" conaffinity="0"/></body></worldbody></mujoco>"""
    s = sim_xml(xml, 2)
    s.set_param("qpos0", 0, 1, 2, [1.0])
    s.reset([1])
    assert s.get("qpos")[1, 0] == 1.0
    s.set("qpos", 0, 1, [[1.0]])
    g, _ = s.poses()
    assert np.array_equal(g[0]["mat"], g[1]["mat"])


# ---- crates/pippin-env/tests/params.rs ----
INCLINE = """<mujoco><option timestep="0.002"/><worldbody>
    <geom name="floor" type="plane" size="5 5 .1" euler="0 25 0" friction="0.1"/>
    <body name="b" pos="0 0 0.1" euler="0 25 0"><freejoint/><geom name="box" type="box" size=".05 .05 .05" friction="0.1"/></body>
</worldbody></mujoco>"""


def test_friction_per_env():
    s = sim_xml(INCLINE, 2)
    s.set_param("geom_friction", 1, 1, 2, [0.9])
    s.set_param("geom_friction", 0, 1, 2, [0.9])
    s.step(0, 2, 500)
    q = s.get("qpos")
    slid = np.hypot(q[:, 0], q[:, 1])
    assert slid[0] > 0.1 and slid[1] < 0.01
    assert np.array_equal(s.get_param("geom_friction", 1)[:, 0], [0.1, 0.9])


def test_size_and_parking_per_env():
    s = sim_xml("""<mujoco><worldbody><geom type="plane" size="5 5 .1"/>
        <body pos="0 0 0.3"><freejoint/><geom name="ball" type="sphere" size=".05"/></body>
    </worldbody></mujoco>""", 3)
    s.set_param("geom_size", 1, 1, 2, [0.1, 0.0, 0.0])
    s.set_param("geom_contype", 1, 2, 3, [0.0])
    s.set_param("geom_conaffinity", 1, 2, 3, [0.0])
    s.step(0, 3, 1000)
    q = s.get("qpos")
    assert abs(q[0, 2] - 0.05) < 2e-3 and abs(q[1, 2] - 0.10) < 2e-3 and q[2, 2] < -1.0
    ap = s.appearance()
    assert ap[1, 1]["scale"][0] == 2.0 and ap[0, 1]["scale"][0] == 1.0


def test_mass_per_env_changes_actuated_motion():
    s = sim_xml("""<mujoco><option gravity="0 0 0"/><worldbody>
        <body name="link"><joint name="j" axis="0 0 1"/><geom type="capsule" fromto="0 0 0 0.3 0 0" size="0.02" mass="1"/></body>
    </worldbody><actuator><motor joint="j"/></actuator></mujoco>""", 2)
    s.set_param("body_mass", 1, 1, 2, [4.0])
    s.set("ctrl", 0, 2, [[1.0], [1.0]])
    s.step(0, 2, 200)
    q = s.get("qpos")
    assert abs(q[0, 0] / q[1, 0] - 4.0) < 1e-6


def test_invalid_params_are_rejected_without_partial_updates():
    s = sim_xml(INCLINE, 2)
    with pytest.raises(OracleError):
        s.set_param("geom_size", 0, 0, 2, [1.0] * 6)
    with pytest.raises(OracleError):
        s.set_param("geom_friction", 99, 0, 1, [1.0])
    with pytest.raises(OracleError):
        s.set_param("body_mass", 1, 0, 1, [-1.0])
    assert np.array_equal(s.get_param("geom_friction", 1)[:, 0], [0.1, 0.1])


# ---- closed forms ----
def test_double_pendulum_mass_matrix_and_gravity_bias():
    m1, m2, l1, c1, c2, i1, i2 = 1.3, 0.7, 0.5, 0.2, 0.3, 0.011, 0.006
    xml = f"""<mujoco><worldbody>
      <body name="a"><joint axis="0 1 0"/><inertial pos="{c1} 0 0" mass="{m1}" diaginertia="0.01 {i1} 0.01"/>
        <body name="b" pos="{l1} 0 0"><joint axis="0 1 0"/><inertial pos="{c2} 0 0" mass="{m2}" diaginertia="0.01 {i2} 0.01"/></body>
      </body></worldbody></mujoco>"""
    s = sim_xml(xml, 5)
    rng = np.random.default_rng(1)
    q = rng.uniform(-2, 2, size=(5, 2))
    s.set("qpos", 0, 5, q)
    s.set("qvel", 0, 5, np.zeros((5, 2)))
    qm, bias = s.mass_bias()
    g = 9.81
    for e in range(5):
        t1, t2 = q[e]
        m11 = i1 + i2 + m1 * c1 ** 2 + m2 * (l1 ** 2 + c2 ** 2 + 2 * l1 * c2 * math.cos(t2))
        m12 = i2 + m2 * (c2 ** 2 + l1 * c2 * math.cos(t2))
        m22 = i2 + m2 * c2 ** 2
        exp = np.array([[m11, m12], [m12, m22]])
        assert np.max(np.abs(qm[e] - exp)) < 1e-12 * m11
        # rotating about +y by t moves +x toward -z: height of a point at x = -x sin t
        # bias = -(generalised gravity force) = dU/dq, U = m g z
        b1 = -g * (m1 * c1 * math.cos(t1) + m2 * (l1 * math.cos(t1) + c2 * math.cos(t1 + t2)))
        b2 = -g * m2 * c2 * math.cos(t1 + t2)
        assert abs(bias[e, 0] - b1) < 1e-12 * 10 and abs(bias[e, 1] - b2) < 1e-12 * 10


def test_ballistic_free_body():
    s = sim_xml("""<mujoco><option timestep="0.01"/><worldbody>
        <body pos="0 0 5"><freejoint/><geom type="sphere" size=".1" contype="0" conaffinity="0"/></body>
    </worldbody></mujoco>""")
    v0 = np.array([[1.0, -0.5, 2.0, 0.0, 0.0, 0.0]])
    s.set("qvel", 0, 1, v0)
    n, h = 50, 0.01
    s.step(0, 1, n)
    q = s.get("qpos")[0]
    # semi-implicit Euler: v_k = v0 + k h g, p_n = p0 + h sum_{k=1..n} v_k
    z = 5 + h * (n * 2.0 - 9.81 * h * n * (n + 1) / 2)
    assert np.allclose(q[:3], [h * n * 1.0, -h * n * 0.5, z], atol=1e-12)
    assert np.allclose(q[3:7], [1, 0, 0, 0], atol=0)


# plane-cylinder (collision.rs plane_cylinder) has no reference test: pin it on
# rest heights, upright (half-height) and lying on its side (radius).
@both
def test_cylinder_rests_upright_and_on_its_side(solver):
    s = sim_xml("""<mujoco><option timestep="0.002"/><worldbody>
        <geom type="plane" size="5 5 .1"/>
        <body pos="0 0 0.3"><freejoint/><geom type="cylinder" size=".05 .08"/></body>
        <body pos="1 0 0.3" euler="90 0 0"><freejoint/><geom type="cylinder" size=".04 .1"/></body>
    </worldbody></mujoco>""", solver=solver)
    run(s, 3.0, 0.002)
    q = s.get("qpos")[0]
    assert abs(q[2] - 0.08) < 2e-3, q[2]
    assert abs(q[9] - 0.04) < 2e-3, q[9]
    assert np.linalg.norm(s.get("qvel")[0]) < 1e-2
