"""The `Physics` trait surface of the HIP backend against the oracle (needs a GPU).

Follows the reference's trait-level tests: crates/pippin-env/tests/params.rs
(per-env parameters, rejection without partial updates) and
crates/pippin/tests/contact.rs:165-188 (rollout equals the step loop), plus the
PippinCpu conventions (crates/pippin-env/src/pippin_cpu.rs): range stepping,
poses refreshed after step and after set(Qpos), reset to the reset state, size
checks.
"""
import sys

import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleError, OracleSim
from pippin_amd.physics import HipPhysics, ParamError

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(arr, n):
    return OracleSim(arr, n), HipPhysics(arr, n)


def _ant(assets, n, seed=0):
    arr = mjcf.load_file(assets / "ant.xml")
    o, g = _pair(arr, n)
    rng = np.random.default_rng(seed)
    q = o.get("qpos") + rng.normal(size=(n, 15)) * 0.01
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    for s in (o, g):
        s.set("qpos", 0, n, q)
    return arr, o, g, rng


def test_range_stepping_and_poses(assets):
    n = 40
    say("range stepping, ant", n)
    arr, o, g, rng = _ant(assets, n)
    for t in range(30):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in (o, g):
            s.set("ctrl", 0, n, c)
            # uneven ranges: the first 13 envs step twice, the rest once
            s.step(0, 13, 1)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    go, _ = o.poses()
    gg, _ = g.poses()
    assert np.array_equal(go["pos"], gg["pos"]) and np.array_equal(go["mat"], gg["mat"])
    # poses follow set(Qpos) without a step (pippin_cpu.rs:183-185)
    q = o.get("qpos")
    q[:, 2] += 0.25
    for s in (o, g):
        s.set("qpos", 5, 9, q[5:9])
    go, _ = o.poses(5, 9)
    gg, _ = g.poses(5, 9)
    assert np.array_equal(go["pos"], gg["pos"]) and np.array_equal(go["mat"], gg["mat"])


def test_reset_matches_oracle(assets):
    n = 16
    say("reset, ant", n)
    arr, o, g, rng = _ant(assets, n, 1)
    ctrl = rng.uniform(-1, 1, size=(n, 8))
    for s in (o, g):
        s.set("ctrl", 0, n, ctrl)
        s.step(0, n, 20)
        s.reset([3, 7, 8])
        s.step(0, n, 20)
    assert np.array_equal(o.get("qpos"), g.get("qpos"))
    assert np.array_equal(o.get("qvel"), g.get("qvel"))
    assert np.array_equal(o.ncon(), g.ncon())


def test_size_mismatch_raises(assets):
    arr = mjcf.load_file(assets / "ant.xml")
    g = HipPhysics(arr, 4)
    with pytest.raises(ValueError):
        g.set("qpos", 0, 4, np.zeros((3, 15)))
    with pytest.raises(Exception):
        g.step(0, 5, 1)


INCLINE = """<mujoco><option timestep="0.002"/><worldbody>
    <geom name="floor" type="plane" size="5 5 .1" euler="0 25 0" friction="0.1"/>
    <body name="b" pos="0 0 0.1" euler="0 25 0"><freejoint/><geom name="box" type="box" size=".05 .05 .05" friction="0.1"/></body>
</worldbody></mujoco>"""


def test_friction_per_env():
    say("friction per env")
    o, g = _pair(mjcf.load_string(INCLINE), 2)
    for s in (o, g):
        s.set_param("geom_friction", 1, 1, 2, [0.9])
        s.set_param("geom_friction", 0, 1, 2, [0.9])
        s.step(0, 2, 500)
    q = g.get("qpos")
    slid = np.hypot(q[:, 0], q[:, 1])
    assert slid[0] > 0.1 and slid[1] < 0.01
    assert np.array_equal(q, o.get("qpos"))
    assert np.array_equal(g.get_param("geom_friction", 1)[:, 0], [0.1, 0.9])


def test_size_and_parking_per_env():
    say("size and parking per env")
    xml = """<mujoco><worldbody><geom type="plane" size="5 5 .1"/>
        <body pos="0 0 0.3"><freejoint/><geom name="ball" type="sphere" size=".05"/></body>
    </worldbody></mujoco>"""
    o, g = _pair(mjcf.load_string(xml), 3)
    for s in (o, g):
        s.set_param("geom_size", 1, 1, 2, [0.1, 0.0, 0.0])
        s.set_param("geom_contype", 1, 2, 3, [0.0])
        s.set_param("geom_conaffinity", 1, 2, 3, [0.0])
        s.step(0, 3, 1000)
    q = g.get("qpos")
    assert abs(q[0, 2] - 0.05) < 2e-3 and abs(q[1, 2] - 0.10) < 2e-3 and q[2, 2] < -1.0
    assert np.array_equal(q, o.get("qpos"))
    ap = g.appearance()
    assert ap[1, 1]["scale"][0] == 2.0 and ap[0, 1]["scale"][0] == 1.0
    assert np.array_equal(ap, o.appearance())


def test_mass_per_env_changes_actuated_motion():
    say("mass per env")
    xml = """<mujoco><option gravity="0 0 0"/><worldbody>
        <body name="link"><joint name="j" axis="0 0 1"/><geom type="capsule" fromto="0 0 0 0.3 0 0" size="0.02" mass="1"/></body>
    </worldbody><actuator><motor joint="j"/></actuator></mujoco>"""
    o, g = _pair(mjcf.load_string(xml), 2)
    for s in (o, g):
        s.set_param("body_mass", 1, 1, 2, [4.0])
        s.set("ctrl", 0, 2, [[1.0], [1.0]])
        s.step(0, 2, 200)
    q = g.get("qpos")
    assert abs(q[0, 0] / q[1, 0] - 4.0) < 1e-6
    assert np.array_equal(q, o.get("qpos"))
    assert np.array_equal(g.invweight()[0], o.invweight()[0])


def test_invalid_params_are_rejected_without_partial_updates():
    g = HipPhysics(mjcf.load_string(INCLINE), 2)
    with pytest.raises(ParamError):
        g.set_param("geom_size", 0, 0, 2, [1.0] * 6)
    with pytest.raises(ParamError):
        g.set_param("geom_friction", 99, 0, 1, [1.0])
    with pytest.raises(ParamError):
        g.set_param("body_mass", 1, 0, 1, [-1.0])
    assert np.array_equal(g.get_param("geom_friction", 1)[:, 0], [0.1, 0.1])


@pytest.mark.parametrize("name", ["cartpole.xml", "ant.xml"])
def test_rollout_device_matches_oracle_rollout(assets, name):
    import torch

    say("rollout_device", name)
    arr = mjcf.load_file(assets / name)
    n, t, nu, nq = 6, 50, int(arr["nu"]), int(arr["nq"])
    ctrl = np.sin(np.arange(n * t * nu) * 0.37).reshape(n, t, nu)
    o, g = _pair(arr, n)
    out = o.rollout(ctrl)
    dctrl = torch.from_numpy(ctrl).to("cuda")
    dq = torch.zeros((n, t, nq), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.rollout_device(dctrl.data_ptr(), dq.data_ptr(), t)
    assert np.array_equal(dq.cpu().numpy(), out)
    assert np.array_equal(g.get("qpos"), o.get("qpos"))


@pytest.mark.parametrize("name", ["cartpole.xml", "pendulum.xml", "tumble.xml", "arm3d.xml"])
def test_contact_free_trajectories(assets, name):
    """Contact-free bar: qpos within 1e-12 over 200 steps (held to exact equality).

    tumble (free body + hinge) and arm3d (hinge, ball, slide; motor, position
    and velocity actuators; joint damping and armature) start from random
    velocities, the others from perturbed positions."""
    arr = mjcf.load_file(assets / name)
    n = 64
    o, g = _pair(arr, n)
    rng = np.random.default_rng(2)
    if name in ("tumble.xml", "arm3d.xml"):
        v = rng.normal(size=(n, int(arr["nv"])))
        for s in (o, g):
            s.set("qvel", 0, n, v)
    else:
        q = o.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.2
        for s in (o, g):
            s.set("qpos", 0, n, q)
    nu = int(arr["nu"])
    for t in range(200):
        c = rng.uniform(-1, 1, size=(n, nu))
        for s in (o, g):
            if nu:
                s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
    err = np.abs(o.get("qpos") - g.get("qpos")).max()
    say(name, "200 steps max qpos err", err)
    assert err <= 1e-12
    assert err == 0.0


def test_masks_switch_general_pairs_on_per_env(assets):
    """arm3d's cylinder meets its boxes and capsules only through the general
    convex narrow phase, and its collision masks keep every pair off. Masks
    that switch them on for one env (a copy-on-write parameter variant) are
    accepted and stored for that env alone, and every env still equals the
    oracle given the same masks."""
    arr = mjcf.load_file(assets / "arm3d.xml")
    n = 3
    g, o = HipPhysics(arr, n), OracleSim(arr, n)
    for s in (g, o):
        for gi in range(int(arr["ngeom"])):
            s.set_param("geom_contype", gi, 1, 2, [1.0])
            s.set_param("geom_conaffinity", gi, 1, 2, [1.0])
    assert np.array_equal(g.get_param("geom_contype", 0)[:, 0], [0, 1, 0])
    assert np.array_equal(g.get_param("geom_conaffinity", 2)[:, 0], [0, 1, 0])
    rng = np.random.default_rng(23)
    v = rng.normal(size=(n, int(arr["nv"])))
    for s in (g, o):
        s.set("qvel", 0, n, v)
    for t in range(150):
        for s in (g, o):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
    assert np.array_equal(o.get("qpos"), g.get("qpos"))


def test_contact_capacity_overflow_fails_loudly_and_freezes_the_env(assets):
    """A user-lowered contact capacity (options.max_contacts) that a step
    exceeds: the step reports the failure (the reference would panic), the
    failed env keeps its last state, the others go on stepping."""
    from pippin_amd.physics import HipError

    arr = mjcf.load_file(assets / "box_drop.xml")
    n = 3
    g = HipPhysics(arr, n, max_contacts=2)  # a box resting on the plane makes 4
    ok = HipPhysics(arr, n)
    q = g.get("qpos")
    q[1, [2, 9, 16]] += 5.0  # env 1's three bodies start high up: no contact in time
    for s in (g, ok):
        s.set("qpos", 0, n, q)
    failed_at = None
    for t in range(200):
        try:
            g.step(0, n, 1)
        except HipError as e:
            assert "capacity" in str(e)
            failed_at = t
            break
        ok.step(0, n, 1)
    assert failed_at is not None
    frozen = g.get("qpos")
    g.step(1, 2, 5)  # the env that has not failed still steps
    after = g.get("qpos")
    assert not np.array_equal(after[1], frozen[1])
    with pytest.raises(HipError):
        g.step(0, 1, 1)  # the failed env reports again and stays put
    assert np.array_equal(g.get("qpos")[0], frozen[0])
    # Data::reset brings the failed env back: it steps again from the reset state
    ok.reset([0])
    g.reset([0])
    q_reset = g.get("qpos")[0].copy()
    assert np.array_equal(q_reset, ok.get("qpos")[0])
    g.step(0, 1, 2)  # no error: the env is no longer frozen
    assert np.all(np.isfinite(g.get("qpos")[0]))
    assert not np.array_equal(g.get("qpos")[0], q_reset)  # and it moves (the drop starts)


def test_set_device_is_ordered_before_get_set_and_step(assets):
    """set_device queues a device-to-device copy on the library stream;
    get and set wait for it, and a host set after it wins."""
    import torch

    arr = mjcf.load_file(assets / "ant.xml")
    n = 64
    g, ref = HipPhysics(arr, n), OracleSim(arr, n)
    rng = np.random.default_rng(5)
    c_dev = rng.uniform(-1, 1, size=(n, 8))
    buf = torch.from_numpy(c_dev).to("cuda")
    g.set_device("ctrl", 0, n, buf.data_ptr())
    assert np.array_equal(g.get("ctrl"), c_dev)
    g.set_device("ctrl", 0, n, buf.data_ptr())
    c_host = rng.uniform(-1, 1, size=(n, 8))
    g.set("ctrl", 0, n, c_host)  # issued after the device copy: it must win
    assert np.array_equal(g.get("ctrl"), c_host)
    ref.set("ctrl", 0, n, c_host)
    for s in (g, ref):
        s.step(0, n, 3)
    assert np.array_equal(g.get("qpos"), ref.get("qpos"))


def test_device_poses_equal_host_poses(assets):
    """pippin_hip_poses_device: the f32 Pose records of every env in device
    memory, for a renderer on the GPU; equal to the host copy of poses() bit
    for bit, after steps and after set(qpos)."""
    from pippin_amd.desc import POSE_DTYPE
    import ctypes

    from pippin_amd import physics

    # the Ant with a camera on its torso, so that camera poses move too
    xml = (assets / "ant.xml").read_text()
    i = xml.index(">", xml.index('<body name="torso"')) + 1
    xml = xml[:i] + '<camera name="follow" pos="0 -1.5 0.6" quat="0.92 0.38 0 0"/>' + xml[i:]
    arr = mjcf.load_string(xml)
    assert int(arr["ncam"]) == 1
    n = 40
    g = HipPhysics(arr, n)
    rng = np.random.default_rng(2)
    hip = physics.lib()  # its HIP runtime (the one torch loaded) copies device to host
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int

    def device_copy(ptr, count):
        out = np.zeros(count, dtype=POSE_DTYPE)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    first = None
    for t in range(3):
        g.set("ctrl", 0, n, rng.uniform(-1, 1, size=(n, 8)))
        g.step(0, n, 5)
        if t == 2:
            q = g.get("qpos")
            q[:, 2] += 0.1
            g.set("qpos", 0, n, q)
        gp, cp = g.poses_device()
        first = first or (gp, cp)
        assert (gp, cp) == first  # the buffers stay put
        hg, hc = g.poses()
        dg = device_copy(gp, n * g.ngeom).reshape(n, g.ngeom)
        assert np.array_equal(dg["pos"], hg["pos"]) and np.array_equal(dg["mat"], hg["mat"])
        if g.ncam:
            dc = device_copy(cp, n * g.ncam).reshape(n, g.ncam)
            assert np.array_equal(dc["pos"], hc["pos"]) and np.array_equal(dc["mat"], hc["mat"])
