"""The kernel body, compiled for the host and run serially (tests/emu), against
the oracle on identical inputs (CPU only).

This checks the device code's arithmetic and control flow without a GPU: the
same kbody.hpp / dcollide.hpp / dmath.hpp source that the HIP kernels run, with
the lane loops executed one after another. Bars as on the GPU: contact counts,
pairs and iteration counts equal, M(q) and c(q, qdot) exact, qpos exact over
200 steps through contact.
"""
import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim
from tests.emu.pyemu import EmuSim
from tests.scenes import (GENERAL_SCENES, MESH_SCENES, MULTIBODY_SCENES, PRIMITIVE_SCENES, general_xml, mesh_xml,
                          primitive_xml)


def _perturbed(q, rng, scale, free=True):
    q = q + rng.normal(size=q.shape) * scale
    if free:
        q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    return q


@pytest.fixture(scope="module")
def ant(assets):
    return mjcf.load_file(assets / "ant.xml")


def test_ant_mass_bias_invweight_exact(ant):
    n = 16
    o, e = OracleSim(ant, n), EmuSim(ant, n)
    rng = np.random.default_rng(0)
    q = _perturbed(o.get("qpos"), rng, 0.3)
    v = rng.normal(size=(n, 14))
    for s in (o, e):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    m1, b1 = o.mass_bias()
    m2, b2 = e.mass_bias()
    assert np.array_equal(m1, m2) and np.array_equal(b1, b2)
    d1, w1 = o.invweight()
    d2, w2 = e.invweight()
    assert np.array_equal(d1[0], d2) and np.array_equal(w1[0], w2)


def test_ant_trajectories_exact(ant):
    n, T = 24, 200
    o, e = OracleSim(ant, n), EmuSim(ant, n)
    rng = np.random.default_rng(0)
    q = _perturbed(o.get("qpos"), rng, 0.01)
    for s in (o, e):
        s.set("qpos", 0, n, q)
    ctrl = rng.uniform(-1, 1, size=(T, n, 8))
    contacts = 0
    for t in range(T):
        for s in (o, e):
            s.set("ctrl", 0, n, ctrl[t])
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        contacts += int(o.ncon().sum())
        if t % 40 == 0:
            for k in range(0, n, 7):
                co, ce = o.contacts(k), e.contacts(k)
                assert np.array_equal(co[["geom1", "geom2", "feature"]], ce[["geom1", "geom2", "feature"]])
                assert np.array_equal(co["depth"], ce["depth"])
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.get("qvel"), e.get("qvel"))
    assert np.array_equal(o.solver_iterations(), e.solver_iterations())
    assert contacts > 0 and not e.status().any()


def test_ant_pressed_into_the_floor_exact(ant):
    """Ant states pushed into the floor with bent legs, so that steps hold 17
    to 32 constraint rows and over 32 rows (more than the 16-row lane solver
    and LDS row block hold: the dense solver, rows in HBM): trajectories,
    contact counts and iteration counts equal the oracle's."""
    n, T = 64, 20
    o, e = OracleSim(ant, n), EmuSim(ant, n)
    rng = np.random.default_rng(1)
    q = o.get("qpos")
    q[:, 2] = rng.uniform(0.05, 0.35, n)
    q[:, 7:] += rng.uniform(-0.8, 0.8, (n, 8))
    for s in (o, e):
        s.set("qpos", 0, n, q)
    mid = wide = 0
    for t in range(T):
        c = rng.uniform(-1, 1, (n, 8))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        it, nr, _ = o.step_stats()
        mid += int(((nr > 16) & (nr <= 32)).sum())
        wide += int((nr > 32).sum())
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(it, e.solver_iterations()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
        assert np.array_equal(o.get("qvel"), e.get("qvel")), t
    assert mid > 20 and wide > 0, (mid, wide)
    assert not e.status().any()


@pytest.mark.parametrize("name", ["cartpole.xml", "pendulum.xml", "tumble.xml", "arm3d.xml"])
def test_contact_free_models_exact(assets, name):
    """Contact-free models: hinge/slide chains, a free body with a hinged
    appendage, and arm3d (ball and slide joints, motor/position/velocity
    actuators, damping, armature, general pairs masked off)."""
    arr = mjcf.load_file(assets / name)
    n = 8
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(5)
    if name in ("tumble.xml", "arm3d.xml"):
        v = rng.normal(size=(n, int(arr["nv"])))
        for s in (o, e):
            s.set("qvel", 0, n, v)
    else:
        q = _perturbed(o.get("qpos"), rng, 0.2, False)
        for s in (o, e):
            s.set("qpos", 0, n, q)
    nu = int(arr["nu"])
    for t in range(200):
        c = rng.uniform(-1, 1, size=(n, nu))
        for s in (o, e):
            if nu:
                s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
    assert np.array_equal(o.get("qpos"), e.get("qpos"))
    assert np.array_equal(o.get("qvel"), e.get("qvel"))


@pytest.mark.parametrize("name", sorted(PRIMITIVE_SCENES))
def test_primitive_pairs_exact(name):
    arr = mjcf.load_string(primitive_xml(name))
    n = 6
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(11)
    q = o.get("qpos")
    q[:, :3] += rng.normal(size=(n, 3)) * 0.02
    v = rng.normal(size=(n, 6)) * 0.3
    for s in (o, e):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    touched = 0
    for t in range(250):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        touched += int((o.ncon() > 0).sum())
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert touched > 0, "the scene never made contact"
    for k in range(n):
        co, ce = o.contacts(k), e.contacts(k)
        for f in ("geom1", "geom2", "feature", "depth", "pos", "normal", "tangent"):
            assert np.array_equal(co[f], ce[f]), f


@pytest.mark.parametrize("name", sorted(MULTIBODY_SCENES))
def test_multibody_islands_exact(assets, name):
    """Several dof blocks: Newton islands formed by contacts (newton.rs:378-473)."""
    xml = MULTIBODY_SCENES[name]
    arr = mjcf.load_file(assets / "box_drop.xml") if xml is None else mjcf.load_string(xml)
    n = 6
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(3)
    nb = (int(arr["nq"]) // 7)
    q = o.get("qpos")
    for b in range(nb):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, e):
        s.set("qpos", 0, n, q)
    for t in range(300):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), e.solver_iterations())
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("name", ["plane_box", "box_box", "capsule_capsule", "plane_cylinder", "sphere_box"])
def test_pgs_primitive_exact(name):
    """PGS (solver.rs:44-289) on the kernel body against the oracle."""
    arr = mjcf.load_string(primitive_xml(name))
    arr["solver_kind"] = 1
    n = 6
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(13)
    q = o.get("qpos")
    q[:, :3] += rng.normal(size=(n, 3)) * 0.02
    v = rng.normal(size=(n, 6)) * 0.3
    for s in (o, e):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    for t in range(250):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.get("qvel"), e.get("qvel"))


@pytest.mark.parametrize("name", sorted(MULTIBODY_SCENES))
def test_pgs_multibody_exact(assets, name):
    xml = MULTIBODY_SCENES[name]
    arr = mjcf.load_file(assets / "box_drop.xml") if xml is None else mjcf.load_string(xml)
    arr["solver_kind"] = 1
    n = 4
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    for t in range(300):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t


def test_pgs_ant_exact(ant):
    arr = dict(ant)
    arr["solver_kind"] = 1
    n = 8
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(1)
    for t in range(150):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t


def _ant_pair(ant, n, seed):
    o, e = OracleSim(ant, n), EmuSim(ant, n)
    rng = np.random.default_rng(seed)
    q = _perturbed(o.get("qpos"), rng, 0.01)
    for s in (o, e):
        s.set("qpos", 0, n, q)
    return o, e, rng


def test_learned_warm_start_logging_is_the_reference_path(ant):
    """The learned-warm-start kernel body (run_env_mlp, mlp_stage, mlp_targets) on
    the host with logging on: trajectories equal the oracle's bit for bit; the
    records hold the touching contacts, predictions equal the FP64 restatement of
    the network (oracle/mlp.py) on the recorded features, targets are
    log1p(lambda / D) >= 0."""
    from oracle import mlp as omlp

    n, h = 8, 32
    o, e, rng = _ant_pair(ant, n, 21)
    P = omlp.random_params(h, np.random.default_rng(3), scale=1.5)
    e.mlp_set(h, 1, P)
    seen = 0
    for t in range(80):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
        for k in range(n):
            f, pr, tg = e.mlp_records(k)
            assert f.shape[0] == min(int((o.contacts(k)["depth"] > 0).sum()), 16)
            if f.shape[0]:
                seen += f.shape[0]
                assert np.isfinite(f).all() and (tg >= 0).all() and np.isfinite(tg).all()
                ref, _ = omlp.forward(P, h, f, bf16=True)
                assert np.abs(pr - ref).max() <= 1e-5 * (1 + np.abs(ref).max()), (t, k)
    assert seen > 0
    assert np.array_equal(o.solver_iterations(), e.solver_iterations())


def test_learned_warm_start_candidate(ant):
    """With the candidate on: an all-zero network (lambda = 0, the candidate is
    a0) changes nothing, bit for bit; a random network gives the same step as
    the reference rule to the solver tolerance, from the same state."""
    from oracle import mlp as omlp

    n, h = 8, 32
    o, e, rng = _ant_pair(ant, n, 22)
    e.mlp_set(h, 3, np.zeros(omlp.nparam(h), dtype=np.float32))
    for t in range(60):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    e.mlp_set(h, 3, omlp.random_params(h, np.random.default_rng(4), scale=1.0))
    for t in range(40):
        c = rng.uniform(-1, 1, size=(n, 8))
        e.set("qpos", 0, n, o.get("qpos"))
        e.set("qvel", 0, n, o.get("qvel"))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.abs(o.get("qvel") - e.get("qvel")).max() < 1e-6, t


@pytest.mark.parametrize("name", ["arm3d.xml", "ant.xml"])
def test_implicitfast_exact(assets, name):
    """integrator="implicitfast" (forward.rs:407-431): the position and velocity
    actuators' kv terms on tree-ancestor dof pairs (arm3d) and joint damping
    (both), through the general factorisation path."""
    xml = (assets / name).read_text()
    xml = xml.replace("<option ", '<option integrator="implicitfast" ', 1)
    arr = mjcf.load_string(xml)
    assert int(arr["integrator"]) == 1
    n = 6
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(17)
    v = rng.normal(size=(n, int(arr["nv"]))) * 0.5
    for s in (o, e):
        s.set("qvel", 0, n, v)
    nu = int(arr["nu"])
    for t in range(150):
        c = rng.uniform(-1, 1, size=(n, nu))
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.get("qvel"), e.get("qvel"))


def test_gripper_exact():
    """Arm + gripper + cube (tests/scenes.py GRIPPER): fixed-tendon position
    actuator, joint equality between the fingers, implicitfast, joint limits,
    and the arm and cube dof blocks joined into one island by contact."""
    from tests.scenes import GRIPPER

    arr = mjcf.load_string(GRIPPER)
    n = 4
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(31)
    q = o.get("qpos")
    q[:, 5:7] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, e):
        s.set("qpos", 0, n, q)
    for t in range(300):
        c = np.stack([np.sin(t * 0.01 + np.arange(n)), 0.4 + 0.3 * np.sin(t * 0.013) + 0 * np.arange(n),
                      -0.9 + 0.5 * np.sin(t * 0.017) + 0 * np.arange(n), 0.03 * np.sin(t * 0.02 + np.arange(n))], 1)
        for s in (o, e):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), e.solver_iterations())
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("seed", [3, 11, 29])
@pytest.mark.parametrize("solver", [0, 1])
def test_random_piles_exact(seed, solver):
    """Random primitive piles (tests/scenes.py random_pile_xml), Newton and PGS."""
    from tests.scenes import random_pile_xml

    arr = mjcf.load_string(random_pile_xml(seed))
    arr["solver_kind"] = solver
    n = 3
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    for t in range(200):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("name", sorted(GENERAL_SCENES))
@pytest.mark.parametrize("solver", [0, 1])
def test_general_pairs_exact(name, solver):
    """The general convex narrow phase (dgeneral.hpp) against its CPU
    restatement (oracle/general.cpp), Newton and PGS, bit for bit."""
    arr = mjcf.load_string(general_xml(name))
    arr["solver_kind"] = solver
    n = 4
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(17)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, e):
        s.set("qpos", 0, n, q)
    for t in range(250):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), e.solver_iterations())
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("seed", [2, 13])
@pytest.mark.parametrize("solver", [0, 1])
def test_random_piles_with_general_pairs_exact(seed, solver):
    """Random piles whose cylinders meet every body (general pairs)."""
    from pippin_amd.workloads import random_pile_xml

    arr = mjcf.load_string(random_pile_xml(seed, general=True))
    arr["solver_kind"] = solver
    n = 3
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    for t in range(200):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("solver", [0, 1])
def test_arm_gripper_exact(solver):
    """The 7-DoF arm with tendon-coupled gripper (pippin_amd/workloads.py
    arm_gripper_xml): implicitfast, position actuators, fixed tendon, joint
    equality, limits and hand/cube/table contacts, on smooth random targets;
    Newton and PGS (whose 80-odd rows use the global row buffers)."""
    from pippin_amd.workloads import arm_gripper_xml, arm_targets

    arr = mjcf.load_string(arm_gripper_xml(3))
    arr["solver_kind"] = solver
    n = 4
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    tg = arm_targets(3, n, 150, int(arr["nu"]))
    for t in range(150):
        for s in (o, e):
            s.set("ctrl", 0, n, tg[t])
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("name", sorted(MESH_SCENES))
@pytest.mark.parametrize("solver", [0, 1])
def test_mesh_scenes_exact(name, solver):
    """Convex hulls (plane against a hull, hulls against every primitive and
    each other) in the kernel body against the oracle, bit for bit."""
    arr = mjcf.load_string(mesh_xml(name))
    arr["solver_kind"] = solver
    n = 3
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(23)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (o, e):
        s.set("qpos", 0, n, q)
    for t in range(150):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    assert o.ncon().sum() > 0


def _bake(arr, plan, env):
    """The tumble model with env `env`'s per-env parameters written into its
    arrays (the emulator keeps one parameter set per model)."""
    out = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
    for name, gid, values in plan.params:
        v = values[env]
        if name in ("geom_contype", "geom_conaffinity"):
            out[name][gid] = int(v[0])
        elif name == "geom_size":
            out[name][gid] = v
        elif name in ("geom_friction", "body_mass"):
            out[name][gid] = v[0]
        elif name == "body_inertia":
            out[name][gid] = np.diag(v).reshape(out[name][gid].shape)
        elif name == "body_ipos":
            out[name][gid] = v
    out["qpos_reset"] = plan.qpos0[env].copy()
    return out


@pytest.mark.parametrize("solver", [0, 1])
def test_tabletop_tumble_exact(solver):
    """The tumble workload (2 to 6 objects per env chosen by contype masks, hull
    wedges among the primitives) in the kernel body against the oracle, one
    env per model with its parameters baked in."""
    from pippin_amd.workloads import tabletop_tumble

    n = 4
    plan = tabletop_tumble(n, seed=solver + 5)
    base = mjcf.load_string(plan.xml)
    base["solver_kind"] = solver
    total = 0
    for env in range(n):
        arr = _bake(base, plan, env)
        o, e = OracleSim(arr, 1), EmuSim(arr, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos"))
        for t in range(120):
            for s in (o, e):
                s.step(0, 1, 1)
            assert np.array_equal(o.ncon(), e.ncon()), (env, t)
            assert np.array_equal(o.get("qpos"), e.get("qpos")), (env, t)
            total += int(o.ncon().sum())
    assert total > 0


@pytest.mark.parametrize("solver", [0, 1])
def test_urdf_arm_exact(tmp_path, solver):
    """The URDF arm (oracle URDF import) through the kernel body on the host:
    positions and contact counts equal the oracle's over 150 steps of a fall
    onto the floor."""
    from oracle import urdf
    from tests.test_urdf import ARM

    path = tmp_path / "arm.urdf"
    path.write_text(ARM)
    arr = urdf.load_file(path)
    arr["solver_kind"] = solver
    n = 4
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(11)
    q = o.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.3
    for s in (o, e):
        s.set("qpos", 0, n, q)
    seen = 0
    for t in range(150):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), e.ncon()), t
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
        seen += int(o.ncon().sum())
    assert seen > 0


def _pgs_pile_pair(n, seed=29):
    from pippin_amd.workloads import random_pile_xml

    arr = mjcf.load_string(random_pile_xml(seed))
    arr["solver_kind"] = 1
    o, e = OracleSim(arr, n), EmuSim(arr, n)
    rng = np.random.default_rng(seed)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (o, e):
        s.set("qpos", 0, n, q)
    return arr, o, e


def _slog(x):
    return np.sign(x) * np.log1p(np.abs(x))


def test_pgs_learned_warm_start_logging_is_the_reference_path():
    """The PGS hook's kernel body (run_env_mlp's PGS branch, mlp_stage_pgs,
    mlp_targets_pgs) on the host with logging on: trajectories equal the
    oracle's bit for bit; the records hold the env's first 16 contacts, their
    targets are sign(x) log1p(|x|) of the solved normal and friction forces
    scaled by 100 / inv_a (normal >= 0, friction within mu times normal),
    and the predictions equal
    the FP64 restatement of the network on the recorded features."""
    from oracle import mlp as omlp

    n, h = 4, 32
    arr, o, e = _pgs_pile_pair(n)
    P = omlp.random_params(h, np.random.default_rng(5), scale=1.5)
    e.mlp_set(h, 1, P)
    seen = 0
    for t in range(120):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
        assert np.array_equal(o.ncon(), e.ncon()), t
        for k in range(n):
            f, pr, tg = e.mlp_records(k)
            assert f.shape[0] == min(int(o.ncon()[k]), 16), (t, k)
            if f.shape[0]:
                seen += f.shape[0]
                assert np.isfinite(f).all() and np.isfinite(tg).all()
                assert (tg[:, 0] >= 0).all() and (tg[:, 3] == 0).all()
                # targets are sign(x) log1p(|x|) of 100 lambda_r / inv_a_r; the
                # features carry 0.1 log(inv_a_r) in float32
                mu = f[:, 11].astype(np.float64)
                ia = np.exp(10.0 * f[:, 3:6].astype(np.float64))
                lam = np.sign(tg[:, :3]) * np.expm1(np.abs(tg[:, :3].astype(np.float64))) * 0.01 * ia
                for j in (1, 2):
                    assert (np.abs(lam[:, j]) <= mu * lam[:, 0] * (1 + 1e-3) + 1e-9).all(), (t, k)
                # the FP64 restatement rounds at the same BF16 points; a value on a
                # rounding boundary may still land one BF16 step apart (2^-8)
                ref, _ = omlp.forward(P, h, f, bf16=True)
                assert np.abs(pr - ref).max() <= 1e-3 * (1 + np.abs(ref).max()), (t, k)
    assert seen > 0
    assert (e.solver_iterations() > 0).any()  # sweeps of the last step


def test_pgs_learned_warm_start_zero_network_and_random_network():
    """With the hook on: an all-zero network predicts zero forces, so the
    unmatched contacts start at zero as in the reference and trajectories equal
    the oracle's bit for bit; with a random network each step from the oracle's
    positions and velocities stays finite and close to the oracle's step (the
    warm start moves where the sweeps stop, not what they solve)."""
    from oracle import mlp as omlp

    n, h = 4, 32
    arr, o, e = _pgs_pile_pair(n)
    e.mlp_set(h, 3, np.zeros(omlp.nparam(h), dtype=np.float32))
    for t in range(100):
        for s in (o, e):
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), e.get("qpos")), t
    e.mlp_set(h, 3, omlp.random_params(h, np.random.default_rng(6), scale=0.5))
    moved = 0.0
    for t in range(30):
        e.set("qpos", 0, n, o.get("qpos"))
        e.set("qvel", 0, n, o.get("qvel"))
        for s in (o, e):
            s.step(0, n, 1)
        d = np.abs(o.get("qvel") - e.get("qvel")).max()
        moved = max(moved, d)
        assert np.isfinite(e.get("qvel")).all() and d < 0.05, (t, d)
