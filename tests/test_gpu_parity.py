"""HIP kernels vs the oracle on identical inputs (needs a GPU).

Bars (from the target): contact counts and pair indices bit-exact per step,
M(q) and c(q, qdot) within 1e-12 relative, qpos within 1e-12 contact-free and
1e-9 through contact. The kernels keep the reference's operation order and the
oracle and device share correctly rounded sin/cos, so the tests check for
exact equality first and report the tolerance-based bar alongside.
"""
import math
import sys

import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim, cr_sincos
from pippin_amd.physics import HipPhysics, load_model, selftest_math
from tests.scenes import (GENERAL_SCENES, MESH_SCENES, MULTIBODY_SCENES, PRIMITIVE_SCENES, general_xml, mesh_xml,
                          primitive_xml)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_math_matches_host():
    rng = np.random.default_rng(7)
    # square roots and reciprocals over the whole range, including both sides
    # of the shortcut bounds (2^-767 for sqrt, [2^-537, 2^512] for 1/x)
    wide = np.exp2(rng.uniform(-1070, 1023, 200000))
    edges = np.array([2.0 ** -767, 2.0 ** -537, 2.0 ** 512, 2.0 ** 1023, 5e-324, 2.2250738585072014e-308,
                      1.7976931348623157e308])
    with np.errstate(over="ignore"):
        edges = np.concatenate([edges, np.nextafter(edges, 0), np.nextafter(edges, np.inf)])
    x = np.concatenate([rng.uniform(-20, 20, 20000), rng.uniform(-1e-3, 1e-3, 2000),
                        [0.0, -0.0, 1e-300, math.pi / 4, math.pi / 2, 1.0, -2.5, 3.0],
                        wide, -wide[:1000], edges, np.array([np.inf, -np.inf, np.nan])])
    fin = np.isfinite(x)
    say("device math on", x.size, "inputs")
    out = selftest_math(x)
    small = fin & (np.abs(x) < 1e6)  # cr_sincos is checked on the range the kernels use
    s, c = cr_sincos(x[small])
    assert np.array_equal(out[small, 0], s) and np.array_equal(out[small, 1], c), "cr_sincos differs"
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        assert np.array_equal(out[:, 2], np.sqrt(np.abs(x)), equal_nan=True), "device sqrt is not correctly rounded"
        assert np.array_equal(out[:, 3], 1.0 / x, equal_nan=True), "device reciprocal is not correctly rounded"


def _ant(assets, n):
    arr = mjcf.load_file(assets / "ant.xml")
    return arr, OracleSim(arr, n), HipPhysics(arr, n)


def _perturbed_qpos(sim, rng, n, scale=0.01):
    q = sim.get("qpos")
    q = q + rng.normal(size=q.shape) * scale
    for i in range(n):
        q[i, 3:7] /= np.linalg.norm(q[i, 3:7])
    return q


def test_ant_mass_and_bias_match(assets):
    n = 64
    arr, o, g = _ant(assets, n)
    rng = np.random.default_rng(0)
    q = _perturbed_qpos(o, rng, n, 0.3)
    v = rng.normal(size=(n, 14))
    for s in (o, g):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    say("ant mass/bias", n, "envs")
    m1, b1 = o.mass_bias()
    m2, b2 = g.mass_bias()
    rel_m = np.abs(m1 - m2).max() / np.abs(m1).max()
    rel_b = np.abs(b1 - b2).max() / np.abs(b1).max()
    say("M rel err", rel_m, "c rel err", rel_b)
    assert rel_m <= 1e-12 and rel_b <= 1e-12
    assert np.array_equal(m1, m2) and np.array_equal(b1, b2)
    d1, w1 = o.invweight()
    d2, w2 = g.invweight()
    assert np.array_equal(d1, d2) and np.array_equal(w1, w2)


def test_ant_trajectories_match(assets):
    """Config 0 shape: qpos0 + N(0, 0.01), ctrl U[-1, 1] per step, Newton, 200 steps."""
    n = 256
    arr, o, g = _ant(assets, n)
    rng = np.random.default_rng(0)
    q = _perturbed_qpos(o, rng, n)
    for s in (o, g):
        s.set("qpos", 0, n, q)
    ctrl = rng.uniform(-1, 1, size=(200, n, 8))
    worst = 0.0
    for t in range(200):
        for s in (o, g):
            s.set("ctrl", 0, n, ctrl[t])
            s.step(0, n, 1)
        nc_o, nc_g = o.ncon(), g.ncon()
        assert np.array_equal(nc_o, nc_g), f"ncon differs at step {t}"
        if t % 50 == 0:
            for e in range(0, n, 37):
                co, cg = o.contacts(e), g.contacts(e)
                assert np.array_equal(co[["geom1", "geom2", "feature"]], cg[["geom1", "geom2", "feature"]])
        err = np.abs(o.get("qpos") - g.get("qpos")).max()
        worst = max(worst, err)
        assert err <= 1e-9, f"qpos error {err} at step {t}"
    say("ant 200 steps: worst qpos err", worst, "contacts", int(o.ncon().sum()))
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())
    assert worst == 0.0


@pytest.mark.parametrize("name", sorted(PRIMITIVE_SCENES))
def test_primitive_pairs_match(name):
    """Every primitive pair of the device narrow phase, 250 steps through contact."""
    arr = mjcf.load_string(primitive_xml(name))
    n = 64
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(11)
    q = o.get("qpos")
    q[:, :3] += rng.normal(size=(n, 3)) * 0.02
    v = rng.normal(size=(n, 6)) * 0.3
    for s in (o, g):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    say("primitive pair", name)
    for t in range(250):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        err = np.abs(o.get("qpos") - g.get("qpos")).max()
        assert err <= 1e-9, (t, err)
    assert np.array_equal(o.get("qpos"), g.get("qpos"))
    for k in range(0, n, 9):
        co, cg = o.contacts(k), g.contacts(k)
        for f in ("geom1", "geom2", "feature", "depth", "pos", "normal", "tangent"):
            assert np.array_equal(co[f], cg[f]), f


@pytest.mark.parametrize("shape", [(4, 1, "ws_bucket"), (1, 0, "grid")])
def test_ant_pressed_into_the_floor_matches(assets, shape):
    """Ant states pushed into the floor with bent legs: steps with 17 to 32
    and over 32 constraint rows, more than the 16-row lane solver and LDS row
    block hold, so they take the dense solver with their rows in HBM. The
    Ant-specialised kernel at 4 envs per wavefront and the generic one at 1
    match the oracle bit for bit (contacts, iterations, qpos, qvel)."""
    n, T = 256, 20
    arr, o, g = _ant(assets, n)
    g.set_launch_shape(*shape)
    rng = np.random.default_rng(1)
    q = o.get("qpos")
    q[:, 2] = rng.uniform(0.05, 0.35, n)
    q[:, 7:] += rng.uniform(-0.8, 0.8, (n, 8))
    for s in (o, g):
        s.set("qpos", 0, n, q)
    mid = wide = 0
    say("ant pressed into the floor", shape)
    for t in range(T):
        c = rng.uniform(-1, 1, (n, 8))
        for s in (o, g):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        it, nr, _ = o.step_stats()
        mid += int(((nr > 16) & (nr <= 32)).sum())
        wide += int((nr > 32).sum())
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(it, g.solver_iterations()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
        assert np.array_equal(o.get("qvel"), g.get("qvel")), t
    say("  steps with 17..32 rows", mid, "over 32", wide)
    assert mid > 50 and wide > 0, (mid, wide)


def test_specialised_kernel_equals_generic(assets, monkeypatch):
    """Ant runs a kernel built with compile-time sizes and layout; it must give
    the generic kernel's results bit for bit (and both the oracle's)."""
    arr = mjcf.load_file(assets / "ant.xml")
    n = 96
    rng = np.random.default_rng(5)
    sims = []
    for spec in ("1", "0"):
        monkeypatch.setenv("PIPPIN_HIP_SPEC", spec)
        sims.append(HipPhysics(arr, n))
    q = sims[0].get("qpos") + rng.normal(size=(n, 15)) * 0.05
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    for s in sims:
        s.set("qpos", 0, n, q)
    say("specialised vs generic kernel, ant", n)
    for t in range(60):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in sims:
            s.set("ctrl", 0, n, c)
            s.step(0, n, 2)
    assert np.array_equal(sims[0].get("qpos"), sims[1].get("qpos"))
    assert np.array_equal(sims[0].get("qvel"), sims[1].get("qvel"))
    assert np.array_equal(sims[0].solver_iterations(), sims[1].solver_iterations())


@pytest.mark.parametrize("k", ["1", "2"])
def test_env_packing_matches(assets, monkeypatch, k):
    """The generic kernel with 1 or 2 envs per wavefront (the packings larger
    models get) against the oracle, at a ragged tail (n = 37), with the
    learned-warm-start kernel's logging variant alongside."""
    arr = mjcf.load_file(assets / "ant.xml")
    n = 37
    monkeypatch.setenv("PIPPIN_HIP_ENVS_PER_WG", k)
    monkeypatch.setenv("PIPPIN_HIP_SPEC", "0")
    g, lg = HipPhysics(arr, n), HipPhysics(arr, n)
    lg.mlp_init(hidden=32, seed=1, max_batch=1024)
    lg.mlp_mode(HipPhysics.MLP_LOG)
    o = OracleSim(arr, n)
    rng = np.random.default_rng(8)
    q = o.get("qpos") + rng.normal(size=(n, 15)) * 0.02
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    for s in (o, g, lg):
        s.set("qpos", 0, n, q)
    say("env packing", k, "ant", n)
    for t in range(80):
        c = rng.uniform(-1, 1, size=(n, 8))
        for s in (o, g, lg):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
        assert np.array_equal(o.get("qpos"), lg.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())


@pytest.mark.parametrize("name", sorted(MULTIBODY_SCENES))
def test_multibody_islands_match(assets, name):
    """Several dof blocks: Newton islands formed by contacts (newton.rs:378-473)."""
    xml = MULTIBODY_SCENES[name]
    arr = mjcf.load_file(assets / "box_drop.xml") if xml is None else mjcf.load_string(xml)
    n = 64
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(3)
    nb = int(arr["nq"]) // 7
    q = o.get("qpos")
    for b in range(nb):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("multibody islands", name)
    for t in range(300):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        err = np.abs(o.get("qpos") - g.get("qpos")).max()
        assert err <= 1e-9, (t, err)
    assert np.array_equal(o.get("qpos"), g.get("qpos"))
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())


@pytest.mark.parametrize("name", ["plane_box", "box_box", "capsule_capsule", "plane_cylinder"])
def test_pgs_primitive_match(name):
    """PGS solver (solver.rs:44-289) on the device against the oracle."""
    arr = mjcf.load_string(primitive_xml(name))
    arr["solver_kind"] = 1
    n = 64
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(13)
    q = o.get("qpos")
    q[:, :3] += rng.normal(size=(n, 3)) * 0.02
    v = rng.normal(size=(n, 6)) * 0.3
    for s in (o, g):
        s.set("qpos", 0, n, q)
        s.set("qvel", 0, n, v)
    say("pgs primitive", name)
    for t in range(250):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        err = np.abs(o.get("qpos") - g.get("qpos")).max()
        assert err <= 1e-9, (t, err)
    assert np.array_equal(o.get("qpos"), g.get("qpos"))


@pytest.mark.parametrize("name", sorted(MULTIBODY_SCENES))
def test_pgs_multibody_match(assets, name):
    xml = MULTIBODY_SCENES[name]
    arr = mjcf.load_file(assets / "box_drop.xml") if xml is None else mjcf.load_string(xml)
    arr["solver_kind"] = 1
    n = 32
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(4)
    nb = int(arr["nq"]) // 7
    q = o.get("qpos")
    for b in range(nb):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("pgs multibody", name)
    for t in range(300):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
    assert np.array_equal(o.get("qpos"), g.get("qpos"))


@pytest.mark.parametrize("name", ["arm3d.xml", "ant.xml"])
def test_implicitfast_match(assets, name):
    """integrator="implicitfast": actuator kv terms on tree-ancestor dof pairs
    and joint damping in the factorisation (forward.rs:392-456)."""
    xml = (assets / name).read_text().replace("<option ", '<option integrator="implicitfast" ', 1)
    arr = mjcf.load_string(xml)
    n = 48
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(17)
    v = rng.normal(size=(n, int(arr["nv"]))) * 0.5
    for s in (o, g):
        s.set("qvel", 0, n, v)
    nu = int(arr["nu"])
    say("implicitfast", name, n)
    contacts = 0
    for t in range(150):
        c = rng.uniform(-1, 1, size=(n, nu))
        for s in (o, g):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        contacts += int(o.ncon().sum())
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert np.array_equal(o.get("qvel"), g.get("qvel"))
    say("  contacts", contacts)


def test_gripper_match():
    """Arm + gripper + cube: fixed tendon, joint equality, implicitfast, limits,
    islands joined by contact."""
    from tests.scenes import GRIPPER

    arr = load_model(GRIPPER, is_xml=True)  # the library's own loader
    n = 32
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(31)
    q = o.get("qpos")
    q[:, 5:7] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("gripper", n)
    k = np.arange(n)
    for t in range(300):
        c = np.stack([np.sin(t * 0.01 + k), 0.4 + 0.3 * np.sin(t * 0.013 + 0 * k),
                      -0.9 + 0.5 * np.sin(t * 0.017 + 0 * k), 0.03 * np.sin(t * 0.02 + k)], 1)
        for s in (o, g):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())


@pytest.mark.parametrize("seed", [3, 11, 29, 41])
@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_random_piles_match(seed, solver):
    """Random primitive piles with per-env perturbed drops, Newton and PGS."""
    from tests.scenes import random_pile_xml

    arr = load_model(random_pile_xml(seed), is_xml=True)  # the library's own loader
    arr["solver_kind"] = 0 if solver == "newton" else 1
    n = 24
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(seed)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("random pile", seed, solver, n)
    for t in range(250):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("name", sorted(GENERAL_SCENES))
@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_general_pairs_match(name, solver):
    """Cylinder against sphere, capsule, box and cylinder through the general
    convex narrow phase (GJK/EPA and feature clipping), against the oracle's
    restatement of the same method."""
    arr = mjcf.load_string(general_xml(name))
    arr["solver_kind"] = 0 if solver == "newton" else 1
    n = 16
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(19)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.01
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("general pairs", name, solver, n)
    for t in range(250):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("seed", [2, 13, 18])
@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_random_piles_with_general_pairs_match(seed, solver):
    """Random piles whose cylinders meet every other body."""
    from pippin_amd.workloads import random_pile_xml

    arr = load_model(random_pile_xml(seed, general=True), is_xml=True)  # the library's own loader
    arr["solver_kind"] = 0 if solver == "newton" else 1
    n = 24
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(seed)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("random pile with general pairs", seed, solver, n)
    for t in range(250):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("solver", [0, 1])
def test_arm_gripper_with_per_env_draws_match(solver):
    """The 7-DoF arm with tendon-coupled gripper, link masses, cube size,
    resting height and friction drawn per env (copy-on-write parameter
    variants, one per env), 200 steps on smooth random targets; Newton and PGS."""
    from pippin_amd.workloads import arm_env_params, arm_gripper_xml, arm_targets

    arr = load_model(arm_gripper_xml(5), is_xml=True)  # the library's own loader
    arr["solver_kind"] = solver
    n = 24
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    params, qpos0 = arm_env_params(arr, n, 5)
    for s in (o, g):
        for name, gid, values in params:
            s.set_param(name, gid, 0, n, np.ascontiguousarray(values))
        s.set_param("qpos0", 0, 0, n, qpos0)
        s.reset(np.arange(n))
    assert np.array_equal(o.get("qpos"), g.get("qpos"))
    tg = arm_targets(5, n, 200, int(arr["nu"]))
    say("arm gripper", n)
    for t in range(200):
        for s in (o, g):
            s.set("ctrl", 0, n, tg[t])
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("name", sorted(MESH_SCENES))
@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_mesh_scenes_match(name, solver):
    """Convex hulls on the device: plane against hull (either order), hulls
    against every primitive and each other, models built by the library's own
    loader, against the oracle bit for bit."""
    arr = load_model(mesh_xml(name), is_xml=True)  # the library's own loader
    arr["solver_kind"] = 0 if solver == "newton" else 1
    n = 12
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(29)
    q = o.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("mesh scene", name, solver, n)
    for t in range(200):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert np.array_equal(o.solver_iterations(), g.solver_iterations())
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_tabletop_tumble_match(solver):
    """The tumble workload: per-env object sets (2 to 6 of box, sphere,
    cylinder, capsule and hull wedge, by contype masks), per-env size,
    density and friction, drops from per-env poses; 64 envs, 300 steps."""
    from pippin_amd.workloads import tabletop_tumble

    n = 64
    plan = tabletop_tumble(n, seed=3)
    arr = load_model(plan.xml, is_xml=True)
    arr["solver_kind"] = 0 if solver == "newton" else 1
    o, g = OracleSim(arr, n), HipPhysics(arr, n, max_contacts=96)
    for s in (o, g):
        plan.apply(s)
    assert np.array_equal(o.get("qpos"), g.get("qpos"))
    say("tabletop tumble", solver, n)
    for t in range(300):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
    assert o.ncon().sum() > 0


@pytest.mark.parametrize("solver", [0, 1])
def test_urdf_arm_matches(tmp_path, solver):
    """A URDF arm (revolute, continuous, prismatic and fixed joints, capsule,
    cylinder, sphere and box links over a box floor) imported by the library's
    URDF loader and stepped 300 steps from a perturbed pose falling onto the
    floor: positions and contact counts equal the oracle's at every step."""
    from pippin_amd.physics import load_urdf
    from tests.test_urdf import ARM

    path = tmp_path / "arm.urdf"
    path.write_text(ARM)
    arr = load_urdf(path)
    arr["solver_kind"] = solver
    n = 16
    o, g = OracleSim(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(11)
    q = o.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.3
    for s in (o, g):
        s.set("qpos", 0, n, q)
    say("urdf arm", n, "solver", solver)
    seen = 0
    for t in range(300):
        for s in (o, g):
            s.step(0, n, 1)
        assert np.array_equal(o.ncon(), g.ncon()), t
        assert np.array_equal(o.get("qpos"), g.get("qpos")), t
        seen += int(o.ncon().sum())
    assert seen > 0
