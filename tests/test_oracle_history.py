"""Checkpoint/resume of the oracle's carried state, and pins of PGS's warm
start (CPU only).

PGS warm-starts each new contact from the nearest previous contact of the same
geom pair within 1 cm, re-projecting the previous friction impulse onto the new
contact's friction frame (crates/pippin/src/solver.rs:253-289). With one soft
sweep (`iterations` = 1, then two relaxation sweeps) the result of a step
depends on that start, so planted histories pin the matching rule through
exact equalities between steps from the same state:
  - a history whose contacts all moved by 0.9 cm is matched as if unmoved
    (the same impulses, so the same step, bit for bit);
  - moved by 1.1 cm, or with another geom pair, nothing matches: the step
    equals a cold start (empty history), bit for bit;
  - the same friction force written in a rotated tangent frame gives the same
    step up to rounding, and differs from feeding the impulses unprojected.
"""
import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim

INCLINE = """<mujoco><option timestep="0.002"/><worldbody>
    <geom name="floor" type="plane" size="5 5 .1" euler="0 25 0" friction="0.9"/>
    <body name="b" pos="0 0 0.06" euler="0 25 0"><freejoint/><geom name="box" type="box" size=".05 .05 .05" friction="0.9"/></body>
</worldbody></mujoco>"""


def _settled_incline(iterations):
    arr = mjcf.load_string(INCLINE)
    arr["solver_kind"] = 1
    s = OracleSim(arr, 1)
    s.step(0, 1, 600)  # the box comes to rest, sticking on the slope
    st = s.save_state()
    arr1 = dict(arr)
    arr1["solver_iterations"] = iterations
    return arr1, st


def _step_from(arr, st, history):
    s = OracleSim(arr, 1)
    s.load_state(st)
    s.set_history(0, 1, *history)
    s.step(0, 1, 1)
    return s.get("qvel")[0]


def _empty(cap=8):
    return (np.zeros(1, dtype=np.int32), np.zeros((1, cap), dtype=st_dtype()), np.zeros((1, cap, 3)))


def st_dtype():
    from pippin_amd.desc import CONTACT_DTYPE
    return CONTACT_DTYPE


def _frame(n, t):
    """Contact::frame (collision.rs:47-55) for a contact whose preferred tangent is set."""
    t = t - n * (n @ t)
    t1 = t / np.linalg.norm(t)
    return t1, np.cross(n, t1)


def test_warm_start_matching_rule():
    arr, st = _settled_incline(iterations=1)
    counts, con, lam = st["history"]
    nc = int(counts[0])
    assert nc == 4 and np.all(lam[0, :nc, 0] > 0)
    assert np.max(np.abs(lam[0, :nc, 1:])) > 1e-4  # friction holds the box on the slope
    base = _step_from(arr, st, (counts, con, lam))
    cold = _step_from(arr, st, _empty(con.shape[1]))
    assert np.max(np.abs(base - cold)) > 1e-7, "the warm start must matter for this check to bite"

    n0 = con[0, 0]["normal"]
    t1, _ = _frame(n0, np.array([1.0, 0.3, -0.2]))
    for shift, matched in ((0.009, True), (0.011, False)):
        moved = con.copy()
        moved[0, :nc]["pos"] += shift * t1
        got = _step_from(arr, st, (counts, moved, lam))
        assert np.array_equal(got, base if matched else cold), (shift, matched)

    other = con.copy()
    other[0, :nc]["geom1"], other[0, :nc]["geom2"] = con[0, :nc]["geom2"], con[0, :nc]["geom1"]
    assert np.array_equal(_step_from(arr, st, (counts, other, lam)), cold)


def test_warm_start_reprojects_friction_onto_the_new_frame():
    arr, st = _settled_incline(iterations=1)
    counts, con, lam = st["history"]
    nc = int(counts[0])
    base = _step_from(arr, st, (counts, con, lam))
    cold = _step_from(arr, st, _empty(con.shape[1]))
    rot, lam_rot = con.copy(), lam.copy()
    for k in range(nc):
        n = con[0, k]["normal"]
        # the saved impulses are in the frame of a contact without a preferred
        # tangent (Vec3::tangents, math.rs:65-70); the same world force f is
        # written in a frame turned about the normal by th
        p1, p2 = _tangents(n)
        f = p1 * lam[0, k, 1] + p2 * lam[0, k, 2]
        th = 0.7 + 0.3 * k
        q1, q2 = _frame(n, np.cos(th) * p1 + np.sin(th) * p2)
        rot[0, k]["tangent"] = q1
        lam_rot[0, k, 1], lam_rot[0, k, 2] = f @ q1, f @ q2
    got = _step_from(arr, st, (counts, rot, lam_rot))
    scale = np.max(np.abs(base - cold))
    assert np.max(np.abs(got - base)) < 1e-9 * scale + 1e-15, (got - base)
    # the turned frame's numbers taken as if in the contact's own frame (no re-projection)
    wrong = _step_from(arr, st, (counts, con, lam_rot))
    assert np.max(np.abs(wrong - base)) > 1e-3 * scale


def _tangents(n):
    """Two unit vectors completing an orthonormal frame with n: n x X, or n x Y
    when n is close to X (math.rs:65-70)."""
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.57 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(n, t1)


@pytest.mark.parametrize("model,solver", [("ant.xml", 0), ("box_drop.xml", 1), ("box_drop.xml", 0)])
def test_save_and_load_resume_bit_for_bit(assets, model, solver):
    arr = mjcf.load_file(assets / model)
    arr["solver_kind"] = solver
    n = 4
    a, b = OracleSim(arr, n), OracleSim(arr, n)
    rng = np.random.default_rng(2)
    q = a.get("qpos")
    if model == "ant.xml":
        q = q + rng.normal(size=q.shape) * 0.01
        q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    a.set("qpos", 0, n, q)
    for t in range(150):
        if a.nu:
            a.set("ctrl", 0, n, rng.uniform(-1, 1, size=(n, a.nu)))
        a.step(0, n, 1)
    st = a.save_state()
    assert st["time"][0, 0] == pytest.approx(150 * arr["timestep"])
    if solver == 1:
        assert st["history"][0].sum() > 0
    b.load_state(st)
    for t in range(60):
        c = rng.uniform(-1, 1, size=(n, a.nu)) if a.nu else None
        for s in (a, b):
            if c is not None:
                s.set("ctrl", 0, n, c)
            s.step(0, n, 1)
    assert np.array_equal(a.get("qpos"), b.get("qpos"))
    assert np.array_equal(a.get("qvel"), b.get("qvel"))
    # without the warm-start state the resumed run departs
    cst = {k: v for k, v in st.items()}
    cst["qacc"] = np.zeros_like(st["qacc"])
    cst["history"] = (np.zeros_like(st["history"][0]),) + st["history"][1:]
    c2 = OracleSim(arr, n)
    c2.load_state(cst)
    rng2 = np.random.default_rng(2)
    for t in range(150):
        if a.nu:
            rng2.uniform(-1, 1, size=(n, a.nu))
    for t in range(60):
        if a.nu:
            c2.set("ctrl", 0, n, rng2.uniform(-1, 1, size=(n, a.nu)))
        c2.step(0, n, 1)
    assert not np.array_equal(a.get("qvel"), c2.get("qvel"))
