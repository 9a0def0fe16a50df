"""Checkpoint/resume of the device state, including the warm-start state
(needs a GPU).

A run split by save_state / load_state must equal the uninterrupted run bit for
bit, on the device as on the oracle: Newton carries qacc between steps
(crates/pippin/src/newton.rs:238-245) and PGS its contact history
(Scratch::prev_contacts / prev_lambda, crates/pippin/src/data.rs:60-80, matched
by solver.rs:253-289). States move between the oracle and the device in both
directions, and planted histories give the same step on both.
"""
import sys

import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim
from pippin_amd.physics import HipError, HipPhysics
from pippin_amd.workloads import random_pile_xml

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _drive(sims, n, steps, rng, nu):
    for _ in range(steps):
        c = rng.uniform(-1, 1, size=(n, nu)) if nu else None
        for s in sims:
            if c is not None:
                s.set("ctrl", 0, n, c)
            s.step(0, n, 1)


def _models(assets):
    ant = mjcf.load_file(assets / "ant.xml")
    pile = mjcf.load_string(random_pile_xml(29))
    pile["solver_kind"] = 1
    return [("ant newton", ant), ("pile pgs", pile)]


@pytest.mark.parametrize("which", [0, 1])
def test_split_run_equals_uninterrupted_run(assets, which):
    name, arr = _models(assets)[which]
    n = 24
    say("checkpoint", name)
    g, o = HipPhysics(arr, n), OracleSim(arr, n)
    rng = np.random.default_rng(4)
    q = o.get("qpos")
    if name.startswith("ant"):
        q = q + rng.normal(size=q.shape) * 0.01
        q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    else:
        for b in range(int(arr["nq"]) // 7):
            q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    for s in (g, o):
        s.set("qpos", 0, n, q)
    nu = int(arr["nu"])
    _drive((g, o), n, 120, rng, nu)
    st = g.save_state()
    so = o.save_state()
    for f in ("qpos", "qvel", "ctrl", "qacc", "qfrc_applied", "time"):
        assert np.array_equal(st[f], so[f]), f
    gc, gcon, glam = st["history"]
    oc, ocon, olam = so["history"]
    assert np.array_equal(gc, oc)
    if name.endswith("pgs"):
        assert gc.sum() > 0
    for e in range(n):
        k = int(gc[e])
        for fld in ("pos", "normal", "tangent", "depth", "geom1", "geom2", "feature"):
            assert np.array_equal(gcon[e, :k][fld], ocon[e, :k][fld]), (e, fld)
        assert np.array_equal(glam[e, :k], olam[e, :k]), e
    # resume on a fresh device sim from the device's state, and on another from the oracle's
    g2, g3 = HipPhysics(arr, n), HipPhysics(arr, n)
    g2.load_state(st)
    g3.load_state({**so, "history": (oc, ocon[:, :g3.history_capacity()], olam[:, :g3.history_capacity()])})
    state_rng = rng.bit_generator.state
    _drive((g, o), n, 80, rng, nu)
    rng.bit_generator.state = state_rng
    _drive((g2, g3), n, 80, rng, nu)
    for s in (g2, g3, o):
        assert np.array_equal(s.get("qpos"), g.get("qpos"))
        assert np.array_equal(s.get("qvel"), g.get("qvel"))


def test_planted_history_steps_like_the_oracle():
    """One PGS sweep, so the warm start shapes the step: histories planted on
    both sides (as saved, moved within and beyond the 1 cm match radius,
    another geom pair, cleared) give the same step bit for bit."""
    xml = """<mujoco><option timestep="0.002"/><worldbody>
        <geom name="floor" type="plane" size="5 5 .1" euler="0 25 0" friction="0.9"/>
        <body pos="0 0 0.06" euler="0 25 0"><freejoint/><geom type="box" size=".05 .05 .05" friction="0.9"/></body>
    </worldbody></mujoco>"""
    arr = mjcf.load_string(xml)
    arr["solver_kind"] = 1
    o = OracleSim(arr, 1)
    o.step(0, 1, 600)
    st = o.save_state()
    arr1 = dict(arr)
    arr1["solver_iterations"] = 1
    counts, con, lam = st["history"]
    g0 = HipPhysics(arr1, 1)
    cap = g0.history_capacity()
    g0.close()
    nc = int(counts[0])
    con, lam = con[:, :cap], lam[:, :cap]
    variants = {"saved": (counts, con, lam), "cleared": (np.zeros_like(counts), con, lam)}
    for shift in (0.009, 0.011):
        c2 = con.copy()
        c2[0, :nc]["pos"] += shift * np.array([0.6, 0.8, 0.0])
        variants[f"moved {shift}"] = (counts, c2, lam)
    c3 = con.copy()
    c3[0, :nc]["geom1"], c3[0, :nc]["geom2"] = con[0, :nc]["geom2"], con[0, :nc]["geom1"]
    variants["swapped pair"] = (counts, c3, lam)
    outs = {}
    for name, h in variants.items():
        got = []
        for s in (HipPhysics(arr1, 1), OracleSim(arr1, 1)):
            s.load_state({**st, "history": h})
            s.step(0, 1, 1)
            got.append(s.get("qvel")[0])
        assert np.array_equal(got[0], got[1]), name
        outs[name] = got[0]
    assert not np.array_equal(outs["saved"], outs["cleared"])
    assert np.array_equal(outs["moved 0.009"], outs["saved"])
    assert np.array_equal(outs["moved 0.011"], outs["cleared"])
    assert np.array_equal(outs["swapped pair"], outs["cleared"])


def test_set_history_validates_before_changing_anything(assets):
    arr = mjcf.load_file(assets / "box_drop.xml")
    arr["solver_kind"] = 1
    n = 3
    g = HipPhysics(arr, n)
    g.step(0, n, 200)
    before = g.get_history()
    cap = g.history_capacity()
    counts, con, lam = (x.copy() for x in before)
    counts[1] = cap + 1
    with pytest.raises(HipError):
        g.set_history(0, n, counts, con, lam)
    counts, con, lam = (x.copy() for x in before)
    con[2, 0]["geom1"] = 999
    counts[2] = max(counts[2], 1)
    with pytest.raises(HipError):
        g.set_history(0, n, counts, con, lam)
    counts, con, lam = (x.copy() for x in before)
    lam[0, 0, 0] = np.nan
    counts[0] = max(counts[0], 1)
    with pytest.raises(HipError):
        g.set_history(0, n, counts, con, lam)
    after = g.get_history()
    for a, b in zip(before, after):
        assert np.array_equal(a, b, equal_nan=False) if a.dtype.names is None else \
            all(np.array_equal(a[f], b[f]) for f in a.dtype.names)
    # reset clears the history (Data::reset, data.rs:122-131) and keeps qacc
    qacc = g.get("qacc")
    g.reset([1])
    c2, _, _ = g.get_history()
    assert c2[1] == 0 and c2[0] == before[0][0]
    assert np.array_equal(g.get("qacc"), qacc)
