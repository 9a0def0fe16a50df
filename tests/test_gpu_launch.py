"""Launch shapes, autotuning and graph replay of the step (needs a GPU).

Every shape (envs per wavefront, generic or model-specialised kernel, grid,
work-stealing or bucketed work-stealing schedule) computes every env with the
same arithmetic, so trajectories must be bit-identical under all of them, also
while the autotuner switches shapes between launches and when launches are
replayed from captured graphs.
"""
import sys

import numpy as np
import pytest

from oracle import mjcf
from pippin_amd.physics import HipError, HipPhysics
from pippin_amd.workloads import random_pile_xml

pytestmark = pytest.mark.gpu

SHAPES = [(k, sp, sc) for sp in (1, 0) for k in (4, 2, 1) for sc in ("grid", "ws", "ws_bucket")]
SPLIT = (1, 0, "pgs_split")  # PGS models only: the step as front, sweeps and back kernels


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ant(assets, n, seed):
    arr = mjcf.load_file(assets / "ant.xml")
    rng = np.random.default_rng(seed)
    g = HipPhysics(arr, n)
    q = g.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.01
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    return arr, q


def _run(arr, q, n, shape, calls, seed, graphs=True, ranges=None):
    g = HipPhysics(arr, n)
    if shape is not None:
        g.set_launch_shape(*shape)
    g.use_graphs(graphs)
    g.set("qpos", 0, n, q)
    rng = np.random.default_rng(seed)
    nu = int(arr["nu"])
    out = []
    for t in range(calls):
        if nu:
            g.set("ctrl", 0, n, rng.uniform(-1, 1, size=(n, nu)))
        b, e = ranges[t % len(ranges)] if ranges else (0, n)
        g.step(b, e, 4)
        out.append((g.get("qpos").copy(), g.ncon().copy()))
    g.close()
    return out


def _valid_shapes(arr, n):
    """The shapes the library accepts for this model (LDS, specialised kernel).
    Only set_launch_shape may refuse: a failure while stepping a valid shape
    fails the test."""
    g = HipPhysics(arr, n)
    ok = []
    for shape in SHAPES + [SPLIT]:
        try:
            g.set_launch_shape(*shape)
        except HipError:
            continue
        ok.append(shape)
    g.close()
    return ok


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


def test_every_shape_matches_on_ant(assets):
    n = 13  # not a multiple of 2 or 4: ragged last wavefront
    arr, q = _ant(assets, n, 3)
    ranges = [(0, n), (0, n), (3, n), (0, 9)]
    ref = _run(arr, q, n, (1, 0, "grid"), 24, 5, graphs=False, ranges=ranges)
    assert sum(int(c.sum()) for _, c in ref) > 0
    shapes = _valid_shapes(arr, n)
    # every generic shape, and the Ant-specialised kernel at 4 envs per wavefront
    assert len(shapes) == 12, shapes
    for shape in shapes:
        say("ant shape", shape)
        got = _run(arr, q, n, shape, 24, 5, ranges=ranges)
        assert _same(ref, got), shape


@pytest.mark.parametrize("solver", [0, 1])
def test_every_shape_matches_on_a_pile(solver):
    arr = mjcf.load_string(random_pile_xml(29))
    arr["solver_kind"] = solver
    n = 11
    rng = np.random.default_rng(29)
    g = HipPhysics(arr, n)
    q = g.get("qpos")
    g.close()
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    ref = _run(arr, q, n, None, 30, 7, graphs=False)
    shapes = _valid_shapes(arr, n)
    assert len(shapes) >= 3, shapes  # at least one K with every schedule
    assert (SPLIT in shapes) == (solver == 1), shapes
    for shape in shapes:
        say("pile shape", solver, shape)
        got = _run(arr, q, n, shape, 30, 7)
        assert _same(ref, got), shape


def test_autotune_keeps_results_and_caches_its_choice(assets, tmp_path):
    n = 64
    arr, q = _ant(assets, n, 11)
    cache = tmp_path / "shapes.txt"
    plain = _run(arr, q, n, None, 1, 13)  # warm the kernels
    g, ref = HipPhysics(arr, n), HipPhysics(arr, n)
    for s in (g, ref):
        s.set("qpos", 0, n, q)
    g.autotune(rounds=1, cache_path=cache)
    rng = np.random.default_rng(17)
    calls = 0
    while g.autotune_pending():
        c = rng.uniform(-1, 1, size=(n, int(arr["nu"])))
        for s in (g, ref):
            s.set("ctrl", 0, n, c)
            s.step(0, n, 2)
        assert np.array_equal(g.get("qpos"), ref.get("qpos")), calls
        calls += 1
        assert calls < 200
    k, sp, sc = g.launch_shape()
    say("autotuned", (k, sp, sc), "after", calls, "launches")
    assert calls >= 10 and k in (1, 2, 4) and sc in ("grid", "ws", "ws_bucket")
    assert cache.read_text().strip().endswith(f"{k} {sp} {HipPhysics.SCHED[sc]}")
    h = HipPhysics(arr, n)
    h.autotune(rounds=1, cache_path=cache)
    assert not h.autotune_pending() and h.launch_shape() == (k, sp, sc)
    for s in (g, ref, h):
        s.close()
    del plain


@pytest.mark.parametrize("schedule", ["grid", "ws_bucket"])
def test_graph_replays_match_eager_launches(assets, schedule):
    """Replays of captured step graphs, with the program's own eager work
    (controls copied in, state read back) between them, against the same
    launches made eagerly."""
    n = 16
    arr, q = _ant(assets, n, 21)
    ranges = [(0, n), (4, n), (0, n), (0, 8)]
    eager = _run(arr, q, n, (4, 0, schedule), 24, 23, graphs=False, ranges=ranges)
    graph = _run(arr, q, n, (4, 0, schedule), 24, 23, graphs=True, ranges=ranges)
    for t, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]), t
        assert np.array_equal(a[1], b[1]), t


@pytest.mark.parametrize("solver", [0, 1])
def test_rollout_matches_under_every_shape_and_while_autotuning(solver):
    """Batch::rollout with per-step controls and qpos outputs in device
    memory, under every shape and during autotuning trials (both solvers: the
    split PGS step takes part for PGS)."""
    import torch

    arr = mjcf.load_string(random_pile_xml(29))
    arr["solver_kind"] = solver
    n, T = 11, 12
    rng = np.random.default_rng(31)
    q0 = None
    outs = []
    shapes = [None] + [s for s in _valid_shapes(arr, n) if s[1] == 0] + ["autotune"]
    for shape in shapes:
        g = HipPhysics(arr, n)
        if q0 is None:
            q0 = g.get("qpos")
            for b in range(int(arr["nq"]) // 7):
                q0[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
        g.set("qpos", 0, n, q0)
        if shape == "autotune":
            g.autotune(rounds=1)
        elif shape is not None:
            g.set_launch_shape(*shape)
        nu = max(int(arr["nu"]), 1)
        ctrl = torch.zeros((n, T, nu), dtype=torch.float64, device="cuda")
        qo = torch.zeros((n, T, int(arr["nq"])), dtype=torch.float64, device="cuda")
        res = []
        for call in range(12):
            g.rollout_device(ctrl.data_ptr(), qo.data_ptr(), T)
            torch.cuda.synchronize()
            res.append(qo.cpu().numpy().copy())
        outs.append(res)
        say("rollout shape", shape)
        g.close()
    for shape, res in zip(shapes[1:], outs[1:]):
        for a, b in zip(outs[0], res):
            assert np.array_equal(a, b), shape


def test_autotune_cache_is_keyed_by_env_count(assets, tmp_path):
    """A shape tuned for one env count is not taken for another: the key holds
    the model's sizes, the env count and the device."""
    arr = mjcf.load_file(assets / "ant.xml")
    cache = tmp_path / "shapes.txt"
    cache.write_text("v1 not-this-model : 1 0 0\n")
    g = HipPhysics(arr, 32)
    g.autotune(rounds=1, cache_path=cache)
    assert g.autotune_pending()
    while g.autotune_pending():
        g.step(0, 32, 1)
    lines = cache.read_text().strip().splitlines()
    assert len(lines) == 2 and " n32 " in lines[1]
    h = HipPhysics(arr, 48)
    h.autotune(rounds=1, cache_path=cache)
    assert h.autotune_pending()
    k = HipPhysics(arr, 32)
    k.autotune(rounds=1, cache_path=cache)
    assert not k.autotune_pending() and k.launch_shape() == g.launch_shape()
    for s in (g, h, k):
        s.close()
