"""pippin_hip_reset_noalloc against pippin_hip_reset, through the C ABI (needs a GPU).

The no-alloc entry point stages the env ids in a fixed device array instead of
a buffer allocated per call; every other effect must be the same as
pippin_hip_reset's (Data::reset, crates/pippin/src/data.rs:122-131), bit for
bit, on consecutive calls of different sizes, and the state must step on
identically afterwards.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mjcf
from pippin_amd import physics
from pippin_amd.physics import STATE_FIELDS, HipPhysics

pytestmark = pytest.mark.gpu
MAX_IDS = 65536


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _same_state(a, b):
    for f in STATE_FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f
    ca, cb = a.get_history(), b.get_history()
    assert np.array_equal(ca[0], cb[0])
    assert np.array_equal(a.ncon(), b.ncon())


@pytest.mark.parametrize("model,solver", [("ant.xml", 0), ("box_drop.xml", 1)])
def test_reset_noalloc_matches_reset(assets, model, solver):
    arr = mjcf.load_file(assets / model)
    arr["solver_kind"] = solver
    n = 256
    a, b = HipPhysics(arr, n), HipPhysics(arr, n)
    rng = np.random.default_rng(8)
    q = a.get("qpos")
    if model == "ant.xml":
        q = q + rng.normal(size=q.shape) * 0.01
        q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    for s in (a, b):
        s.set("qpos", 0, n, q)

    def drive(steps):
        for _ in range(steps):
            c = rng.uniform(-1, 1, size=(n, a.nu)) if a.nu else None
            for s in (a, b):
                if c is not None:
                    s.set("ctrl", 0, n, c)
                s.step(0, n, 1)

    drive(60)
    _same_state(a, b)
    L = physics.lib()
    for k in (97, 13):  # two calls in a row, the second with fewer ids
        ids = np.ascontiguousarray(np.sort(rng.choice(n, k, replace=False)), dtype=np.int64)
        print(f"{model}: reset {k} envs", flush=True)
        assert L.pippin_hip_reset(a.h, _p(ids), len(ids)) == 0
        assert L.pippin_hip_reset_noalloc(b.h, _p(ids), len(ids)) == 0, physics._err()
        _same_state(a, b)
        assert np.array_equal(b.get("qpos")[ids], np.tile(arr["qpos_reset"], (k, 1)))
        drive(30)
        _same_state(a, b)


def test_reset_noalloc_rejects_too_many_ids(assets):
    arr = mjcf.load_file(assets / "ant.xml")
    n = 8
    g = HipPhysics(arr, n)
    g.step(0, n, 5)
    before = g.get("qpos")
    L = physics.lib()
    big = np.zeros(MAX_IDS + 1, dtype=np.int64)
    assert L.pippin_hip_reset_noalloc(g.h, _p(big), len(big)) == -1
    assert "at most" in physics._err()
    bad = np.array([3, n], dtype=np.int64)
    assert L.pippin_hip_reset_noalloc(g.h, _p(bad), len(bad)) == -1
    assert np.array_equal(g.get("qpos"), before)
    # the largest accepted call, on a second sim sharing the device's array
    h = HipPhysics(arr, MAX_IDS)
    h.step(0, MAX_IDS, 2)
    ids = np.arange(MAX_IDS, dtype=np.int64)
    assert L.pippin_hip_reset_noalloc(h.h, _p(ids), len(ids)) == 0
    assert np.array_equal(h.get("qpos"), np.tile(arr["qpos_reset"], (MAX_IDS, 1)))
    assert np.array_equal(g.get("qpos"), before)
