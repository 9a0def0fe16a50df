"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must still reproduce them bit for bit, and the
HIP kernels must reproduce them on the GPU without the oracle in the loop.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import mjcf

GOLDEN = Path(__file__).resolve().parent / "golden"
ASSETS = Path(__file__).resolve().parent.parent / "pippin_amd" / "assets"


def _load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_oracle_reproduces_the_golden_fixtures():
    from tests.golden.make_golden import build_all

    for name, arrays in build_all().items():
        ref = _load(name)
        assert set(ref) == set(arrays), name
        for k in ref:
            assert np.array_equal(ref[k], arrays[k]), (name, k)


def test_fixtures_are_not_trivial():
    a = _load("ant_rollout")
    assert a["ncon"].sum() > 100 and len(a["pairs"]) == a["ncon"].sum()
    g = _load("general_pairs")
    assert g["counts"].sum() == len(g["hits"]) and (g["counts"] > 0).sum() >= 6
    p = _load("pile_general")
    assert p["newton_ncon"].sum() > 0 and p["pgs_ncon"].sum() > 0


@pytest.mark.gpu
def test_device_reproduces_ant_rollout():
    from pippin_amd.physics import HipPhysics

    a = _load("ant_rollout")
    arr = mjcf.load_file(ASSETS / "ant.xml")
    n = a["qpos0"].shape[0]
    g = HipPhysics(arr, n)
    g.set("qpos", 0, n, a["qpos0"])
    pairs = []
    for t in range(a["ctrl"].shape[0]):
        g.set("ctrl", 0, n, a["ctrl"][t])
        g.step(0, n, 1)
        assert np.array_equal(g.ncon(), a["ncon"][t]), t
        for e in range(n):
            for c in g.contacts(e):
                pairs.append((int(c["geom1"]), int(c["geom2"])))
        if t % 10 == 9:
            assert np.array_equal(g.get("qpos"), a["qpos_every10"][t // 10]), t
    assert np.array_equal(np.array(pairs, dtype=np.int32).reshape(-1, 2), a["pairs"])


@pytest.mark.gpu
def test_device_reproduces_ant_mass_and_bias():
    from pippin_amd.physics import HipPhysics

    m = _load("ant_mass_bias")
    arr = mjcf.load_file(ASSETS / "ant.xml")
    n = m["qpos"].shape[0]
    g = HipPhysics(arr, n)
    g.set("qpos", 0, n, m["qpos"])
    g.set("qvel", 0, n, m["qvel"])
    qm, bias = g.mass_bias()
    assert np.array_equal(qm, m["qm"])
    assert np.array_equal(bias, m["bias"])


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["newton", "pgs"])
def test_device_reproduces_general_pile(solver):
    from pippin_amd.physics import HipPhysics
    from pippin_amd.workloads import random_pile_xml

    p = _load("pile_general")
    arr = mjcf.load_string(random_pile_xml(13, general=True))
    arr["solver_kind"] = 0 if solver == "newton" else 1
    q0 = p[f"{solver}_qpos0"]
    n = q0.shape[0]
    g = HipPhysics(arr, n)
    g.set("qpos", 0, n, q0)
    for t in range(p[f"{solver}_ncon"].shape[0]):
        g.step(0, n, 1)
        assert np.array_equal(g.ncon(), p[f"{solver}_ncon"][t]), t
    assert np.array_equal(g.get("qpos"), p[f"{solver}_qpos"])
