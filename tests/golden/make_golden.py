"""Writes the golden fixtures in this folder from the oracle (oracle/, the CPU
restatement of the reference engine). Seeded and deterministic: rerunning it
reproduces the files bit for bit.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickled objects):
- ant_rollout.npz: Ant (pippin_amd/assets/ant.xml), 4 envs, qpos0 + N(0, 0.01)
  (seed 0), 200 steps with controls U[-1, 1] (seed 0): qpos every 10 steps,
  contact count per step, and the (geom1, geom2) pair of every contact of
  every step (flattened, in contact order).
- ant_mass_bias.npz: M(q) and c(q, qdot) at 20 random Ant states (seed 1).
- general_pairs.npz: contacts of posed cylinder pairs through the general
  convex narrow phase.
- pile_general.npz: pile seed 13 with every cylinder meeting every body
  (pippin_amd/workloads.py random_pile_xml), 4 envs, 120 Newton steps and 120
  PGS steps: contact counts per step and final qpos.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import mjcf  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from oracle.pyoracle import OracleSim  # noqa: E402

ASSETS = ROOT / "pippin_amd" / "assets"

# (type a, pos a, mat a, size a, type b, pos b, mat b, size b): sphere 1,
# capsule 2, box 3, cylinder 4; matrices row-major
I3 = np.eye(3).ravel()


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).ravel()


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]).ravel()


GENERAL_CASES = [
    (4, [0, 0, 0], I3, [.04, .06, 0], 1, [0.01, 0.005, .077], I3, [.02, 0, 0]),
    (1, [.057, 0, .01], I3, [.02, 0, 0], 4, [0, 0, 0], I3, [.04, .06, 0]),
    (3, [0, 0, 0], I3, [.05, .05, .05], 4, [0.01, 0, .108], _rz(0.3), [.03, .06, 0]),
    (3, [0, 0, 0], I3, [.2, .2, .05], 4, [0, 0, .089], _rx(np.pi / 2), [.04, .06, 0]),
    (4, [0, 0, 0], I3, [.04, .06, 0], 2, [0, 0, .068], _rx(np.pi / 2), [.01, .05, 0]),
    (4, [0, 0, 0], I3, [.04, .06, 0], 4, [.079, 0, .02], I3, [.04, .06, 0]),
    (4, [0, 0, 0], _rx(0.4), [.04, .06, 0], 4, [.03, .02, .1], _rx(1.2), [.03, .05, 0]),
    (3, [0, 0, 0], _rz(0.2), [.03, .02, .04], 4, [.02, .01, .085], _rx(0.3), [.02, .05, 0]),
]


def ant_rollout():
    arr = mjcf.load_file(ASSETS / "ant.xml")
    n, T = 4, 200
    s = OracleSim(arr, n)
    rng = np.random.default_rng(0)
    q = s.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.01
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    s.set("qpos", 0, n, q)
    ctrl = rng.uniform(-1, 1, size=(T, n, int(arr["nu"])))
    qs, ncon, pairs = [], [], []
    for t in range(T):
        s.set("ctrl", 0, n, ctrl[t])
        s.step(0, n, 1)
        ncon.append(s.ncon().copy())
        for e in range(n):
            for c in s.contacts(e):
                pairs.append((int(c["geom1"]), int(c["geom2"])))
        if t % 10 == 9:
            qs.append(s.get("qpos").copy())
    return dict(qpos0=q, ctrl=ctrl, qpos_every10=np.array(qs), ncon=np.array(ncon, dtype=np.int32),
                pairs=np.array(pairs, dtype=np.int32).reshape(-1, 2))


def ant_mass_bias():
    arr = mjcf.load_file(ASSETS / "ant.xml")
    n = 20
    s = OracleSim(arr, n)
    rng = np.random.default_rng(1)
    q = s.get("qpos") + rng.normal(size=(n, int(arr["nq"]))) * 0.3
    q[:, 3:7] /= np.linalg.norm(q[:, 3:7], axis=1, keepdims=True)
    v = rng.normal(size=(n, int(arr["nv"])))
    s.set("qpos", 0, n, q)
    s.set("qvel", 0, n, v)
    qm, bias = s.mass_bias()
    return dict(qpos=q, qvel=v, qm=qm, bias=bias)


def general_pairs():
    rows, counts = [], []
    for ta, pa, ma, sa, tb, pb, mb, sb in GENERAL_CASES:
        h = po.collide(ta, pa, ma, sa, tb, pb, mb, sb, margin=1e-3)
        counts.append(len(h))
        for c in h:
            rows.append(np.concatenate([c["pos"], c["normal"], [c["depth"]]]))
    return dict(counts=np.array(counts, dtype=np.int32), hits=np.array(rows).reshape(-1, 7))


def pile_general():
    from pippin_amd.workloads import random_pile_xml

    out = {}
    for solver, name in ((0, "newton"), (1, "pgs")):
        arr = mjcf.load_string(random_pile_xml(13, general=True))
        arr["solver_kind"] = solver
        n, T = 4, 120
        s = OracleSim(arr, n)
        rng = np.random.default_rng(13)
        q = s.get("qpos")
        for b in range(int(arr["nq"]) // 7):
            q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
        s.set("qpos", 0, n, q)
        ncon = []
        for t in range(T):
            s.step(0, n, 1)
            ncon.append(s.ncon().copy())
        out[f"{name}_qpos0"] = q
        out[f"{name}_ncon"] = np.array(ncon, dtype=np.int32)
        out[f"{name}_qpos"] = s.get("qpos").copy()
    return out


def build_all():
    return {"ant_rollout": ant_rollout(), "ant_mass_bias": ant_mass_bias(), "general_pairs": general_pairs(),
            "pile_general": pile_general()}


if __name__ == "__main__":
    for name, arrays in build_all().items():
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})
