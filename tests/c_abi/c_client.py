"""Run the plain C client (ant_golden.c) on the golden Ant rollout and compare.

The client is started as a fresh child process: no torch and no other HIP
runtime is loaded in it, so the library binds ROCm's own, as a Rust caller's
process would.
"""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
EXE = HERE / "_build" / "ant_golden"
GOLDEN = ROOT / "tests" / "golden" / "ant_rollout.npz"
MODEL = ROOT / "pippin_amd" / "assets" / "ant.xml"


def run_c_client(timeout: float = 300.0) -> str:
    """Steps the golden Ant rollout through the C ABI in a child process and
    checks qpos (every 10 steps) and contact counts bit for bit. Returns the
    child's output line; raises AssertionError on any difference."""
    if not EXE.exists():
        raise AssertionError(f"{EXE} is missing: build it with `make -C tests/c_abi`")
    g = np.load(GOLDEN)
    q0, ctrl = g["qpos0"], g["ctrl"]
    n, nq = q0.shape
    T, _, nu = ctrl.shape
    with tempfile.TemporaryDirectory() as d:
        inp, out = Path(d) / "in.bin", Path(d) / "out.bin"
        with open(inp, "wb") as f:
            np.array([n, T, nq, nu], dtype=np.int64).tofile(f)
            np.ascontiguousarray(q0, dtype=np.float64).tofile(f)
            np.ascontiguousarray(ctrl, dtype=np.float64).tofile(f)
        r = subprocess.run([str(EXE), str(MODEL), str(inp), str(out)], capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise AssertionError(f"C client exited with {r.returncode}: {r.stderr.strip()}")
        raw = out.read_bytes()
    nqb = T // 10 * n * nq * 8
    q = np.frombuffer(raw[:nqb], dtype=np.float64).reshape(T // 10, n, nq)
    nc = np.frombuffer(raw[nqb:], dtype=np.int32).reshape(T, n)
    if not np.array_equal(nc, g["ncon"]):
        raise AssertionError("C client: contact counts differ from the golden rollout")
    if not np.array_equal(q, g["qpos_every10"]):
        err = float(np.abs(q - g["qpos_every10"]).max())
        raise AssertionError(f"C client: qpos differs from the golden rollout by {err}")
    return r.stdout.strip()
