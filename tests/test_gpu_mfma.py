"""Newton's dense Hessian on the FP64 matrix cores (ph_options.hessian_mfma), needs a GPU.

The option replaces the reference's row-by-row sums H = M + sum (d_r J_ri) J_rj
(crates/pippin/src/newton.rs:279-300) by v_mfma_f64_16x16x4f64 tiles. Fused
accumulation rounds differently, so one step from the same state must land
within rounding distance of the oracle's step (and not on it bit for bit, which
shows the matrix cores were used); with the option off the step is the
oracle's, bit for bit.
"""
import sys

import numpy as np
import pytest

from oracle.pyoracle import OracleSim
from pippin_amd.physics import HipPhysics, load_model
from pippin_amd.workloads import random_pile_xml

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mfma_hessian_step_is_within_rounding_of_the_oracle():
    arr = load_model(random_pile_xml(29), is_xml=True)
    n = 48
    rng = np.random.default_rng(29)
    base = HipPhysics(arr, n)
    base.set_launch_shape(1, 0, "grid")
    q = base.get("qpos")
    for b in range(int(arr["nq"]) // 7):
        q[:, 7 * b:7 * b + 2] += rng.normal(size=(n, 2)) * 0.005
    base.set("qpos", 0, n, q)
    say("mfma hessian: settling the pile")
    base.step(0, n, 120)  # bodies land on each other: islands over 16 dofs
    state = base.save_state()
    o = OracleSim(arr, n)
    o.load_state({**state, "history": (state["history"][0], state["history"][1], state["history"][2])})
    plain, mfma = HipPhysics(arr, n), HipPhysics(arr, n, hessian_mfma=True)
    for s in (plain, mfma):
        s.set_launch_shape(1, 0, "grid")
        s.load_state(state)
    diffs = []
    for t in range(5):
        for s in (o, plain, mfma):
            s.step(0, n, 1)
        assert np.array_equal(plain.get("qvel"), o.get("qvel")), t
        d = float(np.abs(mfma.get("qvel") - o.get("qvel")).max())
        diffs.append(d)
        say("step", t, "max |dqvel|", d)
        # restart every step from the oracle's state: one step's worth of rounding
        mfma.load_state(o.save_state() | {"history": plain.get_history()})
    assert max(diffs) < 1e-9, diffs
    assert max(diffs) > 0.0, "the matrix-core Hessian was not used"
