"""The C ABI from a plain C process with no torch in it (needs a GPU).

physics.py loads torch before the library so that both share torch's HIP
runtime; a Rust FFI caller loads ROCm's own. The C client in tests/c_abi runs
the golden Ant rollout (tests/golden/ant_rollout.npz, written by the oracle)
through pippin_hip_model_load_file / create / set / step / ncon / get in a
fresh child process, and must reproduce it bit for bit.
"""
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "c_abi"))
from c_client import run_c_client  # noqa: E402

pytestmark = pytest.mark.gpu


def test_plain_c_client_reproduces_the_golden_ant_rollout():
    print(run_c_client(), flush=True)
