"""Test-only ctypes front end of the serial kernel emulation (tests/emu/emu.hip)."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from pippin_amd.desc import CONTACT_DTYPE, Desc

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libemu.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        # the emulation links libamdhip64; bind it to torch's copy (one HIP runtime per process)
        import torch  # noqa: F401

        L = C.CDLL(str(LIB))
        vp, i64 = C.c_void_p, C.c_int64
        L.emu_create.restype = vp
        L.emu_create.argtypes = [vp, i64]
        L.emu_destroy.argtypes = [vp]
        L.emu_last_error.restype = C.c_char_p
        for n, a in {"emu_step": [vp, i64, i64, i64], "emu_get": [vp, C.c_int, i64, i64, vp],
                     "emu_set": [vp, C.c_int, i64, i64, vp], "emu_ncon": [vp, i64, i64, vp],
                     "emu_iters": [vp, i64, i64, vp], "emu_status": [vp, i64, i64, vp],
                     "emu_contacts": [vp, i64, vp, C.c_int32, vp], "emu_mass_bias": [vp, i64, i64, vp, vp],
                     "emu_invweight": [vp, vp, vp], "emu_poses": [vp, i64, i64, vp, vp],
                     "emu_mlp_set": [vp, C.c_int, C.c_int, vp],
                     "emu_mlp_records": [vp, i64, vp, vp, vp, vp]}.items():
            getattr(L, n).argtypes = a
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class EmuSim:
    FIELD = {"qpos": 0, "qvel": 1, "ctrl": 2}

    def __init__(self, arrays: dict, n: int):
        self.L = lib()
        self.desc = Desc(arrays)
        self.h = self.L.emu_create(self.desc.ptr(), n)
        if not self.h:
            raise RuntimeError(self.L.emu_last_error().decode())
        self.n = n
        self.nq, self.nv, self.nu = int(arrays["nq"]), int(arrays["nv"]), int(arrays["nu"])
        self.nbody, self.ngeom = int(arrays["nbody"]), int(arrays["ngeom"])

    def __del__(self):
        if getattr(self, "h", None):
            self.L.emu_destroy(self.h)

    def width(self, f):
        return {"qpos": self.nq, "qvel": self.nv, "ctrl": self.nu}[f]

    def step(self, b, e, nstep=1):
        self.L.emu_step(self.h, b, e, nstep)

    def get(self, f, b=0, e=None):
        e = self.n if e is None else e
        out = np.zeros((e - b, self.width(f)))
        self.L.emu_get(self.h, self.FIELD[f], b, e, _p(out))
        return out

    def set(self, f, b, e, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        self.L.emu_set(self.h, self.FIELD[f], b, e, _p(src))

    def _ints(self, fn, b, e):
        e = self.n if e is None else e
        out = np.zeros(e - b, dtype=np.int32)
        fn(self.h, b, e, _p(out))
        return out

    def ncon(self, b=0, e=None):
        return self._ints(self.L.emu_ncon, b, e)

    def solver_iterations(self, b=0, e=None):
        return self._ints(self.L.emu_iters, b, e)

    def status(self, b=0, e=None):
        return self._ints(self.L.emu_status, b, e)

    def contacts(self, env):
        n = C.c_int32(0)
        buf = np.zeros(512, dtype=CONTACT_DTYPE)
        self.L.emu_contacts(self.h, env, _p(buf), 512, C.byref(n))
        return buf[: n.value].copy()

    def mass_bias(self, b=0, e=None):
        e = self.n if e is None else e
        qm = np.zeros((e - b, self.nv, self.nv))
        bias = np.zeros((e - b, self.nv))
        self.L.emu_mass_bias(self.h, b, e, _p(qm), _p(bias))
        return qm, bias

    def invweight(self):
        dof = np.zeros(self.nv)
        body = np.zeros((self.nbody, 2))
        self.L.emu_invweight(self.h, _p(dof), _p(body))
        return dof, body

    def poses(self, b=0, e=None):
        e = self.n if e is None else e
        gp = np.zeros((e - b, self.ngeom, 3))
        gm = np.zeros((e - b, self.ngeom, 9))
        self.L.emu_poses(self.h, b, e, _p(gp), _p(gm))
        return gp, gm

    # ---- learned warm start (the kernel's network stage, run on the host) ----
    def mlp_set(self, h, mode, params):
        self._mlp_params = np.ascontiguousarray(params, dtype=np.float32)
        self.L.emu_mlp_set(self.h, h, mode, _p(self._mlp_params))

    def mlp_records(self, env):
        f = np.zeros((16, 24), dtype=np.float32)
        pr = np.zeros((16, 4), dtype=np.float32)
        tg = np.zeros((16, 4), dtype=np.float32)
        c = C.c_int32(0)
        self.L.emu_mlp_records(self.h, env, _p(f), _p(pr), _p(tg), C.byref(c))
        k = c.value
        return f[:k], pr[:k], tg[:k]
