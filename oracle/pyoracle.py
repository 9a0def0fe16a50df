"""ORACLE — test infrastructure only. ctypes front end of oracle/_build/liborc.so.

`OracleSim` mirrors the `Physics` trait (crates/pippin-env/src/lib.rs:110-155)
over the C++ restatement of PippinCpu (crates/pippin-env/src/pippin_cpu.rs), so
the tests drive the oracle and the HIP backend through the same calls.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from pippin_amd.desc import (APPEARANCE_DTYPE, CONTACT_DTYPE, POSE_DTYPE, Desc)

from . import mjcf

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liborc.so"
_lib = None

FIELD = {"qpos": 0, "qvel": 1, "ctrl": 2, "qacc": 3, "qfrc_applied": 4, "time": 5}
STATE_FIELDS = ["qpos", "qvel", "ctrl", "qacc", "qfrc_applied", "time"]
PARAM = {n: i for i, n in enumerate([
    "geom_friction", "geom_size", "geom_rgba", "geom_contype", "geom_conaffinity", "body_mass",
    "dof_damping", "actuator_gain", "actuator_bias", "body_inertia", "body_ipos", "qpos0"])}


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def use_host_tuned_build() -> bool:
    """Load the x86-64-v3 build (same arithmetic, vectorised) when the host CPU
    has AVX2 and nothing has loaded the library yet; returns whether it did."""
    global LIB_PATH
    if _lib is not None:
        return LIB_PATH.name == "liborc_v3.so"
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    v3 = HERE / "_build" / "liborc_v3.so"
    if " avx2" in flags and " bmi2" in flags and " fma" in flags and v3.exists():
        LIB_PATH = v3
        return True
    return False


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, i64, i32, f64p = C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, i64]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        for name, args in {
            "orc_step": [vp, i64, i64, i64, C.c_int],
            "orc_rollout": [vp, i64, i64, i64, vp, vp, C.c_int],
            "orc_reset": [vp, vp, i64],
            "orc_get": [vp, C.c_int, i64, i64, vp],
            "orc_set": [vp, C.c_int, i64, i64, vp],
            "orc_set_applied": [vp, i64, vp],
            "orc_poses": [vp, i64, i64, vp, vp],
            "orc_set_param": [vp, C.c_int, i64, i64, i64, vp],
            "orc_get_param": [vp, C.c_int, i64, i64, i64, vp],
            "orc_appearance": [vp, i64, i64, vp],
            "orc_ncon": [vp, i64, i64, vp],
            "orc_contacts": [vp, i64, vp, i32, vp],
            "orc_step_stats": [vp, i64, i64, vp, vp, vp],
            "orc_mass_bias": [vp, i64, i64, vp, vp],
            "orc_invweight": [vp, i64, i64, vp, vp],
            "orc_collision_pairs": [vp, i64, vp, i32, vp],
            "orc_model_info": [vp, i64, vp, vp, vp, vp, vp, vp, vp],
            "orc_cholesky_block": [vp, i64, i64, i64],
            "orc_cholesky_solve_block": [vp, i64, i64, i64, vp],
            "orc_quat_from_axis_angle": [vp, C.c_double, vp],
            "orc_quat_to_mat": [vp, vp],
            "orc_quat_from_mat": [vp, vp],
            "orc_quat_integrate": [vp, vp, C.c_double, vp],
            "orc_collide": [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, C.c_double, vp, i32],
            "orc_get_history": [vp, i64, i64, i32, vp, vp, vp],
            "orc_set_history": [vp, i64, i64, i32, vp, vp, vp],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            if name.startswith("orc_quat") or name == "orc_cholesky_solve_block":
                fn.restype = None
            else:
                fn.restype = C.c_int
        L.orc_impedance.restype = C.c_double
        L.orc_impedance.argtypes = [vp, C.c_double]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [vp, vp, i64]
        L.orc_num_envs.restype = i64
        L.orc_num_envs.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    pass


class OracleSim:
    """PippinCpu restated in C++: `num_envs` environments sharing one model."""

    name = "pippin-cpu-oracle"

    def __init__(self, model: dict | str | Path, num_envs: int):
        if not isinstance(model, dict):
            model = mjcf.load_file(model)
        self.arrays = model
        self.desc = Desc(model)
        self.L = lib()
        self.h = self.L.orc_create(self.desc.ptr(), int(num_envs))
        self.n = int(num_envs)
        self.nq, self.nv, self.nu = int(model["nq"]), int(model["nv"]), int(model["nu"])
        self.ngeom, self.ncam, self.nbody = int(model["ngeom"]), int(model["ncam"]), int(model["nbody"])

    @classmethod
    def from_xml(cls, xml: str, num_envs: int):
        return cls(mjcf.load_string(xml), num_envs)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.L.orc_last_error().decode())

    def num_envs(self):
        return self.n

    def dims(self):
        return dict(nq=self.nq, nv=self.nv, nu=self.nu, ngeom=self.ngeom, ncam=self.ncam)

    def width(self, field):
        return {"qpos": self.nq, "qvel": self.nv, "ctrl": self.nu, "qacc": self.nv, "qfrc_applied": self.nv,
                "time": 1}[field]

    def get_history(self, begin=0, end=None, cap=256):
        """Scratch::prev_contacts / prev_lambda per env: counts, contacts (n, cap), impulses (n, cap, 3)."""
        end = self.n if end is None else end
        counts = np.zeros(end - begin, dtype=np.int32)
        con = np.zeros((end - begin, cap), dtype=CONTACT_DTYPE)
        lam = np.zeros((end - begin, cap, 3))
        self._check(self.L.orc_get_history(self.h, begin, end, cap, _p(counts), _p(con), _p(lam)))
        return counts, con, lam

    def set_history(self, begin, end, counts, contacts, lam):
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        contacts = np.ascontiguousarray(contacts, dtype=CONTACT_DTYPE)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._check(self.L.orc_set_history(self.h, begin, end, contacts.shape[1], _p(counts), _p(contacts),
                                           _p(lam)))

    def save_state(self, begin=0, end=None):
        end = self.n if end is None else end
        st = {f: self.get(f, begin, end) for f in STATE_FIELDS}
        st["history"] = self.get_history(begin, end)
        return st

    def load_state(self, st, begin=0, end=None):
        end = self.n if end is None else end
        for f in STATE_FIELDS:
            self.set(f, begin, end, st[f])
        self.set_history(begin, end, *st["history"])

    def step(self, begin, end, nstep=1, threads=1):
        self._check(self.L.orc_step(self.h, begin, end, nstep, threads))

    def rollout(self, ctrl, begin=0, end=None, threads=1, want_qpos=True):
        end = self.n if end is None else end
        ctrl = np.ascontiguousarray(ctrl, dtype=np.float64)
        nstep = ctrl.shape[1]
        out = np.zeros((end - begin, nstep, self.nq)) if want_qpos else None
        self._check(self.L.orc_rollout(self.h, begin, end, nstep, _p(ctrl),
                                       _p(out) if want_qpos else None, threads))
        return out

    def reset(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        self._check(self.L.orc_reset(self.h, _p(ids), len(ids)))

    def get(self, field, begin=0, end=None):
        end = self.n if end is None else end
        out = np.zeros((end - begin, self.width(field)))
        self._check(self.L.orc_get(self.h, FIELD[field], begin, end, _p(out)))
        return out

    def set(self, field, begin, end, src):
        src = np.ascontiguousarray(src, dtype=np.float64).reshape(end - begin, self.width(field))
        self._check(self.L.orc_set(self.h, FIELD[field], begin, end, _p(src)))

    def set_applied(self, env, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._check(self.L.orc_set_applied(self.h, env, _p(f)))

    def poses(self, begin=0, end=None):
        end = self.n if end is None else end
        g = np.zeros((end - begin, self.ngeom), dtype=POSE_DTYPE)
        c = np.zeros((end - begin, self.ncam), dtype=POSE_DTYPE)
        self._check(self.L.orc_poses(self.h, begin, end, _p(g), _p(c)))
        return g, c

    def param_width(self, p):
        return {"geom_size": 3, "actuator_bias": 3, "body_inertia": 3, "body_ipos": 3,
                "geom_rgba": 4, "qpos0": self.nq}.get(p, 1)

    def set_param(self, p, pid, begin, end, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        self._check(self.L.orc_set_param(self.h, PARAM[p], pid, begin, end, _p(v)))

    def get_param(self, p, pid, begin=0, end=None):
        end = self.n if end is None else end
        out = np.zeros((end - begin, self.param_width(p)))
        self._check(self.L.orc_get_param(self.h, PARAM[p], pid, begin, end, _p(out)))
        return out

    def appearance(self, begin=0, end=None):
        end = self.n if end is None else end
        out = np.zeros((end - begin, self.ngeom), dtype=APPEARANCE_DTYPE)
        self._check(self.L.orc_appearance(self.h, begin, end, _p(out)))
        return out

    def ncon(self, begin=0, end=None):
        end = self.n if end is None else end
        out = np.zeros(end - begin, dtype=np.int32)
        self._check(self.L.orc_ncon(self.h, begin, end, _p(out)))
        return out

    def contacts(self, env):
        n = C.c_int32(0)
        buf = np.zeros(256, dtype=CONTACT_DTYPE)
        self._check(self.L.orc_contacts(self.h, env, _p(buf), len(buf), C.byref(n)))
        return buf[: n.value].copy()

    def step_stats(self, begin=0, end=None):
        end = self.n if end is None else end
        it, nr, ls = (np.zeros(end - begin, dtype=np.int32) for _ in range(3))
        self._check(self.L.orc_step_stats(self.h, begin, end, _p(it), _p(nr), _p(ls)))
        return it, nr, ls

    def solver_iterations(self, begin=0, end=None):
        return self.step_stats(begin, end)[0]

    def mass_bias(self, begin=0, end=None):
        end = self.n if end is None else end
        qm = np.zeros((end - begin, self.nv, self.nv))
        bias = np.zeros((end - begin, self.nv))
        self._check(self.L.orc_mass_bias(self.h, begin, end, _p(qm), _p(bias)))
        return qm, bias

    def invweight(self, begin=0, end=None):
        end = self.n if end is None else end
        dof = np.zeros((end - begin, self.nv))
        body = np.zeros((end - begin, self.nbody, 2))
        self._check(self.L.orc_invweight(self.h, begin, end, _p(dof), _p(body)))
        return dof, body

    def collision_pairs(self, env=0):
        n = C.c_int32(0)
        buf = np.zeros((4096, 2), dtype=np.int32)
        self._check(self.L.orc_collision_pairs(self.h, env, _p(buf), len(buf), C.byref(n)))
        return buf[: n.value].copy()

    def model_info(self, env=0):
        nb, nv = self.nbody, self.nv
        mass = np.zeros(nb)
        ipos = np.zeros((nb, 3))
        inertia = np.zeros((nb, 9))
        dof_parent = np.zeros(max(nv, 1), dtype=np.int32)
        weld = np.zeros(nb, dtype=np.int32)
        blocks = np.zeros((max(nv, 1), 2), dtype=np.int32)
        nblk = C.c_int32(0)
        self._check(self.L.orc_model_info(self.h, env, _p(mass), _p(ipos), _p(inertia), _p(dof_parent),
                                          _p(weld), _p(blocks), C.byref(nblk)))
        return dict(body_mass=mass, body_ipos=ipos, body_inertia=inertia, dof_parent=dof_parent[:nv],
                    body_weldid=weld, dof_blocks=blocks[: nblk.value].copy())


# ---- known-answer hooks ----
def impedance(solimp, pos):
    s = np.ascontiguousarray(solimp, dtype=np.float64)
    return lib().orc_impedance(_p(s), float(pos))


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_dot(_p(a), _p(b), min(len(a), len(b)))


def cholesky_block(a, stride, s0, e):
    return bool(lib().orc_cholesky_block(_p(a), stride, s0, e))


def cholesky_solve_block(l, stride, s0, e, x):
    lib().orc_cholesky_solve_block(_p(l), stride, s0, e, _p(x))


def quat_from_axis_angle(axis, angle):
    a = np.ascontiguousarray(axis, dtype=np.float64)
    q = np.zeros(4)
    lib().orc_quat_from_axis_angle(_p(a), float(angle), _p(q))
    return q


def quat_to_mat(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    m = np.zeros(9)
    lib().orc_quat_to_mat(_p(q), _p(m))
    return m


def quat_from_mat(m):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    q = np.zeros(4)
    lib().orc_quat_from_mat(_p(m), _p(q))
    return q


def quat_integrate(q, w, h):
    q = np.ascontiguousarray(q, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_quat_integrate(_p(q), _p(w), float(h), _p(out))
    return out


def cr_sincos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    L = lib()
    L.orc_cr_sincos.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.orc_cr_sincos.restype = None
    L.orc_cr_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


def collide(ta, pa, ma, sa, tb, pb, mb, sb, margin=1e-3):
    arr = lambda x, n: np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n))
    pa, ma, sa, pb, mb, sb = arr(pa, 3), arr(ma, 9), arr(sa, 3), arr(pb, 3), arr(mb, 9), arr(sb, 3)
    out = np.zeros(16, dtype=CONTACT_DTYPE)
    n = lib().orc_collide(ta, _p(pa), _p(ma), _p(sa), tb, _p(pb), _p(mb), _p(sb), float(margin), _p(out), 16)
    return out[:n].copy()


# ---- op-counting build (liborc_count.so) ----
_clib = None


def count_lib():
    global _clib
    if _clib is None:
        path = HERE / "_build" / "liborc_count.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.orcc_create.restype = C.c_void_p
        L.orcc_create.argtypes = [C.c_void_p, C.c_int64]
        L.orcc_destroy.argtypes = [C.c_void_p]
        L.orcc_set.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_void_p]
        L.orcc_step.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_ulonglong)]
        _clib = L
    return _clib


def ops_per_env_step(arrays: dict, qpos: np.ndarray, ctrl_seq: np.ndarray, substeps: int, skip: int = 0) -> float:
    """FP64 operations of the reference step, averaged over envs and steps.

    qpos: (n, nq) initial state; ctrl_seq: (actions, n, nu), held for `substeps` steps
    each. The first `skip` actions are stepped but not counted.
    """
    L = count_lib()
    d = Desc(arrays)
    n = qpos.shape[0]
    h = L.orcc_create(d.ptr(), n)
    try:
        q = np.ascontiguousarray(qpos, dtype=np.float64)
        L.orcc_set(h, FIELD["qpos"], 0, n, _p(q))
        total = 0
        for a in range(ctrl_seq.shape[0]):
            c = np.ascontiguousarray(ctrl_seq[a], dtype=np.float64)
            L.orcc_set(h, FIELD["ctrl"], 0, n, _p(c))
            ops = C.c_ulonglong(0)
            if L.orcc_step(h, 0, n, substeps, C.byref(ops)) != 0:
                raise OracleError("mass matrix is not positive definite")
            if a >= skip:
                total += ops.value
        return total / (n * substeps * (ctrl_seq.shape[0] - skip))
    finally:
        L.orcc_destroy(h)
