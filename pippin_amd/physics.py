"""Python front end of libpippin_hip.so, mirroring Pippin's `Physics` trait.

`HipPhysics` exposes the methods of `trait Physics` (crates/pippin-env/src/lib.rs:110-155)
with PippinCpu's conventions (crates/pippin-env/src/pippin_cpu.rs): buffers shaped
(envs, width), size mismatches raise `ValueError` (the trait panics through
`check_len`), parameter errors raise `ParamError` (the trait returns `Err(String)`).
All stepping runs in the HIP kernels; there is no CPU path behind this class, and
constructing it fails loudly when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .desc import (APPEARANCE_DTYPE, CONTACT_DTYPE, POSE_DTYPE, Desc, arrays_from_desc, ph_dims,
                   ph_launch_shape, ph_model_desc, ph_options)

LIB_PATH = Path(__file__).resolve().parent / "libpippin_hip.so"
FIELD = {"qpos": 0, "qvel": 1, "ctrl": 2, "qacc": 3, "qfrc_applied": 4, "time": 5}
# the state a resumed run needs, in the order save_state stores it
STATE_FIELDS = ["qpos", "qvel", "ctrl", "qacc", "qfrc_applied", "time"]
PARAMS = ["geom_friction", "geom_size", "geom_rgba", "geom_contype", "geom_conaffinity", "body_mass",
          "dof_damping", "actuator_gain", "actuator_bias", "body_inertia", "body_ipos", "qpos0"]
PARAM = {n: i for i, n in enumerate(PARAMS)}
SOLVER = {"newton": 0, "pgs": 1}

PH_ERR_SIZE = -2

_lib = None
BOUND: list = []  # entry points given ctypes signatures


class HipError(RuntimeError):
    pass


class ParamError(ValueError):
    pass


def lib():
    """Load libpippin_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HipError(f"{LIB_PATH} is missing: build it with `python -m pippin_amd.build`")
        # One HIP runtime per process: torch ships its own libamdhip64 under the
        # same soname. Loading torch first binds this library to that copy, so
        # torch tensors and this library share the device; loading it second
        # would start a second runtime that finds no GPU.
        import torch  # noqa: F401

        L = C.CDLL(str(LIB_PATH))
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        sigs = {
            "pippin_hip_model_load_xml": ([C.c_char_p, C.c_char_p, C.POINTER(vp)], C.c_int),
            "pippin_hip_model_load_file": ([C.c_char_p, C.POINTER(vp)], C.c_int),
            "pippin_hip_model_load_urdf": ([C.c_char_p, i32, C.c_char_p, C.POINTER(vp)], C.c_int),
            "pippin_hip_urdf_to_mjcf": ([C.c_char_p, C.c_char_p, i32, C.c_char_p, C.c_char_p, i64, C.POINTER(i64)],
                                        C.c_int),
            "pippin_hip_model_desc": ([vp], C.POINTER(ph_model_desc)),
            "pippin_hip_model_free": ([vp], None),
            "pippin_hip_create": ([vp, i64, vp, C.POINTER(vp)], C.c_int),
            "pippin_hip_destroy": ([vp], None),
            "pippin_hip_dims": ([vp, vp], C.c_int),
            "pippin_hip_num_envs": ([vp], i64),
            "pippin_hip_step": ([vp, i64, i64, i64], C.c_int),
            "pippin_hip_reset": ([vp, vp, i64], C.c_int),
            "pippin_hip_reset_noalloc": ([vp, vp, i64], C.c_int),
            "pippin_hip_get": ([vp, i32, i64, i64, vp, i64], C.c_int),
            "pippin_hip_set": ([vp, i32, i64, i64, vp, i64], C.c_int),
            "pippin_hip_poses": ([vp, i64, i64, vp, i64, vp, i64], C.c_int),
            "pippin_hip_poses_device": ([vp, C.POINTER(vp), C.POINTER(vp)], C.c_int),
            "pippin_hip_set_param": ([vp, i32, i64, i64, i64, vp, i64], C.c_int),
            "pippin_hip_get_param": ([vp, i32, i64, i64, i64, vp, i64], C.c_int),
            "pippin_hip_appearance": ([vp, i64, i64, vp, i64], C.c_int),
            "pippin_hip_ncon": ([vp, i64, i64, vp, i64], C.c_int),
            "pippin_hip_contacts": ([vp, i64, vp, i32, vp], C.c_int),
            "pippin_hip_solver_iterations": ([vp, i64, i64, vp, i64], C.c_int),
            "pippin_hip_mass_bias": ([vp, i64, i64, vp, i64, vp, i64], C.c_int),
            "pippin_hip_invweight": ([vp, i64, i64, vp, i64, vp, i64], C.c_int),
            "pippin_hip_collision_pairs": ([vp, i64, vp, i32, vp], C.c_int),
            "pippin_hip_rollout_device": ([vp, i64, i64, i64, vp, vp], C.c_int),
            "pippin_hip_state_device": ([vp, i32, C.POINTER(vp)], C.c_int),
            "pippin_hip_stream": ([vp], vp),
            "pippin_hip_last_error": ([], C.c_char_p),
            "pippin_hip_selftest_math": ([vp, i64, vp, i32], C.c_int),
            "pippin_hip_set_device": ([vp, i32, i64, i64, vp], C.c_int),
            "pippin_hip_kernel_timing": ([vp, i32, C.POINTER(C.c_double), C.POINTER(i64)], C.c_int),
            "pippin_hip_get_launch_shape": ([vp, C.POINTER(ph_launch_shape)], C.c_int),
            "pippin_hip_set_launch_shape": ([vp, C.POINTER(ph_launch_shape)], C.c_int),
            "pippin_hip_autotune": ([vp, i32, C.c_char_p], C.c_int),
            "pippin_hip_autotune_pending": ([vp], C.c_int),
            "pippin_hip_use_graphs": ([vp, i32], C.c_int),
            "pippin_hip_mlp_init": ([vp, i32, C.c_uint64, i64, i64], C.c_int),
            "pippin_hip_mlp_mode": ([vp, i32], C.c_int),
            "pippin_hip_mlp_collect": ([vp, C.POINTER(i64)], C.c_int),
            "pippin_hip_mlp_train": ([vp, i32, i64, C.c_float, C.c_uint64, vp], C.c_int),
            "pippin_hip_mlp_num_params": ([vp], i64),
            "pippin_hip_mlp_get_params": ([vp, vp, i64], C.c_int),
            "pippin_hip_mlp_set_params": ([vp, vp, i64], C.c_int),
            "pippin_hip_mlp_forward": ([vp, vp, i64, vp], C.c_int),
            "pippin_hip_mlp_loss_grad": ([vp, vp, vp, i64, vp, vp], C.c_int),
            "pippin_hip_mlp_records": ([vp, i64, vp, vp, vp, vp], C.c_int),
            "pippin_hip_history_capacity": ([vp, C.POINTER(i32)], C.c_int),
            "pippin_hip_get_history": ([vp, i64, i64, i32, vp, vp, vp], C.c_int),
            "pippin_hip_set_history": ([vp, i64, i64, i32, vp, vp, vp], C.c_int),
        }
        global BOUND
        BOUND = sorted(sigs)
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _err():
    return lib().pippin_hip_last_error().decode()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def load_model(path_or_xml: str | Path, is_xml: bool = False) -> dict:
    """Parse MJCF with the library's loader; returns the raw descriptor as numpy arrays."""
    L = lib()
    h = C.c_void_p()
    if is_xml:
        rc = L.pippin_hip_model_load_xml(str(path_or_xml).encode(), b".", C.byref(h))
    else:
        rc = L.pippin_hip_model_load_file(str(path_or_xml).encode(), C.byref(h))
    if rc != 0:
        raise HipError(_err())
    try:
        return arrays_from_desc(L.pippin_hip_model_desc(h).contents)
    finally:
        L.pippin_hip_model_free(h)


def load_urdf(path: str | Path, floating_base: bool = False, package_dirs=()) -> dict:
    """Import a URDF file with the library's loader (pippin::urdf::load_file,
    urdf.rs:330-339): the raw descriptor as numpy arrays."""
    L = lib()
    h = C.c_void_p()
    dirs = ":".join(str(d) for d in package_dirs).encode()
    rc = L.pippin_hip_model_load_urdf(str(path).encode(), int(bool(floating_base)), dirs, C.byref(h))
    if rc != 0:
        raise HipError(_err())
    try:
        return arrays_from_desc(L.pippin_hip_model_desc(h).contents)
    finally:
        L.pippin_hip_model_free(h)


def urdf_to_mjcf(urdf: str, base: str | Path = ".", floating_base: bool = False, package_dirs=()) -> str:
    """The MJCF text the library makes of a URDF document (urdf::to_mjcf, urdf.rs:205)."""
    L = lib()
    n = C.c_int64(0)
    args = (urdf.encode(), str(base).encode(), int(bool(floating_base)), ":".join(str(d) for d in package_dirs).encode())
    rc = L.pippin_hip_urdf_to_mjcf(*args, None, 0, C.byref(n))
    if rc not in (0, -2) or (rc == -2 and n.value <= 0):
        raise HipError(_err())
    buf = C.create_string_buffer(n.value + 1)
    if L.pippin_hip_urdf_to_mjcf(*args, buf, n.value + 1, C.byref(n)) != 0:
        raise HipError(_err())
    return buf.value.decode()


def selftest_math(x, device: int = 0):
    """(sin, cos, sqrt|x|, 1/x) of each x as computed on the device."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((x.size, 4))
    if lib().pippin_hip_selftest_math(_p(x), x.size, _p(out), device) != 0:
        raise HipError(_err())
    return out


class HipPhysics:
    """A batch of environments stepped by the HIP kernels on one GPU."""

    name = "pippin-hip"

    def __init__(self, model, num_envs: int, device: int = 0, solver: str | None = None,
                 max_contacts: int = 0, hessian_mfma: bool = False):
        if isinstance(model, (str, Path)):
            model = load_model(model)
        self.arrays = model
        self._desc = Desc(model)
        opt = ph_options()
        opt.device = device
        opt.solver_override = -1 if solver is None else SOLVER[solver]
        opt.max_contacts = max_contacts
        opt.hessian_mfma = 1 if hessian_mfma else 0
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.pippin_hip_create(self._desc.ptr(), int(num_envs), C.byref(opt), C.byref(self.h))
        if rc != 0:
            self.h = None
            raise HipError(_err())
        d = ph_dims()
        self._check(self.L.pippin_hip_dims(self.h, C.byref(d)))
        self.nq, self.nv, self.nu, self.ngeom, self.ncam = d.nq, d.nv, d.nu, d.ngeom, d.ncam
        self.nbody = int(model["nbody"])
        self.n = int(num_envs)

    @classmethod
    def from_xml(cls, xml: str, num_envs: int, **kw):
        return cls(load_model(xml, is_xml=True), num_envs, **kw)

    def close(self):
        if getattr(self, "h", None):
            self.L.pippin_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc == PH_ERR_SIZE:
            raise ValueError(_err())
        if rc != 0:
            raise HipError(_err())

    def _range(self, begin, end):
        return (0, self.n) if begin is None else (begin, self.n if end is None else end)

    # ---- Physics trait ----
    def num_envs(self):
        return self.n

    def dims(self):
        return dict(nq=self.nq, nv=self.nv, nu=self.nu, ngeom=self.ngeom, ncam=self.ncam)

    def width(self, field):
        return {"qpos": self.nq, "qvel": self.nv, "ctrl": self.nu, "qacc": self.nv, "qfrc_applied": self.nv,
                "time": 1}[field]

    def step(self, begin, end, nstep=1):
        self._check(self.L.pippin_hip_step(self.h, begin, end, nstep))

    # ---- checkpoint/resume (no trait counterpart) ----
    def history_capacity(self):
        c = C.c_int32(0)
        self._check(self.L.pippin_hip_history_capacity(self.h, C.byref(c)))
        return c.value

    def get_history(self, begin=None, end=None, cap=None):
        """PGS contact history per env: counts (n,), contacts (n, cap), impulses (n, cap, 3)."""
        begin, end = self._range(begin, end)
        cap = self.history_capacity() if cap is None else cap
        counts = np.zeros(end - begin, dtype=np.int32)
        con = np.zeros((end - begin, cap), dtype=CONTACT_DTYPE)
        lam = np.zeros((end - begin, cap, 3))
        self._check(self.L.pippin_hip_get_history(self.h, begin, end, cap, _p(counts), _p(con), _p(lam)))
        return counts, con, lam

    def set_history(self, begin, end, counts, contacts, lam):
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        contacts = np.ascontiguousarray(contacts, dtype=CONTACT_DTYPE)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        cap = contacts.shape[1]
        if counts.shape != (end - begin,) or contacts.shape[0] != end - begin or lam.shape != (end - begin, cap, 3):
            raise ValueError("set_history: arrays do not match the env range and capacity")
        self._check(self.L.pippin_hip_set_history(self.h, begin, end, cap, _p(counts), _p(contacts), _p(lam)))

    def save_state(self, begin=None, end=None):
        """Everything a resumed run needs to continue bit for bit (Data's carried state)."""
        begin, end = self._range(begin, end)
        st = {f: self.get(f, begin, end) for f in STATE_FIELDS}
        st["history"] = self.get_history(begin, end)
        return st

    def load_state(self, st, begin=None, end=None):
        begin, end = self._range(begin, end)
        for f in STATE_FIELDS:
            self.set(f, begin, end, st[f])
        self.set_history(begin, end, *st["history"])

    def reset(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        self._check(self.L.pippin_hip_reset(self.h, _p(ids), len(ids)))

    def get(self, field, begin=None, end=None):
        begin, end = self._range(begin, end)
        out = np.zeros((end - begin, self.width(field)))
        self._check(self.L.pippin_hip_get(self.h, FIELD[field], begin, end, _p(out), out.size))
        return out

    def set(self, field, begin, end, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        self._check(self.L.pippin_hip_set(self.h, FIELD[field], begin, end, _p(src), src.size))

    def poses(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        g = np.zeros((end - begin, self.ngeom), dtype=POSE_DTYPE)
        c = np.zeros((end - begin, self.ncam), dtype=POSE_DTYPE)
        self._check(self.L.pippin_hip_poses(self.h, begin, end, _p(g), g.size, _p(c), c.size))
        return g, c

    def poses_device(self):
        """Device addresses of the sim's f32 pose buffers (envs x ngeom, envs x ncam),
        refreshed to the current state."""
        g, c = C.c_void_p(), C.c_void_p()
        self._check(self.L.pippin_hip_poses_device(self.h, C.byref(g), C.byref(c)))
        return g.value, c.value

    def param_width(self, p):
        return {"geom_size": 3, "actuator_bias": 3, "body_inertia": 3, "body_ipos": 3,
                "geom_rgba": 4, "qpos0": self.nq}.get(p, 1)

    def set_param(self, p, pid, begin, end, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        rc = self.L.pippin_hip_set_param(self.h, PARAM[p], pid, begin, end, _p(v), v.size)
        if rc == -1:
            raise ParamError(_err())
        self._check(rc)

    def get_param(self, p, pid, begin=None, end=None):
        begin, end = self._range(begin, end)
        out = np.zeros((end - begin, self.param_width(p)))
        rc = self.L.pippin_hip_get_param(self.h, PARAM[p], pid, begin, end, _p(out), out.size)
        if rc == -1:
            raise ParamError(_err())
        self._check(rc)
        return out

    def appearance(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        out = np.zeros((end - begin, self.ngeom), dtype=APPEARANCE_DTYPE)
        self._check(self.L.pippin_hip_appearance(self.h, begin, end, _p(out), out.size))
        return out

    # ---- diagnostics ----
    def ncon(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        out = np.zeros(end - begin, dtype=np.int32)
        self._check(self.L.pippin_hip_ncon(self.h, begin, end, _p(out), out.size))
        return out

    def contacts(self, env):
        n = C.c_int32(0)
        buf = np.zeros(512, dtype=CONTACT_DTYPE)
        self._check(self.L.pippin_hip_contacts(self.h, env, _p(buf), len(buf), C.byref(n)))
        return buf[: n.value].copy()

    def solver_iterations(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        out = np.zeros(end - begin, dtype=np.int32)
        self._check(self.L.pippin_hip_solver_iterations(self.h, begin, end, _p(out), out.size))
        return out

    def mass_bias(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        qm = np.zeros((end - begin, self.nv, self.nv))
        bias = np.zeros((end - begin, self.nv))
        self._check(self.L.pippin_hip_mass_bias(self.h, begin, end, _p(qm), qm.size, _p(bias), bias.size))
        return qm, bias

    def invweight(self, begin=None, end=None):
        begin, end = self._range(begin, end)
        dof = np.zeros((end - begin, self.nv))
        body = np.zeros((end - begin, self.nbody, 2))
        self._check(self.L.pippin_hip_invweight(self.h, begin, end, _p(dof), dof.size, _p(body), body.size))
        return dof, body

    def collision_pairs(self, env=0):
        n = C.c_int32(0)
        buf = np.zeros((8192, 2), dtype=np.int32)
        self._check(self.L.pippin_hip_collision_pairs(self.h, env, _p(buf), len(buf), C.byref(n)))
        return buf[: n.value].copy()

    # ---- device-resident paths ----
    def state_ptr(self, field):
        p = C.c_void_p()
        self._check(self.L.pippin_hip_state_device(self.h, FIELD[field], C.byref(p)))
        return p.value

    def set_device(self, field, begin, end, dev_ptr: int):
        """Copy a field from device memory, ordered before the next step."""
        self._check(self.L.pippin_hip_set_device(self.h, FIELD[field], begin, end, C.c_void_p(dev_ptr)))

    SCHED = {"grid": 0, "ws": 1, "ws_bucket": 2, "pgs_split": 3}

    def launch_shape(self):
        """(envs per wavefront, specialised kernel, schedule name) of the plain step."""
        sh = ph_launch_shape()
        self._check(self.L.pippin_hip_get_launch_shape(self.h, C.byref(sh)))
        names = {v: k for k, v in self.SCHED.items()}
        return int(sh.envs_per_wave), int(sh.specialised), names[int(sh.schedule)]

    def set_launch_shape(self, envs_per_wave: int, specialised: int, schedule: str):
        sh = ph_launch_shape(int(envs_per_wave), int(specialised), self.SCHED[schedule], 0)
        if self.L.pippin_hip_set_launch_shape(self.h, C.byref(sh)) != 0:
            raise HipError(_err())

    def autotune(self, rounds: int = 3, cache_path: str | None = None):
        """Try every launch shape on the next plain steps, then keep the fastest
        (results are identical under every shape)."""
        path = None if cache_path is None else str(cache_path).encode()
        self._check(self.L.pippin_hip_autotune(self.h, int(rounds), path))

    def autotune_pending(self) -> bool:
        return bool(self.L.pippin_hip_autotune_pending(self.h))

    def use_graphs(self, on: bool):
        self._check(self.L.pippin_hip_use_graphs(self.h, int(bool(on))))

    def kernel_timing(self, reset=False):
        ms, n = C.c_double(0), C.c_int64(0)
        self._check(self.L.pippin_hip_kernel_timing(self.h, int(reset), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def rollout_device(self, ctrl_ptr: int, qpos_out_ptr: int | None, nstep: int, begin=0, end=None):
        """Batch::rollout with device pointers (e.g. torch.cuda tensors' data_ptr())."""
        end = self.n if end is None else end
        self._check(self.L.pippin_hip_rollout_device(self.h, begin, end, nstep, C.c_void_p(ctrl_ptr),
                                                      C.c_void_p(qpos_out_ptr) if qpos_out_ptr else None))

    # ---- learned Newton warm start (no counterpart in the reference; mlp.hpp) ----
    MLP_LOG, MLP_WARM = 1, 2

    def mlp_init(self, hidden=64, seed=0, capacity=None, max_batch=8192):
        cap = capacity if capacity is not None else max(2_000_000, 16 * self.n, max_batch)
        self._check(self.L.pippin_hip_mlp_init(self.h, hidden, seed, cap, max_batch))

    def mlp_mode(self, mode: int):
        self._check(self.L.pippin_hip_mlp_mode(self.h, mode))

    def mlp_collect(self) -> int:
        n = C.c_int64(0)
        self._check(self.L.pippin_hip_mlp_collect(self.h, C.byref(n)))
        return n.value

    def mlp_train(self, iters, batch=4096, lr=1e-3, seed=0):
        loss = np.zeros(iters, dtype=np.float32)
        self._check(self.L.pippin_hip_mlp_train(self.h, iters, batch, lr, seed, _p(loss)))
        return loss

    def mlp_params(self):
        n = self.L.pippin_hip_mlp_num_params(self.h)
        out = np.zeros(n, dtype=np.float32)
        self._check(self.L.pippin_hip_mlp_get_params(self.h, _p(out), n))
        return out

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.pippin_hip_mlp_set_params(self.h, _p(p), p.size))

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 4), dtype=np.float32)
        self._check(self.L.pippin_hip_mlp_forward(self.h, _p(x), x.shape[0], _p(y)))
        return y

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.L.pippin_hip_mlp_num_params(self.h), dtype=np.float32)
        loss = C.c_float(0.0)
        self._check(self.L.pippin_hip_mlp_loss_grad(self.h, _p(x), _p(t), x.shape[0], _p(g), C.byref(loss)))
        return loss.value, g

    def mlp_records(self, env):
        f = np.zeros((16, 24), dtype=np.float32)
        pr = np.zeros((16, 4), dtype=np.float32)
        tg = np.zeros((16, 4), dtype=np.float32)
        c = C.c_int32(0)
        self._check(self.L.pippin_hip_mlp_records(self.h, env, _p(f), _p(pr), _p(tg), C.byref(c)))
        k = c.value
        return f[:k], pr[:k], tg[:k]
