"""ctypes mirror of the plain-C structures in include/pippin_hip.h.

`ph_model_desc` is the raw compiled-model descriptor that crosses the C ABI:
the fields Pippin's MJCF compiler fills before `Model::compile`
(crates/pippin/src/model.rs:103-249). `desc_from_arrays` packs a dict of numpy
arrays into one, keeping the arrays alive on the returned object.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)

# (name, ctype, numpy dtype, width) for the array fields, in header order.
ARRAY_FIELDS = [
    ("body_parent", _i32p, np.int32, 1), ("body_pos", _f64p, np.float64, 3),
    ("body_quat", _f64p, np.float64, 4), ("body_jntadr", _i32p, np.int32, 1),
    ("body_jntnum", _i32p, np.int32, 1), ("body_dofadr", _i32p, np.int32, 1),
    ("body_dofnum", _i32p, np.int32, 1), ("body_mass", _f64p, np.float64, 1),
    ("body_ipos", _f64p, np.float64, 3), ("body_inertia", _f64p, np.float64, 9),
    ("jnt_type", _i32p, np.int32, 1), ("jnt_body", _i32p, np.int32, 1),
    ("jnt_qposadr", _i32p, np.int32, 1), ("jnt_dofadr", _i32p, np.int32, 1),
    ("jnt_pos", _f64p, np.float64, 3), ("jnt_axis", _f64p, np.float64, 3),
    ("jnt_limited", _i32p, np.int32, 1), ("jnt_range", _f64p, np.float64, 2),
    ("jnt_stiffness", _f64p, np.float64, 1), ("jnt_springref", _f64p, np.float64, 1),
    ("dof_body", _i32p, np.int32, 1), ("dof_jnt", _i32p, np.int32, 1),
    ("dof_damping", _f64p, np.float64, 1), ("dof_armature", _f64p, np.float64, 1),
    ("geom_type", _i32p, np.int32, 1), ("geom_body", _i32p, np.int32, 1),
    ("geom_pos", _f64p, np.float64, 3), ("geom_quat", _f64p, np.float64, 4),
    ("geom_size", _f64p, np.float64, 3), ("geom_friction", _f64p, np.float64, 1),
    ("geom_contype", _u32p, np.uint32, 1), ("geom_conaffinity", _u32p, np.uint32, 1),
    ("geom_rgba", _f32p, np.float32, 4), ("geom_group", _i32p, np.int32, 1),
    ("geom_dataid", _i32p, np.int32, 1),
    ("cam_body", _i32p, np.int32, 1), ("cam_pos", _f64p, np.float64, 3),
    ("cam_quat", _f64p, np.float64, 4), ("cam_fovy", _f64p, np.float64, 1),
    ("actuator_gear", _f64p, np.float64, 1), ("actuator_gain", _f64p, np.float64, 1),
    ("actuator_bias", _f64p, np.float64, 3), ("actuator_ctrllimited", _i32p, np.int32, 1),
    ("actuator_ctrlrange", _f64p, np.float64, 2), ("actuator_forcelimited", _i32p, np.int32, 1),
    ("actuator_forcerange", _f64p, np.float64, 2), ("actuator_moment_adr", _i32p, np.int32, 1),
    ("actuator_moment_num", _i32p, np.int32, 1), ("moment_jnt", _i32p, np.int32, 1),
    ("moment_coef", _f64p, np.float64, 1),
    ("eq_joint1", _i32p, np.int32, 1), ("eq_joint2", _i32p, np.int32, 1),
    ("eq_polycoef", _f64p, np.float64, 5), ("eq_solref", _f64p, np.float64, 2),
    ("eq_solimp", _f64p, np.float64, 5),
    ("exclude_pairs", _i32p, np.int32, 2),
    ("qpos0", _f64p, np.float64, 1), ("qpos_reset", _f64p, np.float64, 1),
]

SCALAR_INTS = ["nq", "nv", "nu", "nbody", "njnt", "ngeom", "ncam", "neq", "nexclude", "nmoment"]
# mesh hulls: counts, then arrays (name, ctype, dtype, width, count field)
MESH_INTS = ["nmesh", "nmeshvert", "nmeshface", "nmeshfacevert"]
MESH_FIELDS = [
    ("mesh_vertadr", _i32p, np.int32, 1, "nmesh"), ("mesh_vertnum", _i32p, np.int32, 1, "nmesh"),
    ("mesh_faceadr", _i32p, np.int32, 1, "nmesh"), ("mesh_facenum", _i32p, np.int32, 1, "nmesh"),
    ("mesh_vert", _f64p, np.float64, 3, "nmeshvert"), ("face_vertadr", _i32p, np.int32, 1, "nmeshface"),
    ("face_vertnum", _i32p, np.int32, 1, "nmeshface"), ("face_vert", _i32p, np.int32, 1, "nmeshfacevert"),
]


def no_meshes() -> dict:
    """The mesh fields of a model without mesh assets."""
    out = {n: 0 for n in MESH_INTS}
    for name, _c, dtype, width, _n in MESH_FIELDS:
        out[name] = np.zeros((0, width) if width > 1 else 0, dtype=dtype)
    return out


class ph_model_desc(C.Structure):
    _fields_ = (
        [(n, C.c_int32) for n in SCALAR_INTS]
        + [("timestep", C.c_double), ("gravity", C.c_double * 3), ("integrator", C.c_int32),
           ("solver_kind", C.c_int32), ("solver_iterations", C.c_int32),
           ("tolerance", C.c_double), ("contact_margin", C.c_double),
           ("solref", C.c_double * 2), ("solimp", C.c_double * 5), ("impratio", C.c_double),
           ("contact_hertz", C.c_double), ("contact_damping_ratio", C.c_double),
           ("contact_slop", C.c_double)]
        + [(n, t) for (n, t, _, _) in ARRAY_FIELDS]
        + [(n, C.c_int32) for n in MESH_INTS]
        + [(n, t) for (n, t, _, _, _) in MESH_FIELDS]
    )


class ph_pose(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("mat", C.c_float * 9)]


class ph_appearance(C.Structure):
    _fields_ = [("rgba", C.c_float * 4), ("scale", C.c_float * 3), ("pad", C.c_float)]


class ph_dims(C.Structure):
    _fields_ = [("nq", C.c_int64), ("nv", C.c_int64), ("nu", C.c_int64), ("ngeom", C.c_int64),
                ("ncam", C.c_int64)]


class ph_contact(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("normal", C.c_double * 3), ("tangent", C.c_double * 3),
                ("depth", C.c_double), ("geom1", C.c_int32), ("geom2", C.c_int32),
                ("feature", C.c_int32), ("pad", C.c_int32)]


class ph_launch_shape(C.Structure):
    _fields_ = [("envs_per_wave", C.c_int32), ("specialised", C.c_int32), ("schedule", C.c_int32),
                ("pad", C.c_int32)]


class ph_options(C.Structure):
    _fields_ = [("device", C.c_int32), ("solver_override", C.c_int32), ("max_contacts", C.c_int32),
                ("hessian_mfma", C.c_int32), ("reserved", C.c_int32 * 4)]


CONTACT_DTYPE = np.dtype([("pos", "<f8", 3), ("normal", "<f8", 3), ("tangent", "<f8", 3),
                          ("depth", "<f8"), ("geom1", "<i4"), ("geom2", "<i4"),
                          ("feature", "<i4"), ("pad", "<i4")])
POSE_DTYPE = np.dtype([("pos", "<f4", 3), ("mat", "<f4", 9)])
APPEARANCE_DTYPE = np.dtype([("rgba", "<f4", 4), ("scale", "<f4", 3), ("pad", "<f4")])


class Desc:
    """A `ph_model_desc` plus the numpy arrays its pointers refer to."""

    def __init__(self, arrays: dict):
        self.arrays = {}
        d = ph_model_desc()
        for n in SCALAR_INTS:
            setattr(d, n, int(arrays[n]))
        d.timestep = float(arrays["timestep"])
        d.gravity[:] = [float(x) for x in arrays["gravity"]]
        d.integrator = int(arrays["integrator"])
        d.solver_kind = int(arrays["solver_kind"])
        d.solver_iterations = int(arrays["solver_iterations"])
        for n in ["tolerance", "contact_margin", "impratio", "contact_hertz",
                  "contact_damping_ratio", "contact_slop"]:
            setattr(d, n, float(arrays[n]))
        d.solref[:] = [float(x) for x in arrays["solref"]]
        d.solimp[:] = [float(x) for x in arrays["solimp"]]
        for name, ctype, dtype, width in ARRAY_FIELDS:
            a = np.ascontiguousarray(np.asarray(arrays[name], dtype=dtype).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)  # keep a valid pointer for empty arrays
            self.arrays[name] = a
            setattr(d, name, a.ctypes.data_as(ctype))
        mesh = arrays if "nmesh" in arrays else no_meshes()
        for n in MESH_INTS:
            setattr(d, n, int(mesh[n]))
        for name, ctype, dtype, width, _cnt in MESH_FIELDS:
            a = np.ascontiguousarray(np.asarray(mesh[name], dtype=dtype).reshape(-1))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            self.arrays[name] = a
            setattr(d, name, a.ctypes.data_as(ctype))
        self.struct = d
        self.scalars = {k: arrays[k] for k in arrays if k not in self.arrays}

    def ptr(self):
        return C.byref(self.struct)


def arrays_from_desc(d: ph_model_desc) -> dict:
    """Copy a descriptor (e.g. one the C++ loader returned) into numpy arrays."""
    out = {n: int(getattr(d, n)) for n in SCALAR_INTS}
    out["timestep"] = d.timestep
    out["gravity"] = np.array(d.gravity[:])
    out["integrator"] = d.integrator
    out["solver_kind"] = d.solver_kind
    out["solver_iterations"] = d.solver_iterations
    for n in ["tolerance", "contact_margin", "impratio", "contact_hertz", "contact_damping_ratio",
              "contact_slop"]:
        out[n] = getattr(d, n)
    out["solref"] = np.array(d.solref[:])
    out["solimp"] = np.array(d.solimp[:])
    counts = {
        "body": d.nbody, "jnt": d.njnt, "dof": d.nv, "geom": d.ngeom, "cam": d.ncam,
        "actuator": d.nu, "moment": d.nmoment, "eq": d.neq, "exclude": d.nexclude, "qpos": d.nq,
    }
    for name, _ctype, dtype, width in ARRAY_FIELDS:
        prefix = name.split("_")[0]
        n = counts.get(prefix, d.nu if name.startswith("actuator") else None)
        if name == "exclude_pairs":
            n = d.nexclude
        elif name in ("qpos0", "qpos_reset"):
            n = d.nq
        elif name in ("moment_jnt", "moment_coef"):
            n = d.nmoment
        total = n * width
        p = getattr(d, name)
        out[name] = np.ctypeslib.as_array(p, shape=(total,)).copy() if total > 0 else np.zeros(0, dtype)
    for n in MESH_INTS:
        out[n] = int(getattr(d, n))
    for name, _ctype, dtype, width, cnt in MESH_FIELDS:
        total = out[cnt] * width
        p = getattr(d, name)
        out[name] = np.ctypeslib.as_array(p, shape=(total,)).copy() if total > 0 else np.zeros(0, dtype)
    return out
