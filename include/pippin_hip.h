/*
 * pippin_hip.h — C ABI of the MI355X batched FP64 physics backend.
 *
 * This is the drop-in seam for Pippin's `Physics` trait
 * (crates/pippin-env/src/lib.rs:110-155). A Rust FFI crate `impl Physics` over
 * these entry points; AsyncEnv and the renderers stay unchanged because
 * `scene()` is still built on the Rust side from the same `pippin::Model`.
 *
 * Conventions (mirroring the trait and PippinCpu, crates/pippin-env/src/pippin_cpu.rs):
 *   - plain pointers and sizes only; every buffer is caller-owned host memory
 *     unless a function says it takes device pointers;
 *   - every function returns PH_OK (0) or a negative status; the text of the
 *     last error of the calling thread is available from pippin_hip_last_error();
 *   - a buffer-size mismatch returns PH_ERR_SIZE (the trait panics via
 *     `check_len`, lib.rs:241-243; the FFI crate maps this status to a panic);
 *   - set_param validates everything before touching any environment, so a
 *     failing call leaves no partial update (pippin_cpu.rs:122-131);
 *   - `step` is synchronous: it returns after the device finished the steps.
 *   - Arrays of structures are row-major: (envs, width).
 */
#ifndef PIPPIN_HIP_H
#define PIPPIN_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PH_OK = 0,
  PH_ERR_ARG = -1,     /* invalid argument (ids, ranges, values) */
  PH_ERR_SIZE = -2,    /* buffer length does not match envs x width */
  PH_ERR_MODEL = -3,   /* model cannot be loaded or is unsupported */
  PH_ERR_DEVICE = -4,  /* HIP runtime error */
  PH_ERR_NOTPD = -5,   /* mass matrix not positive definite (forward.rs:397-399 panics) */
};

/* Joint types, in the order of crates/pippin/src/model.rs:7-12. */
enum { PH_JNT_FREE = 0, PH_JNT_BALL = 1, PH_JNT_SLIDE = 2, PH_JNT_HINGE = 3 };
/* Geom types, in the order of crates/pippin/src/model.rs:32-40 (the narrow
 * phase orders pairs by this value, collision.rs:62). */
enum { PH_GEOM_PLANE = 0, PH_GEOM_SPHERE = 1, PH_GEOM_CAPSULE = 2, PH_GEOM_BOX = 3,
       PH_GEOM_CYLINDER = 4, PH_GEOM_MESH = 5 };
enum { PH_INTEGRATOR_EULER = 0, PH_INTEGRATOR_IMPLICITFAST = 1 };
enum { PH_SOLVER_NEWTON = 0, PH_SOLVER_PGS = 1 };
/* `Field` of crates/pippin-env/src/lib.rs:22-27, then the rest of `Data`'s
 * carried state (crates/pippin/src/data.rs:8-80) for checkpoint/resume:
 * qacc (Newton's warm start, nv wide), qfrc_applied (nv) and time (1). */
enum { PH_FIELD_QPOS = 0, PH_FIELD_QVEL = 1, PH_FIELD_CTRL = 2,
       PH_FIELD_QACC = 3, PH_FIELD_QFRC_APPLIED = 4, PH_FIELD_TIME = 5 };
/* `Param` of crates/pippin/src/params.rs:10-39, same order. */
enum {
  PH_PARAM_GEOM_FRICTION = 0, PH_PARAM_GEOM_SIZE, PH_PARAM_GEOM_RGBA, PH_PARAM_GEOM_CONTYPE,
  PH_PARAM_GEOM_CONAFFINITY, PH_PARAM_BODY_MASS, PH_PARAM_DOF_DAMPING, PH_PARAM_ACTUATOR_GAIN,
  PH_PARAM_ACTUATOR_BIAS, PH_PARAM_BODY_INERTIA, PH_PARAM_BODY_IPOS, PH_PARAM_QPOS0
};

/*
 * Raw model: the fields the MJCF/URDF compiler fills before `Model::compile`
 * (crates/pippin/src/model.rs:103-249, crates/pippin/src/mjcf.rs:211-395).
 * Derived data (dof tree, dof_blocks, weld ids, bounding boxes, collision
 * pairs, invweight0) is recomputed by the library from these fields, exactly
 * as `Model::compile` and `forward::set_const` do. Vectors are flattened:
 * Vec3 -> 3 doubles, Quat -> 4 (w, x, y, z), Mat3 -> 9 row-major,
 * bools -> int32 0/1, usize::MAX -> -1.
 */
typedef struct ph_model_desc {
  int32_t nq, nv, nu, nbody, njnt, ngeom, ncam, neq, nexclude, nmoment;
  double timestep;
  double gravity[3];
  int32_t integrator;
  /* SolverOptions, model.rs:60-100 */
  int32_t solver_kind, solver_iterations;
  double tolerance, contact_margin, solref[2], solimp[5], impratio;
  double contact_hertz, contact_damping_ratio, contact_slop;

  /* bodies (body 0 is the world) */
  const int32_t* body_parent;
  const double* body_pos;      /* nbody x 3 */
  const double* body_quat;     /* nbody x 4 */
  const int32_t* body_jntadr;
  const int32_t* body_jntnum;
  const int32_t* body_dofadr;
  const int32_t* body_dofnum;
  const double* body_mass;
  const double* body_ipos;     /* nbody x 3 */
  const double* body_inertia;  /* nbody x 9 */
  /* joints */
  const int32_t* jnt_type;
  const int32_t* jnt_body;
  const int32_t* jnt_qposadr;
  const int32_t* jnt_dofadr;
  const double* jnt_pos;       /* njnt x 3 */
  const double* jnt_axis;      /* njnt x 3 */
  const int32_t* jnt_limited;
  const double* jnt_range;     /* njnt x 2 */
  const double* jnt_stiffness;
  const double* jnt_springref;
  /* dofs */
  const int32_t* dof_body;
  const int32_t* dof_jnt;
  const double* dof_damping;
  const double* dof_armature;
  /* geoms */
  const int32_t* geom_type;
  const int32_t* geom_body;
  const double* geom_pos;      /* ngeom x 3 */
  const double* geom_quat;     /* ngeom x 4 */
  const double* geom_size;     /* ngeom x 3 */
  const double* geom_friction;
  const uint32_t* geom_contype;
  const uint32_t* geom_conaffinity;
  const float* geom_rgba;      /* ngeom x 4 */
  const int32_t* geom_group;
  const int32_t* geom_dataid;  /* mesh index or -1 */
  /* cameras */
  const int32_t* cam_body;
  const double* cam_pos;       /* ncam x 3 */
  const double* cam_quat;      /* ncam x 4 */
  const double* cam_fovy;
  /* actuators: transmission = moment entries [adr, adr+num) of (joint, coef) */
  const double* actuator_gear;
  const double* actuator_gain;
  const double* actuator_bias;        /* nu x 3 */
  const int32_t* actuator_ctrllimited;
  const double* actuator_ctrlrange;   /* nu x 2 */
  const int32_t* actuator_forcelimited;
  const double* actuator_forcerange;  /* nu x 2 */
  const int32_t* actuator_moment_adr;
  const int32_t* actuator_moment_num;
  const int32_t* moment_jnt;          /* nmoment */
  const double* moment_coef;          /* nmoment */
  /* joint equality: (q1 - q1_0) = poly(q2 - q2_0) */
  const int32_t* eq_joint1;
  const int32_t* eq_joint2;           /* -1: fixed to a constant */
  const double* eq_polycoef;          /* neq x 5 */
  const double* eq_solref;            /* neq x 2 */
  const double* eq_solimp;            /* neq x 5 */
  /* <contact><exclude body1 body2> */
  const int32_t* exclude_pairs;       /* nexclude x 2 */
  const double* qpos0;                /* nq: kinematic reference */
  const double* qpos_reset;           /* nq: state Data::reset returns to */
  /* Convex hulls of the mesh assets (Model::mesh_hull, model.rs:176-182, built by
   * mesh::convex_hull, mesh.rs:63-69), indexed by geom_dataid. Per mesh: its hull
   * vertices in the geom frame (scaled), sorted by (x, y, z), and its faces, each
   * a convex polygon of hull-vertex indices (within the mesh) that runs
   * counter-clockwise seen from outside and starts at its smallest index; faces
   * sorted by their index lists. */
  int32_t nmesh, nmeshvert, nmeshface, nmeshfacevert;
  const int32_t* mesh_vertadr;        /* nmesh */
  const int32_t* mesh_vertnum;        /* nmesh */
  const int32_t* mesh_faceadr;        /* nmesh */
  const int32_t* mesh_facenum;        /* nmesh */
  const double* mesh_vert;            /* nmeshvert x 3 */
  const int32_t* face_vertadr;        /* nmeshface: first entry in face_vert */
  const int32_t* face_vertnum;        /* nmeshface */
  const int32_t* face_vert;           /* nmeshfacevert */
} ph_model_desc;

/* World pose, crates/pippin-env/src/lib.rs:102-107 (#[repr(C)]). */
typedef struct ph_pose { float pos[3]; float mat[9]; } ph_pose;
/* Per-env geom appearance, crates/pippin-env/src/lib.rs:159-165 (#[repr(C)]). */
typedef struct ph_appearance { float rgba[4]; float scale[3]; float pad; } ph_appearance;
/* `Dims`, crates/pippin-env/src/lib.rs:29-36. */
typedef struct ph_dims { int64_t nq, nv, nu, ngeom, ncam; } ph_dims;
/* One contact, crates/pippin/src/collision.rs:10-19 (geom ids, pair-local hit index). */
typedef struct ph_contact {
  double pos[3], normal[3], tangent[3], depth;
  int32_t geom1, geom2, feature, pad;
} ph_contact;

typedef struct ph_sim ph_sim;   /* opaque handle: one batch of environments on one GPU */
typedef struct ph_model ph_model; /* opaque handle: a raw model owned by the library */

/* Options for pippin_hip_create. Zero-initialise and set what you need. */
typedef struct ph_options {
  int32_t device;           /* HIP device ordinal */
  int32_t solver_override;  /* -1: keep the model's; else PH_SOLVER_* */
  int32_t max_contacts;     /* per-env contact capacity; 0 = derived from the model */
  /* 1: Newton's dense Hessian J'DJ (islands over 16 dofs or rows, at most 64
   * dofs) on the FP64 matrix cores when one env owns a wavefront (launch shapes
   * with 1 env per wavefront). No counterpart in the reference (newton.rs:279-300 sums
   * (d_r J_ri) J_rj row by row); the fused accumulation is not bit-identical
   * with it, so results then depend on the launch shape. 0 (default): off. */
  int32_t hessian_mfma;
  int32_t reserved[4];
} ph_options;

/* Text of the last error raised on the calling thread ("" if none). */
const char* pippin_hip_last_error(void);

/* ---- model loading (host; replaces pippin::mjcf::load_file, mjcf.rs:161-170) ---- */
/* Parse an MJCF document (subset of mjcf.rs). `dir` resolves relative asset paths. */
int pippin_hip_model_load_xml(const char* xml, const char* dir, ph_model** out);
int pippin_hip_model_load_file(const char* path, ph_model** out);
/* URDF import (replaces pippin::urdf::load_file, crates/pippin/src/urdf.rs:330-339,
 * with UrdfOptions :22-28): the file is rewritten as MJCF (urdf.rs:205-327) and
 * loaded as pippin_hip_model_load_xml loads it. floating_base != 0 gives the
 * root link a free joint. package_dirs: ':'-separated directories searched for
 * package://<pkg>/... URIs (NULL or "" for none); ROS_PACKAGE_PATH and the
 * folders above the URDF file are searched too. Skipped geoms and unsupported
 * joint types are reported on stderr, as the reference does. */
int pippin_hip_model_load_urdf(const char* path, int32_t floating_base, const char* package_dirs, ph_model** out);
/* The MJCF text a URDF document becomes (urdf::to_mjcf, urdf.rs:205). `base`
 * resolves relative mesh paths. Writes the document and its terminating zero
 * to out when cap allows (else PH_ERR_SIZE); *len receives its length. */
int pippin_hip_urdf_to_mjcf(const char* urdf, const char* base, int32_t floating_base, const char* package_dirs,
                            char* out, int64_t cap, int64_t* len);
/* Borrow the flat descriptor of a loaded model (valid until pippin_hip_model_free). */
const ph_model_desc* pippin_hip_model_desc(const ph_model* m);
void pippin_hip_model_free(ph_model* m);

/* ---- lifecycle: PippinCpu::new (pippin_cpu.rs:267-273) ---- */
/* Copies `desc` (the caller may free it afterwards), compiles derived data,
 * uploads it and computes invweight0 on the device. */
int pippin_hip_create(const ph_model_desc* desc, int64_t num_envs, const ph_options* opt, ph_sim** out);
void pippin_hip_destroy(ph_sim* s);
int pippin_hip_dims(const ph_sim* s, ph_dims* out);
int64_t pippin_hip_num_envs(const ph_sim* s);

/* ---- Physics::step / reset / get / set / poses (lib.rs:116-128) ---- */
/* Advance envs [begin, end) by nstep steps, ctrl held. Poses are a function of
 * qpos and are recomputed by the next pippin_hip_poses call (what PippinCpu's
 * kinematics rerun after step amounts to, pippin_cpu.rs:107-110). */
int pippin_hip_step(ph_sim* s, int64_t begin, int64_t end, int64_t nstep);
int pippin_hip_reset(ph_sim* s, const int64_t* envs, int64_t n);
/* pippin_hip_reset for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops that reset finished envs every step,
 * harnesses that forbid hipMalloc). No counterpart in the reference. It calls
 * no hipMalloc or hipFree: the env ids are staged in a fixed device array of
 * 65536 ids (512 KiB per GPU, held until the process ends); a call with more
 * ids returns PH_ERR_ARG and changes nothing. Every sim on one GPU shares that
 * array, so calls that overlap in time on one GPU (other threads, other
 * streams) race; calls from one thread are ordered. */
int pippin_hip_reset_noalloc(ph_sim* s, const int64_t* envs, int64_t n);
/* `out` / `src` hold (end - begin) x width doubles, width = dims.{nq,nv,nu} for
 * qpos/qvel/ctrl, nv for qacc and qfrc_applied, 1 for time. */
int pippin_hip_get(const ph_sim* s, int32_t field, int64_t begin, int64_t end, double* out, int64_t len);
int pippin_hip_set(ph_sim* s, int32_t field, int64_t begin, int64_t end, const double* src, int64_t len);
/* Geom poses (envs x ngeom) and camera poses (envs x ncam). */
int pippin_hip_poses(const ph_sim* s, int64_t begin, int64_t end, ph_pose* geoms, int64_t ngeoms_len,
                     ph_pose* cams, int64_t ncams_len);
/* The same poses for every env in DEVICE memory, for a renderer on the GPU
 * (the `Renderer` trait's consumer, crates/pippin-env/src/lib.rs:224-238; the
 * host copy is pippin_cpu.rs:189-205): *geoms (envs x ngeom) and *cams
 * (envs x ncam) point at buffers the sim owns until pippin_hip_destroy. Each
 * call refreshes them to the current state, equal to pippin_hip_poses bit for
 * bit, on the library stream, and returns after they are written. */
int pippin_hip_poses_device(ph_sim* s, ph_pose** geoms, ph_pose** cams);

/* ---- per-env parameters: Physics::{set,get}_param (lib.rs:130-141), params.rs:111-177 ---- */
/* `values` holds (end - begin) x width doubles (width = Param::width_in). */
int pippin_hip_set_param(ph_sim* s, int32_t param, int64_t id, int64_t begin, int64_t end,
                         const double* values, int64_t len);
int pippin_hip_get_param(const ph_sim* s, int32_t param, int64_t id, int64_t begin, int64_t end,
                         double* out, int64_t len);
int pippin_hip_appearance(const ph_sim* s, int64_t begin, int64_t end, ph_appearance* out, int64_t len);

/* ---- diagnostics for parity and measurement ---- */
/* Contacts of the last step per env (Sim.ncon, pippin-py/src/lib.rs:171-173). */
int pippin_hip_ncon(const ph_sim* s, int64_t begin, int64_t end, int32_t* out, int64_t len);
/* Contacts of env `env`; writes at most `cap`, returns the count in *n. */
int pippin_hip_contacts(const ph_sim* s, int64_t env, ph_contact* out, int32_t cap, int32_t* n);
/* Newton iterations of the last step per env (Data::solver_iterations, data.rs:53-55);
 * with the learned warm start on a PGS model, the sweeps of the last step. */
int pippin_hip_solver_iterations(const ph_sim* s, int64_t begin, int64_t end, int32_t* out, int64_t len);
/* Run forward dynamics (kinematics, CRBA, RNE; no stepping) and copy M(q)
 * (envs x nv x nv) and c(q, qdot) = qfrc_bias (envs x nv). */
int pippin_hip_mass_bias(ph_sim* s, int64_t begin, int64_t end, double* qm, int64_t qm_len, double* bias,
                         int64_t bias_len);
/* dof_invweight (envs x nv) and body_invweight (envs x nbody x 2) as set_const computes them. */
int pippin_hip_invweight(const ph_sim* s, int64_t begin, int64_t end, double* dof, int64_t dof_len,
                         double* body, int64_t body_len);
/* Number of candidate collision pairs and their (g1, g2) list (compile-time filtering). */
int pippin_hip_collision_pairs(const ph_sim* s, int64_t env, int32_t* pairs, int32_t cap, int32_t* n);

/* ---- checkpoint/resume of the warm-start state ----
 * The reference keeps no such entry point; these expose what a resumed run
 * needs beyond the PH_FIELD_* state to continue bit for bit: PGS's contact
 * history, Scratch::prev_contacts and prev_lambda (crates/pippin/src/data.rs:60-80,
 * matched by solver.rs:253-289, cleared by Data::reset, data.rs:122-131).
 * Newton's warm start is PH_FIELD_QACC. */
/* Contacts of history one env can hold (the per-env contact capacity). */
int pippin_hip_history_capacity(const ph_sim* s, int32_t* cap);
/* counts: (end - begin); contacts: (end - begin) x cap; lambda: (end - begin) x cap x 3
 * (normal, tangent 1, tangent 2 impulse of each contact). get: PH_ERR_SIZE if an env
 * holds more than cap contacts. set: validates every env first (count within the
 * capacity, geom ids in range, finite values) and changes nothing on an error. */
int pippin_hip_get_history(const ph_sim* s, int64_t begin, int64_t end, int32_t cap, int32_t* counts,
                           ph_contact* contacts, double* lambda);
int pippin_hip_set_history(ph_sim* s, int64_t begin, int64_t end, int32_t cap, const int32_t* counts,
                           const ph_contact* contacts, const double* lambda);

/* Device arithmetic check: for each x[i], out[4i..4i+3] = (sin x, cos x, sqrt|x|, 1/x)
 * as the kernels compute them (sin/cos correctly rounded). No counterpart in the
 * reference; it lets callers confirm the device math matches the host bit for bit. */
int pippin_hip_selftest_math(const double* x, int64_t n, double* out, int32_t device);

/* ---- device-resident rollout (Batch::rollout, crates/pippin/src/batch.rs:46-59) ---- */
/* All pointers are DEVICE pointers: ctrl (envs x nstep x nu), qpos_out (envs x nstep x nq)
 * for envs [begin, end). Runs on the library's stream and synchronises. */
int pippin_hip_rollout_device(ph_sim* s, int64_t begin, int64_t end, int64_t nstep, const double* ctrl,
                              double* qpos_out);
/* Device pointers of the resident state (envs x width), for zero-copy callers.
 * Once a qpos pointer has been handed out, pippin_hip_poses recomputes poses on
 * every call, since writes through the pointer are not seen by the library. */
int pippin_hip_state_device(ph_sim* s, int32_t field, double** ptr);
/* Copy `field` (PH_FIELD_QVEL or PH_FIELD_CTRL) for envs [begin, end) from a DEVICE
 * buffer, ordered on the library stream before the next step (no host sync);
 * pippin_hip_get and pippin_hip_set wait for it. */
int pippin_hip_set_device(ph_sim* s, int32_t field, int64_t begin, int64_t end, const double* src);
/* Accumulated device time of the step kernel (hipEvents around each launch on the
 * library stream) and the number of launches since the last reset. */
int pippin_hip_kernel_timing(ph_sim* s, int32_t reset, double* total_ms, int64_t* launches);
/* Stream the library launches on (a hipStream_t). */
void* pippin_hip_stream(ph_sim* s);

/* ---- launch shapes, autotuning and graph replay of the step ----
 * No counterpart in the reference: its CPU backend chunks envs over threads
 * (crates/pippin/src/batch.rs:70-72) and its Metal backend dispatches one thread
 * per env (crates/pippin-metal/src/lib.rs:174-175). Every shape computes every
 * env with the same arithmetic, so results are identical under all of them. */
enum {
  PH_SCHED_GRID = 0,      /* one workgroup per K envs, dispatched by the hardware */
  PH_SCHED_WS = 1,        /* persistent workgroups taking K envs at a time from a counter */
  PH_SCHED_WS_BUCKET = 2, /* as WS, envs ordered by Newton iterations of their last launch */
  PH_SCHED_PGS_SPLIT = 3  /* PGS models, K = 1, generic kernel: each physics step as three launches
                             (up to the sweeps, the sweeps with one env per wavefront and no LDS,
                             the rest), so that many sweep wavefronts share a SIMD */
};
typedef struct ph_launch_shape {
  int32_t envs_per_wave; /* K: 1, 2 or 4 envs per 64-lane wavefront */
  int32_t specialised;   /* 1: the kernel compiled for this model's sizes (when one matches) */
  int32_t schedule;      /* PH_SCHED_* */
  int32_t pad;
} ph_launch_shape;
int pippin_hip_get_launch_shape(const ph_sim* s, ph_launch_shape* out);
/* PH_ERR_ARG for a shape the model or device cannot take (LDS, no specialised kernel). */
int pippin_hip_set_launch_shape(ph_sim* s, const ph_launch_shape* shape);
/* Online autotuning: the following pippin_hip_step and pippin_hip_rollout_device
 * calls (without the learned warm start) cycle through every valid shape for
 * 1 + `rounds` rounds, timing each
 * launch per env-step (the first round is not timed), then keep the shape with
 * the lowest median. With `cache_path` (or NULL), a shape stored there for this
 * model's sizes, env count and device is taken at once, and a tuned one is
 * appended as a line "key : K specialised schedule". */
int pippin_hip_autotune(ph_sim* s, int32_t rounds, const char* cache_path);
/* 1 while autotuning trials remain, else 0. */
int pippin_hip_autotune_pending(const ph_sim* s);
/* Graph replay (default on): each plain pippin_hip_step launch sequence is
 * captured in a hipGraph on first use and replayed while its range, step count
 * and shape repeat (the most recent 8 are kept). */
int pippin_hip_use_graphs(ph_sim* s, int32_t on);

/* ---- learned warm start ----
 * No counterpart in the reference. A 24 -> H -> H -> H -> 4 GELU MLP (BF16
 * weights, FP32 accumulation on the matrix cores) reads features of each of an
 * env's first 16 contacts.
 * Newton (the reference warm-starts from the previous acceleration when its
 * cost is lower than that of a0, crates/pippin/src/newton.rs:238-245): the
 * network predicts the contact's four pyramid-row forces, and a0 + M^-1 J^T
 * lambda becomes a third candidate, used only if its Newton cost is lower than
 * the reference's choice.
 * PGS (the reference warm-starts each contact from the nearest previous
 * contact of the same geom pair and leaves the others at zero,
 * crates/pippin/src/solver.rs:253-289): the network predicts the normal and
 * friction forces, which become the warm start of the contacts the matching
 * leaves at zero; in MLP mode solver_iterations reports the step's sweeps. */
#define PH_MLP_LOG 1  /* record features and solved forces of every step */
#define PH_MLP_WARM 2 /* use the predicted candidate (hidden width <= 64) */
/* Allocate the network (hidden width H, a multiple of 16 up to 256; weights
 * U(+-1/sqrt(fan_in)) from `seed`), the ring of recorded contacts (`capacity`
 * samples, at least envs x 16) and the training scratch (batches up to max_batch,
 * a multiple of 16). Mode starts at 0 (off). */
int pippin_hip_mlp_init(ph_sim* s, int32_t hidden, uint64_t seed, int64_t capacity, int64_t max_batch);
/* PH_MLP_LOG | PH_MLP_WARM, or 0: steps run the reference's path unchanged. */
int pippin_hip_mlp_mode(ph_sim* s, int32_t mode);
/* Append the contacts recorded by the last logged launch to the ring (env order);
 * *stored = samples held (at most capacity; oldest overwritten). */
int pippin_hip_mlp_collect(ph_sim* s, int64_t* stored);
/* `iters` Adam steps (lr, betas 0.9 / 0.999, eps 1e-8) on batches of `batch` samples
 * drawn from the ring by a hash of (seed, step); loss_out[iters] (optional) gets each
 * batch's mean squared error on log1p(lambda / D). */
int pippin_hip_mlp_train(ph_sim* s, int32_t iters, int64_t batch, float lr, uint64_t seed, float* loss_out);
/* Parameter count, -1 before mlp_init. Layout: W1 [32][H] b1 [H] W2 [H][H] b2 W3 [H][H] b3
 * W4 [H][16] b4 [16] (FP32, row-major [in][out]; input rows 24..31, output columns 4..15
 * and b4[4..15] are padding and must stay 0). */
int64_t pippin_hip_mlp_num_params(const ph_sim* s);
int pippin_hip_mlp_get_params(ph_sim* s, float* out, int64_t n);
/* Replace the parameters (restarts Adam). */
int pippin_hip_mlp_set_params(ph_sim* s, const float* in, int64_t n);
/* Host arrays: x (n x 24 features) -> y (n x 4 outputs); n a multiple of 16. */
int pippin_hip_mlp_forward(ph_sim* s, const float* x, int64_t n, float* y);
/* Mean squared error of the outputs against t (n x 4) and its gradient w.r.t. the
 * parameters (num_params floats); no update. */
int pippin_hip_mlp_loss_grad(ph_sim* s, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* Records of env `env` from the last step: features (16 x 24), predictions and
 * targets (16 x 4 each; Newton: log1p(lambda / D) of the four pyramid rows; PGS:
 * sign(x) log1p(|x|) of 100 lambda / inv_a for the normal and two friction rows,
 * output 3 zero) and the number of contacts recorded. */
int pippin_hip_mlp_records(ph_sim* s, int64_t env, float* feat, float* pred, float* tgt, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* PIPPIN_HIP_H */
