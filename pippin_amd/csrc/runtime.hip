// Device runtime and C ABI of the batched FP64 physics backend (include/pippin_hip.h).
//
// One environment per 64-lane wavefront (one workgroup); per-env working set in
// LDS; per-env state, contacts and large row buffers in HBM. The C entry points
// follow the `Physics` trait (crates/pippin-env/src/lib.rs:110-155) and the CPU
// backend's conventions (crates/pippin-env/src/pippin_cpu.rs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pippin_hip.h"
#include "host_model.hpp"
#include "kstep.hpp"
#include "mlp_host.hpp"
#include "spec_models.hpp"

using namespace phk;

// ---------------------------------------------------------------- kernels
static const void* step_kernel(int K, int spec) {
  if (spec == 1 && K == 4) return (const void*)k_step<4, SpecAnt>;
  return K == 4 ? (const void*)k_step<4, DModel> : K == 2 ? (const void*)k_step<2, DModel> : (const void*)k_step<1, DModel>;
}
static const void* ws_kernel(int K, int spec) {
  if (spec == 1 && K == 4) return (const void*)k_step_ws<4, SpecAnt>;
  return K == 4 ? (const void*)k_step_ws<4, DModel>
                : K == 2 ? (const void*)k_step_ws<2, DModel> : (const void*)k_step_ws<1, DModel>;
}

// Kinematics only (poses after set(qpos) / reset), for listed envs or a range.
__global__ void __launch_bounds__(64) k_kin(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                            int64_t begin, const int64_t* ids) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  extern __shared__ double lds[];
  Env E;
  E.s = lds;
  E.si = reinterpret_cast<int*>(lds + M.L.ndouble);
  E.e = ids ? ids[blockIdx.x] : begin + blockIdx.x;
  E.var = M.env_var[E.e];
  load_state<DevX>(M, S, E);
  kinematics<DevX>(M, E);
  store_poses<DevX>(M, S, E);
}

// Poses as f32 `Pose` records (pippin-env/src/lib.rs:102-107) in device memory:
// geoms from the double poses k_kin wrote, cameras from their body's pose
// (pippin_cpu.rs:441-446), the same arithmetic as the host path of
// pippin_hip_poses, then rounded to f32.
__global__ void __launch_bounds__(256) k_pose_f32(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                  const int* cam_body, const Real* cam_pos, const Real* cam_quat,
                                                  ph_pose* geoms, ph_pose* cams) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ng = M.ngeom, nc = M.ncam, nb = M.nbody;
  if (t < S.n * ng) {
    ph_pose& p = geoms[t];
    for (int k = 0; k < 3; k++) p.pos[k] = (float)S.geom_xpos[t * 3 + k];
    for (int k = 0; k < 9; k++) p.mat[k] = (float)S.geom_xmat[t * 9 + k];
  } else if (t < S.n * ng + S.n * nc) {
    const int64_t u = t - S.n * ng, e = u / nc;
    const int c = (int)(u % nc), b = cam_body[c];
    const M3 xm = ldm(S.xmat + (e * nb + b) * 9);
    const V3 pos = ld3(S.xpos + (e * nb + b) * 3) + mul_vec(xm, ld3(cam_pos + 3 * c));
    const M3 mat = mul_mat(xm, to_mat(ldq(cam_quat + 4 * c)));
    ph_pose& p = cams[u];
    p.pos[0] = (float)pos.x;
    p.pos[1] = (float)pos.y;
    p.pos[2] = (float)pos.z;
    for (int k = 0; k < 9; k++) p.mat[k] = (float)mat.m[k];
  }
}

// Data::reset (data.rs:122-131) for listed envs: qacc is kept.
__global__ void __launch_bounds__(64) k_reset(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                              const int64_t* ids) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  const int64_t e = ids[blockIdx.x];
  const int var = M.env_var[e];
  for (int k = threadIdx.x; k < M.nq; k += 64) S.qpos[e * M.nq + k] = M.v_qpos_reset[(int64_t)var * M.nq + k];
  for (int k = threadIdx.x; k < M.nv; k += 64) {
    S.qvel[e * M.nv + k] = 0.0;
    S.applied[e * M.nv + k] = 0.0;
  }
  for (int k = threadIdx.x; k < M.nu; k += 64) S.ctrl[e * M.nu + k] = 0.0;
  if (threadIdx.x == 0) {
    S.ncon[e] = 0;
    S.npcon[e] = 0;
    S.time[e] = 0.0;
    S.status[e] = 0;  // a frozen env (not positive definite, capacity) steps again from the reset state
  }
}

// Staging of env ids for pippin_hip_reset_noalloc: a file-scope device array,
// so that entry point allocates nothing. Each GPU holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_on_device).
static constexpr int64_t RESET_SCRATCH_MAX = 65536;  // env ids per call (512 KiB)
static __device__ __align__(256) int64_t scratch[RESET_SCRATCH_MAX];

// Forward dynamics up to RNE; copies M(q) and c(q, qdot) out.
__global__ void __launch_bounds__(64) k_mass_bias(const DModel* __restrict__ Mp, const DState* __restrict__ Sp,
                                                  int64_t begin, Real* qm, Real* bias) {
  const DModel& M = *Mp;
  const DState& S = *Sp;
  extern __shared__ double lds[];
  Env E;
  E.s = lds;
  E.si = reinterpret_cast<int*>(lds + M.L.ndouble);
  E.e = begin + blockIdx.x;
  E.var = M.env_var[E.e];
  load_state<DevX>(M, S, E);
  kinematics<DevX>(M, E);
  crba<DevX>(M, E);
  velocity_rne<DevX>(M, E);
  const int64_t rel = blockIdx.x;
  DevX::par(M.nv * M.nv, [&](int k) { qm[rel * M.nv * M.nv + k] = lds[M.L.qm + k]; });
  DevX::par(M.nv, [&](int k) { bias[rel * M.nv + k] = lds[M.L.bias + k]; });
}

// set_const for listed variants.
__global__ void __launch_bounds__(64) k_setconst(const DModel* __restrict__ Mp, const int* vars, Real* dof_iw,
                                                 Real* body_iw) {
  const DModel& M = *Mp;
  extern __shared__ double lds[];
  Env E;
  E.s = lds;
  E.si = reinterpret_cast<int*>(lds + M.L.ndouble);
  E.e = 0;
  E.var = vars[blockIdx.x];
  DevX::par(M.nq, [&](int k) { lds[M.L.qpos + k] = M.v_qpos_reset[(int64_t)E.var * M.nq + k]; });
  set_const_env<DevX>(M, E, dof_iw + (int64_t)E.var * M.nv, body_iw + (int64_t)E.var * M.nbody * 2);
}

// Device math check: correctly rounded sin/cos, sqrt and reciprocal per input.
__global__ void __launch_bounds__(64) k_selftest_math(const Real* x, int64_t n, Real* out) {
  int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (i < n) {
    Real s, c;
    cr_sincos(x[i], &s, &c);
    out[4 * i] = s;
    out[4 * i + 1] = c;
    out[4 * i + 2] = sqrt_cr(fabs(x[i]));  // the kernels' square root
    // the kernels' reciprocal of a root over its domain, division elsewhere
    const Real a = x[i];
    out[4 * i + 3] = (a >= 0x1p-537 && a <= 0x1p512) ? recip_of_root(a) : 1.0 / a;
  }
}

// ---------------------------------------------------------------- host runtime
static thread_local std::string g_err;
void ph_set_error(const std::string& msg) { g_err = msg; }
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
static hipError_t scratch_on_device(int device, int64_t** out) {
  static std::mutex mu;
  static std::map<int, int64_t*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  addr[device] = static_cast<int64_t*>(p);
  *out = static_cast<int64_t*>(p);
  return hipSuccess;
}
#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(PH_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    if (count == 0) count = 1;
    return hipMalloc(&p, count * sizeof(T));
  }
  hipError_t upload(const std::vector<T>& v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  hipError_t zero() { return hipMemset(p, 0, (n ? n : 1) * sizeof(T)); }
};

struct Variants {
  // host mirrors, [nvar][...]
  std::vector<double> geom_size, geom_aabb, friction, body_mass, body_ipos, body_inertia, dof_damping, act_gain,
      act_bias, dof_invweight, body_invweight, qpos_reset;
  std::vector<uint32_t> contype, conaffinity;
  std::vector<float> rgba;
  std::vector<int> damped;
  int nvar = 0;
};

// A captured launch sequence of pippin_hip_step, replayed while its
// arguments (range, step count, launch shape) repeat.
struct StepGraph {
  int64_t begin = 0, end = 0, nstep = 0;
  ph_launch_shape shape{};
  hipGraphExec_t exec = nullptr;
};

struct ph_sim {
  phh::HostModel hm;
  int64_t n = 0;
  int device = 0;
  int envs_per_wg = 1;  // environments packed into one k_step workgroup
  int spec = 0;         // 1: the model matches SpecAnt (compile-time sizes and layout)
  int sched = PH_SCHED_GRID;  // schedule of the plain step (PH_SCHED_*)
  int cus = 0;                // compute units of the device
  int ws_per_cu[3][2] = {};   // resident work-stealing workgroups per CU, by K = 1, 2, 4 and kernel instance
  std::string arch;           // device architecture name (autotune cache key)
  // work-stealing schedule: env order (offsets from the range start) and slot counter
  DBuf<int> d_order;
  DBuf<uint32_t> d_counter;
  // online autotuning: candidate shapes tried in turn on live steps
  std::vector<ph_launch_shape> tune_cand;
  std::vector<std::vector<double>> tune_ms;  // ms per env-step, per candidate
  int tune_next = 0, tune_round = 0, tune_rounds = 0;
  std::string tune_cache;
  // graph replay of pippin_hip_step launches
  bool use_graphs = true;
  std::vector<StepGraph> graphs;
  // Poses are a function of qpos (PippinCpu reruns kinematics after step and
  // set(Qpos), pippin_cpu.rs:107-110, 183-185); they are recomputed for all
  // envs on the first poses() call after qpos changed.
  mutable bool poses_stale = true;
  int64_t* reset_scratch = nullptr;  // this device's `scratch` (pippin_hip_reset_noalloc)
  // f32 poses in device memory (pippin_hip_poses_device), allocated on first use
  DBuf<ph_pose> d_pose_geom, d_pose_cam;
  // the split PGS step's velocities between its kernels (n x nv each), for PGS models
  DBuf<double> d_pgs_vel, d_pgs_vpos;
  DBuf<int> d_cam_body;
  DBuf<double> d_cam_pos, d_cam_quat;
  bool qpos_shared = false;  // a device qpos pointer was handed out (pippin_hip_state_device)
  hipStream_t stream = nullptr;
  DModel M;
  DState S;
  Variants V;
  std::vector<int> env_var;  // host mirror
  std::vector<int> var_refs;  // envs per variant
  // device buffers
  DBuf<int> d_i[56];
  DBuf<double> d_f[32];
  DBuf<uint64_t> d_u64[2];
  DBuf<uint32_t> d_u32[2];
  DBuf<int> d_env_var;
  // variant buffers
  DBuf<double> dv_geom_size, dv_geom_aabb, dv_friction, dv_body_mass, dv_body_ipos, dv_body_inertia, dv_dof_damping,
      dv_act_gain, dv_act_bias, dv_dof_invweight, dv_body_invweight, dv_qpos_reset;
  DBuf<uint32_t> dv_contype, dv_conaffinity;
  DBuf<int> dv_damped;
  // state buffers
  DBuf<double> s_qpos, s_qvel, s_ctrl, s_qacc, s_applied, s_time, s_gxpos, s_gxmat, s_xpos, s_xmat, s_con, s_pcon,
      s_plambda, s_stage, s_gjac, s_grow, s_islf, s_gmj;
  DBuf<int> s_ncon, s_iters, s_nrows, s_status, s_cgeom, s_pcgeom, s_npcon, s_cost;
  DBuf<int64_t> d_ids;
  DBuf<DModel> d_model;  // device copies of M and S, read through the scalar cache
  DBuf<DState> d_state;
  // k_step timing on the library stream (hipEvents around every launch)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_ms = 0.0;
  int64_t kernel_launches = 0;
  // learned warm start (mlp.hpp): network, optimiser, batch scratch, ring of
  // recorded contacts, per-env records of the last step
  bool mlp_ready = false;
  phm::Net net;
  phm::Batch batch;
  DBuf<float> m_P, m_m, m_v, m_G, m_z, m_y, m_loss, m_gp, m_rfeat, m_rtgt, m_feat, m_pred, m_tgt, m_xin, m_tin;
  DBuf<uint16_t> m_wbf, m_wtbf, m_x0, m_a, m_dz;
  DBuf<int> m_cnt;
  DBuf<int64_t> m_total;
  int64_t ring_cap = 0, ring_head = 0, ring_size = 0, adam_t = 0;
  int64_t log_begin = 0, log_end = 0;  // envs recorded by the last launch with logging on
};

extern "C" const char* pippin_hip_last_error(void) { return g_err.c_str(); }

template <class T>
static hipError_t up(DBuf<T>& b, const std::vector<T>& v, gp<const T>* out) {
  hipError_t e = b.upload(v);
  *out = b.p;
  return e;
}

// Refresh the device copies of the model and state descriptors.
static int sync_desc(ph_sim* s) {
  if (!s->d_model.p) HIPCHK(s->d_model.alloc(1));
  if (!s->d_state.p) HIPCHK(s->d_state.alloc(1));
  HIPCHK(hipMemcpy(s->d_model.p, &s->M, sizeof(DModel), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->d_state.p, &s->S, sizeof(DState), hipMemcpyHostToDevice));
  return PH_OK;
}

static int upload_variants(ph_sim* s) {
  Variants& V = s->V;
  DModel& M = s->M;
  HIPCHK(up(s->dv_geom_size, V.geom_size, &M.v_geom_size));
  HIPCHK(up(s->dv_geom_aabb, V.geom_aabb, &M.v_geom_aabb));
  HIPCHK(up(s->dv_friction, V.friction, &M.v_friction));
  HIPCHK(up(s->dv_contype, V.contype, &M.v_contype));
  HIPCHK(up(s->dv_conaffinity, V.conaffinity, &M.v_conaffinity));
  HIPCHK(up(s->dv_body_mass, V.body_mass, &M.v_body_mass));
  HIPCHK(up(s->dv_body_ipos, V.body_ipos, &M.v_body_ipos));
  HIPCHK(up(s->dv_body_inertia, V.body_inertia, &M.v_body_inertia));
  HIPCHK(up(s->dv_dof_damping, V.dof_damping, &M.v_dof_damping));
  HIPCHK(up(s->dv_act_gain, V.act_gain, &M.v_act_gain));
  HIPCHK(up(s->dv_act_bias, V.act_bias, &M.v_act_bias));
  HIPCHK(up(s->dv_dof_invweight, V.dof_invweight, &M.v_dof_invweight));
  HIPCHK(up(s->dv_body_invweight, V.body_invweight, &M.v_body_invweight));
  HIPCHK(up(s->dv_qpos_reset, V.qpos_reset, &M.v_qpos_reset));
  HIPCHK(up(s->dv_damped, V.damped, &M.v_damped));
  return sync_desc(s);
}

// Recompute invweight0 on the device for the listed variants and pull it back.
static int run_setconst(ph_sim* s, const std::vector<int>& vars) {
  if (vars.empty()) return PH_OK;
  DBuf<int> dv;
  HIPCHK(dv.upload(vars));
  hipLaunchKernelGGL(k_setconst, dim3((unsigned)vars.size()), dim3(64), s->M.L.bytes, s->stream, s->d_model.p, dv.p,
                     s->dv_dof_invweight.p, s->dv_body_invweight.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  Variants& V = s->V;
  HIPCHK(hipMemcpy(V.dof_invweight.data(), s->dv_dof_invweight.p, V.dof_invweight.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(V.body_invweight.data(), s->dv_body_invweight.p, V.body_invweight.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
  return PH_OK;
}

static void variant_init(ph_sim* s) {
  const phh::HostModel& hm = s->hm;
  Variants& V = s->V;
  V.nvar = 1;
  V.geom_size = hm.geom_size;
  V.geom_aabb = hm.geom_aabb;
  V.friction = hm.geom_friction;
  V.contype = hm.geom_contype;
  V.conaffinity = hm.geom_conaffinity;
  V.rgba = hm.geom_rgba;
  V.body_mass = hm.body_mass;
  V.body_ipos = hm.body_ipos;
  V.body_inertia = hm.body_inertia;
  V.dof_damping = hm.dof_damping;
  V.act_gain = hm.actuator_gain;
  V.act_bias = hm.actuator_bias;
  V.dof_invweight.assign(hm.d.nv, 0.0);
  V.body_invweight.assign(2 * hm.d.nbody, 0.0);
  V.qpos_reset = hm.qpos_reset;
  V.damped.assign(1, hm.damped ? 1 : 0);
}

// Copy variant `src` to a new variant; returns its index.
static int variant_clone(ph_sim* s, int src) {
  Variants& V = s->V;
  auto dup = [&](auto& vec, size_t w) {
    vec.insert(vec.end(), vec.begin() + src * w, vec.begin() + (src + 1) * w);
  };
  const ph_model_desc& d = s->hm.d;
  dup(V.geom_size, 3 * d.ngeom);
  dup(V.geom_aabb, 6 * d.ngeom);
  dup(V.friction, d.ngeom);
  dup(V.contype, d.ngeom);
  dup(V.conaffinity, d.ngeom);
  dup(V.rgba, 4 * d.ngeom);
  dup(V.body_mass, d.nbody);
  dup(V.body_ipos, 3 * d.nbody);
  dup(V.body_inertia, 9 * d.nbody);
  dup(V.dof_damping, d.nv);
  dup(V.act_gain, d.nu);
  dup(V.act_bias, 3 * d.nu);
  dup(V.dof_invweight, d.nv);
  dup(V.body_invweight, 2 * d.nbody);
  dup(V.qpos_reset, d.nq);
  dup(V.damped, 1);
  return V.nvar++;
}

extern "C" int pippin_hip_create(const ph_model_desc* desc, int64_t num_envs, const ph_options* opt, ph_sim** out) {
  if (!desc || !out || num_envs < 0) return fail(PH_ERR_ARG, "null argument or negative env count");
  std::unique_ptr<ph_sim> s(new ph_sim());
  std::string err;
  if (!phh::build_host_model(*desc, s->hm, &err)) return fail(PH_ERR_MODEL, err);
  phh::HostModel& hm = s->hm;
  const ph_model_desc& d = hm.d;
  int solver = d.solver_kind;
  if (opt && opt->solver_override >= 0) solver = opt->solver_override;
  if (hm.block_adr.size() > 33)  // island masks are 32-bit
    return fail(PH_ERR_MODEL, "models with more than 32 dof blocks are not supported on the device");
  s->n = num_envs;
  s->device = opt ? opt->device : 0;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&s->ev0));
  HIPCHK(hipEventCreate(&s->ev1));

  DModel& M = s->M;
  phh::init_dmodel_scalars(hm, solver, opt ? opt->max_contacts : 0, M);
  M.hess_mfma = opt && opt->hessian_mfma ? 1 : 0;
  if (M.L.bytes > 64 * 1024) return fail(PH_ERR_MODEL, "model too large for the per-env LDS working set");
  // Pack as many envs per workgroup as fit 64 KiB of LDS (PIPPIN_HIP_ENVS_PER_WG overrides: 1, 2 or 4,
  // up to the device's per-workgroup LDS).
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, s->device));
  const int64_t lds_max = (int64_t)prop.sharedMemPerBlock;
  s->envs_per_wg = 1;
  for (int k : {4, 2})
    if (k * M.L.bytes <= 64 * 1024) {
      s->envs_per_wg = k;
      break;
    }
  if (const char* ev = std::getenv("PIPPIN_HIP_ENVS_PER_WG")) {
    int k = std::atoi(ev);
    if ((k == 1 || k == 2 || k == 4) && (int64_t)k * M.L.bytes <= lds_max) s->envs_per_wg = k;
  }
  // model-specialised kernel when the model's sizes and layout match one built
  // in (PIPPIN_HIP_SPEC=0 forces the generic kernel)
  s->spec = SpecAnt::matches(M) ? 1 : 0;
  if (const char* ev = std::getenv("PIPPIN_HIP_SPEC"))
    if (std::atoi(ev) == 0) s->spec = 0;
  for (int k : {1, 2, 4}) {
    const int bytes = k * M.L.bytes;
    if (bytes > 64 * 1024 && bytes <= lds_max) {
      HIPCHK(hipFuncSetAttribute(step_kernel(k, 0), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
      HIPCHK(hipFuncSetAttribute(ws_kernel(k, 0), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    }
  }
  HIPCHK(scratch_on_device(s->device, &s->reset_scratch));
  s->cus = prop.multiProcessorCount;
  s->arch = prop.gcnArchName;
  for (int ki = 0; ki < 3; ki++)
    for (int sp = 0; sp < 2; sp++) {
      const int k = ki == 2 ? 4 : ki + 1;
      if ((sp == 1 && (k != 4 || !SpecAnt::matches(M))) || (int64_t)k * M.L.bytes > lds_max) continue;
      int per_cu = 0;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ws_kernel(k, sp), 64, (size_t)k * M.L.bytes));
      s->ws_per_cu[ki][sp] = per_cu;
    }
  HIPCHK(s->d_order.alloc(std::max<int64_t>(num_envs, 1)));
  if (M.solver_kind == 1) {
    HIPCHK(s->d_pgs_vel.alloc((size_t)std::max<int64_t>(num_envs * M.nv, 1)));
    HIPCHK(s->d_pgs_vpos.alloc((size_t)std::max<int64_t>(num_envs * M.nv, 1)));
  }
  HIPCHK(s->d_counter.alloc(1));
  HIPCHK(s->d_counter.zero());

  int ii = 0, fi = 0;
  constexpr int NI = sizeof(s->d_i) / sizeof(s->d_i[0]), NF = sizeof(s->d_f) / sizeof(s->d_f[0]);
  auto ui = [&](const std::vector<int32_t>& v, gp<const int>* dst) -> hipError_t {
    return ii < NI ? up(s->d_i[ii++], v, dst) : hipErrorInvalidValue;
  };
  auto uf = [&](const std::vector<double>& v, gp<const Real>* dst) -> hipError_t {
    return fi < NF ? up(s->d_f[fi++], v, dst) : hipErrorInvalidValue;
  };
  HIPCHK(ui(hm.body_parent, &M.body_parent));
  HIPCHK(ui(hm.body_jntadr, &M.body_jntadr));
  HIPCHK(ui(hm.body_jntnum, &M.body_jntnum));
  HIPCHK(ui(hm.body_dofadr, &M.body_dofadr));
  HIPCHK(ui(hm.body_dofnum, &M.body_dofnum));
  HIPCHK(ui(hm.body_lastdof, &M.body_lastdof));
  HIPCHK(ui(hm.level_adr, &M.level_adr));
  HIPCHK(ui(hm.level_body, &M.level_body));
  HIPCHK(ui(hm.acc_level_adr, &M.acc_level_adr));
  HIPCHK(ui(hm.acc_parent, &M.acc_parent));
  HIPCHK(ui(hm.acc_child_adr, &M.acc_child_adr));
  HIPCHK(ui(hm.acc_child, &M.acc_child));
  HIPCHK(ui(hm.jnt_type, &M.jnt_type));
  HIPCHK(ui(hm.jnt_qposadr, &M.jnt_qposadr));
  HIPCHK(ui(hm.jnt_dofadr, &M.jnt_dofadr));
  HIPCHK(ui(hm.jnt_limited, &M.jnt_limited));
  HIPCHK(ui(hm.dof_body, &M.dof_body));
  HIPCHK(ui(hm.dof_parent, &M.dof_parent));
  HIPCHK(ui(hm.dof_jnt, &M.dof_jnt));
  HIPCHK(ui(hm.block_adr, &M.block_adr));
  HIPCHK(ui(hm.dof_block, &M.dof_block));
  HIPCHK(ui(hm.geom_type, &M.geom_type));
  HIPCHK(ui(hm.geom_body, &M.geom_body));
  HIPCHK(ui(hm.pair_g1, &M.pair_g1));
  HIPCHK(ui(hm.pair_g2, &M.pair_g2));
  HIPCHK(ui(hm.actuator_ctrllimited, &M.act_ctrllimited));
  HIPCHK(ui(hm.actuator_forcelimited, &M.act_forcelimited));
  HIPCHK(ui(hm.actuator_moment_adr, &M.act_madr));
  HIPCHK(ui(hm.actuator_moment_num, &M.act_mnum));
  HIPCHK(ui(hm.moment_jnt, &M.mom_jnt));
  HIPCHK(ui(hm.dofact_adr, &M.dofact_adr));
  HIPCHK(ui(hm.dofact_num, &M.dofact_num));
  HIPCHK(ui(hm.dofact_u, &M.dofact_u));
  HIPCHK(ui(hm.eq_joint1, &M.eq_j1));
  HIPCHK(ui(hm.eq_joint2, &M.eq_j2));
  HIPCHK(uf(hm.body_pos, &M.body_pos));
  HIPCHK(uf(hm.body_quat, &M.body_quat));
  HIPCHK(uf(hm.jnt_pos, &M.jnt_pos));
  HIPCHK(uf(hm.jnt_axis, &M.jnt_axis));
  HIPCHK(uf(hm.jnt_range, &M.jnt_range));
  HIPCHK(uf(hm.jnt_stiffness, &M.jnt_stiffness));
  HIPCHK(uf(hm.jnt_springref, &M.jnt_springref));
  HIPCHK(uf(hm.dof_armature, &M.dof_armature));
  HIPCHK(uf(hm.geom_pos, &M.geom_pos));
  HIPCHK(uf(hm.geom_quat, &M.geom_quat));
  HIPCHK(uf(hm.actuator_gear, &M.act_gear));
  HIPCHK(uf(hm.actuator_ctrlrange, &M.act_ctrlrange));
  HIPCHK(uf(hm.actuator_forcerange, &M.act_forcerange));
  HIPCHK(uf(hm.moment_coef, &M.mom_coef));
  HIPCHK(uf(hm.dofact_coef, &M.dofact_coef));
  HIPCHK(uf(hm.eq_polycoef, &M.eq_poly));
  HIPCHK(uf(hm.eq_solref, &M.eq_solref));
  HIPCHK(uf(hm.eq_solimp, &M.eq_solimp));
  HIPCHK(uf(hm.qpos0, &M.qpos0));
  HIPCHK(ui(hm.geom_dataid, &M.geom_dataid));
  HIPCHK(ui(hm.mesh_vertadr, &M.mesh_vertadr));
  HIPCHK(ui(hm.mesh_vertnum, &M.mesh_vertnum));
  HIPCHK(ui(hm.face_vertadr, &M.face_vertadr));
  HIPCHK(ui(hm.face_vertnum, &M.face_vertnum));
  HIPCHK(ui(hm.face_vert, &M.face_vert));
  HIPCHK(ui(hm.vert_faceadr, &M.vert_faceadr));
  HIPCHK(ui(hm.vert_facenum, &M.vert_facenum));
  HIPCHK(ui(hm.vert_face, &M.vert_face));
  HIPCHK(uf(hm.mesh_vert, &M.mesh_vert));
  HIPCHK(uf(hm.face_normal, &M.face_normal));
  HIPCHK(up(s->d_u64[0], hm.body_chain, &M.body_chain));
  HIPCHK(up(s->d_u64[1], hm.dof_anc, &M.dof_anc));

  variant_init(s.get());
  s->env_var.assign(num_envs, 0);
  s->var_refs.assign(1, (int)num_envs);
  HIPCHK(up(s->d_env_var, s->env_var, &M.env_var));
  if (int rc = upload_variants(s.get())) return rc;

  // per-env state
  const int64_t n = std::max<int64_t>(num_envs, 1);
  DState& S = s->S;
  S.n = num_envs;
  std::vector<double> q0(n * d.nq);
  for (int64_t e = 0; e < n; e++)
    for (int k = 0; k < d.nq; k++) q0[e * d.nq + k] = hm.qpos_reset[k];
  HIPCHK(s->s_qpos.upload(q0));
  HIPCHK(s->s_qvel.alloc(n * d.nv));
  HIPCHK(s->s_qvel.zero());
  HIPCHK(s->s_ctrl.alloc(n * d.nu));
  HIPCHK(s->s_ctrl.zero());
  HIPCHK(s->s_qacc.alloc(n * d.nv));
  HIPCHK(s->s_qacc.zero());
  HIPCHK(s->s_applied.alloc(n * d.nv));
  HIPCHK(s->s_applied.zero());
  HIPCHK(s->s_time.alloc(n));
  HIPCHK(s->s_time.zero());
  HIPCHK(s->s_gxpos.alloc(n * 3 * d.ngeom));
  HIPCHK(s->s_gxmat.alloc(n * 9 * d.ngeom));
  HIPCHK(s->s_xpos.alloc(n * 3 * d.nbody));
  HIPCHK(s->s_xmat.alloc(n * 9 * d.nbody));
  HIPCHK(s->s_con.alloc(n * M.ncmax * CW));
  HIPCHK(s->s_cgeom.alloc(n * M.ncmax * 3));
  HIPCHK(s->s_pcon.alloc(n * M.ncmax * CW));
  HIPCHK(s->s_pcgeom.alloc(n * M.ncmax * 3));
  HIPCHK(s->s_plambda.alloc(n * M.ncmax * 3));
  HIPCHK(s->s_stage.alloc(n * std::max(M.npair, 1) * MAXHIT * CW));
  HIPCHK(s->s_gjac.alloc(n * (int64_t)M.nrmax * d.nv));
  HIPCHK(s->s_grow.alloc(n * (int64_t)M.nrmax * 10));
  HIPCHK(s->s_islf.alloc(std::max<int64_t>(n * (int64_t)M.isl_fstride, 1)));
  HIPCHK(s->s_gmj.alloc(M.solver_kind == 1 ? n * (int64_t)M.nrmax * d.nv : 1));
  for (auto* b : {&s->s_ncon, &s->s_iters, &s->s_nrows, &s->s_status, &s->s_npcon, &s->s_cost}) {
    HIPCHK(b->alloc(n));
    HIPCHK(b->zero());
  }
  S.qpos = s->s_qpos.p;
  S.qvel = s->s_qvel.p;
  S.ctrl = s->s_ctrl.p;
  S.qacc = s->s_qacc.p;
  S.applied = s->s_applied.p;
  S.time = s->s_time.p;
  S.geom_xpos = s->s_gxpos.p;
  S.geom_xmat = s->s_gxmat.p;
  S.xpos = s->s_xpos.p;
  S.xmat = s->s_xmat.p;
  S.ncon = s->s_ncon.p;
  S.iters = s->s_iters.p;
  S.nrows = s->s_nrows.p;
  S.status = s->s_status.p;
  S.cost = s->s_cost.p;
  S.con = s->s_con.p;
  S.cgeom = s->s_cgeom.p;
  S.pcon = s->s_pcon.p;
  S.pcgeom = s->s_pcgeom.p;
  S.plambda = s->s_plambda.p;
  S.npcon = s->s_npcon.p;
  S.stage = s->s_stage.p;
  S.gjac = s->s_gjac.p;
  S.grow = s->s_grow.p;
  S.isl_f = s->s_islf.p;
  S.gmj = s->s_gmj.p;

  // invweight0 (forward::set_const) for the base variant, then poses (PippinCpu::new runs kinematics)
  if (int rc = run_setconst(s.get(), {0})) return rc;
  if (int rc = upload_variants(s.get())) return rc;
  if (num_envs > 0) {
    if (int rc = sync_desc(s.get())) return rc;
  }
  *out = s.release();
  return PH_OK;
}

extern "C" void pippin_hip_destroy(ph_sim* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  hipStream_t st = s->stream;
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  delete s;
  if (st) (void)hipStreamDestroy(st);
}

extern "C" int64_t pippin_hip_num_envs(const ph_sim* s) { return s ? s->n : -1; }

extern "C" int pippin_hip_dims(const ph_sim* s, ph_dims* out) {
  if (!s || !out) return fail(PH_ERR_ARG, "null argument");
  out->nq = s->hm.d.nq;
  out->nv = s->hm.d.nv;
  out->nu = s->hm.d.nu;
  out->ngeom = s->hm.d.ngeom;
  out->ncam = s->hm.d.ncam;
  return PH_OK;
}

static int check_range(const ph_sim* s, int64_t b, int64_t e) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  if (b < 0 || e < b || e > s->n) return fail(PH_ERR_ARG, "env range out of bounds");
  return PH_OK;
}
static int check_len(const char* what, int64_t got, int64_t envs, int64_t width) {
  if (got != envs * width)
    return fail(PH_ERR_SIZE, std::string(what) + ": buffer has " + std::to_string(got) + " elements, expected " +
                                 std::to_string(envs) + " x " + std::to_string(width));
  return PH_OK;
}

static int status_of(ph_sim* s, int64_t b, int64_t e) {
  std::vector<int> st(e - b);
  if (e > b) HIPCHK(hipMemcpy(st.data(), s->s_status.p + b, (e - b) * sizeof(int), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < e - b; i++) {
    if (st[i] & ST_NOTPD) return fail(PH_ERR_NOTPD, "mass matrix is not positive definite (env " + std::to_string(b + i) + ")");
    if (st[i] & (ST_ROW_OVERFLOW | ST_CONTACT_OVERFLOW))
      return fail(PH_ERR_DEVICE, "constraint capacity exceeded (env " + std::to_string(b + i) + ")");
  }
  return PH_OK;
}

static int field_width(const ph_sim* s, int32_t f, double** dptr) {
  switch (f) {
    case PH_FIELD_QPOS: *dptr = s->s_qpos.p; return s->hm.d.nq;
    case PH_FIELD_QVEL: *dptr = s->s_qvel.p; return s->hm.d.nv;
    case PH_FIELD_CTRL: *dptr = s->s_ctrl.p; return s->hm.d.nu;
    case PH_FIELD_QACC: *dptr = s->s_qacc.p; return s->hm.d.nv;
    case PH_FIELD_QFRC_APPLIED: *dptr = s->s_applied.p; return s->hm.d.nv;
    case PH_FIELD_TIME: *dptr = s->s_time.p; return 1;
    default: return -1;
  }
}

// Workgroups of the work-stealing kernel that the device holds at once
// (occupancy taken at create, so that nothing but launches runs while a
// step is captured in a graph).
static unsigned ws_grid(const ph_sim* s, int K, int spec, int64_t nslots) {
  const int per_cu = std::max(s->ws_per_cu[K == 4 ? 2 : K - 1][spec], 1);
  return (unsigned)std::min<int64_t>((int64_t)per_cu * s->cus, nslots);
}

// The step without the learned warm start in the given launch shape, with
// optional per-step controls and qpos outputs (Batch::rollout). Only kernels
// go on the stream, so the sequence can be captured in a graph.
static hipError_t launch_plain(ph_sim* s, const ph_launch_shape& sh, int64_t begin, int64_t end, int64_t nstep,
                               const Real* ctrl = nullptr, Real* qpos_out = nullptr) {
  const int K = sh.envs_per_wave, spec = sh.specialised;
  const size_t lds = (size_t)K * s->M.L.bytes;
  const DModel* mp = s->d_model.p;
  const DState* sp = s->d_state.p;
  if (sh.schedule == PH_SCHED_PGS_SPLIT) {  // three kernels per physics step, one env per workgroup
    const unsigned grid = (unsigned)(end - begin);
    double* vel = s->d_pgs_vel.p;
    double* vpos = s->d_pgs_vpos.p;
    for (int64_t t = 0; t < nstep; t++) {
      hipLaunchKernelGGL(k_pgs_front<1>, dim3(grid), dim3(64), lds, s->stream, mp, sp, begin, end, nstep, t, ctrl, vel,
                         vpos);
      hipLaunchKernelGGL(k_pgs_sweeps<1>, dim3(grid), dim3(64), 0, s->stream, mp, sp, begin, end, vel, vpos);
      hipLaunchKernelGGL(k_pgs_back<1>, dim3(grid), dim3(64), lds, s->stream, mp, sp, begin, end, nstep, t, qpos_out,
                         (const double*)vel, (const double*)vpos);
      if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
  }
  if (sh.schedule == PH_SCHED_GRID) {
    const unsigned grid = (unsigned)((end - begin + K - 1) / K);
    void* args[] = {&mp, &sp, &begin, &end, &nstep, &ctrl, &qpos_out};
    return hipLaunchKernel(step_kernel(K, spec), dim3(grid), dim3(64), args, lds, s->stream);
  }
  int* order = nullptr;
  unsigned* counter = s->d_counter.p;
  if (sh.schedule == PH_SCHED_WS_BUCKET) {
    order = s->d_order.p;
    hipLaunchKernelGGL(k_bucket, dim3(1), dim3(1024), 0, s->stream, sp, begin, end, order, counter);
  } else {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s->stream, counter);
  }
  if (hipError_t e = hipGetLastError()) return e;
  const int64_t nslots = (end - begin + K - 1) / K;
  const int* corder = order;
  void* args[] = {&mp, &sp, &begin, &end, &nstep, &ctrl, &qpos_out, &corder, &counter};
  return hipLaunchKernel(ws_kernel(K, spec), dim3(ws_grid(s, K, spec, nslots)), dim3(64), args, lds, s->stream);
}

static ph_launch_shape current_shape(const ph_sim* s) {
  ph_launch_shape sh{};
  sh.envs_per_wave = s->envs_per_wg;
  sh.specialised = s->spec;
  sh.schedule = s->sched;
  return sh;
}

// One step launch over [begin, end) in the current launch shape, or the
// learned-warm-start kernel when that is on.
static hipError_t launch_step(ph_sim* s, int64_t begin, int64_t end, int64_t nstep, const Real* ctrl,
                              Real* qpos_out) {
  if (!s->M.mlp.mode) return launch_plain(s, current_shape(s), begin, end, nstep, ctrl, qpos_out);
  const int K = s->envs_per_wg;
  const unsigned grid = (unsigned)((end - begin + K - 1) / K);
  const size_t lds = (size_t)K * s->M.L.bytes;
  phm::launch_step(K, s->spec, grid, lds, s->stream, s->d_model.p, s->d_state.p, begin, end, nstep, ctrl, qpos_out);
  if ((s->M.mlp.mode & MLP_LOG) && nstep > 0) {
    s->log_begin = begin;
    s->log_end = end;
  }
  return hipGetLastError();
}

static bool same_shape(const ph_launch_shape& a, const ph_launch_shape& b) {
  return a.envs_per_wave == b.envs_per_wave && a.specialised == b.specialised && a.schedule == b.schedule;
}

static bool shape_valid(const ph_sim* s, const ph_launch_shape& sh) {
  const int K = sh.envs_per_wave;
  if (K != 1 && K != 2 && K != 4) return false;
  if (sh.specialised != 0 && !(sh.specialised == 1 && K == 4 && SpecAnt::matches(s->M))) return false;
  if (sh.schedule < PH_SCHED_GRID || sh.schedule > PH_SCHED_PGS_SPLIT) return false;
  // the split PGS step: PGS models, one env per workgroup, the generic kernels
  if (sh.schedule == PH_SCHED_PGS_SPLIT && (s->M.solver_kind != 1 || K != 1 || sh.specialised != 0)) return false;
  int dev_lds = 0;
  if (hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) != hipSuccess) return false;
  return (int64_t)K * s->M.L.bytes <= dev_lds;
}

// Key of an autotuned shape: the model's sizes and layout, env count, device.
static std::string tune_key(const ph_sim* s) {
  const DModel& M = s->M;
  char buf[256];
  snprintf(buf, sizeof buf, "v1 nq%d nv%d nu%d nb%d ng%d np%d nc%d nr%d sk%d it%d L%d n%lld %s cu%d", M.nq, M.nv, M.nu,
           M.nbody, M.ngeom, M.npair, M.ncmax, M.nrmax, M.solver_kind, M.iterations, M.L.bytes, (long long)s->n,
           s->arch.c_str(), s->cus);
  std::string k = buf;
  for (char& c : k)
    if (c == ':') c = '_';
  return k;
}

static void tune_finish(ph_sim* s) {
  double best = 0.0;
  int bi = -1;
  for (size_t c = 0; c < s->tune_cand.size(); c++) {
    std::vector<double> v = s->tune_ms[c];
    if (v.empty()) continue;
    std::sort(v.begin(), v.end());
    const double med = v[v.size() / 2];
    if (bi < 0 || med < best) {
      best = med;
      bi = (int)c;
    }
  }
  if (bi >= 0) {
    const ph_launch_shape& sh = s->tune_cand[bi];
    s->envs_per_wg = sh.envs_per_wave;
    s->spec = sh.specialised;
    s->sched = sh.schedule;
    if (!s->tune_cache.empty())
      if (FILE* f = fopen(s->tune_cache.c_str(), "a")) {
        fprintf(f, "%s : %d %d %d\n", tune_key(s).c_str(), sh.envs_per_wave, sh.specialised, sh.schedule);
        fclose(f);
      }
  }
  s->tune_cand.clear();
  s->tune_ms.clear();
  s->tune_rounds = 0;
}

extern "C" int pippin_hip_get_launch_shape(const ph_sim* s, ph_launch_shape* out) {
  if (!s || !out) return fail(PH_ERR_ARG, "null argument");
  *out = current_shape(s);
  return PH_OK;
}

extern "C" int pippin_hip_set_launch_shape(ph_sim* s, const ph_launch_shape* sh) {
  if (!s || !sh) return fail(PH_ERR_ARG, "null argument");
  if (!shape_valid(s, *sh)) return fail(PH_ERR_ARG, "launch shape not available for this model or device");
  s->envs_per_wg = sh->envs_per_wave;
  s->spec = sh->specialised;
  s->sched = sh->schedule;
  s->tune_cand.clear();
  s->tune_ms.clear();
  s->tune_rounds = 0;
  return PH_OK;
}

extern "C" int pippin_hip_autotune(ph_sim* s, int32_t rounds, const char* cache_path) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  if (rounds < 1) return fail(PH_ERR_ARG, "autotune needs at least one round");
  s->tune_cache = cache_path ? cache_path : "";
  if (!s->tune_cache.empty())
    if (FILE* f = fopen(s->tune_cache.c_str(), "r")) {
      const std::string key = tune_key(s);
      char line[512];
      ph_launch_shape found{};
      bool have = false;
      while (fgets(line, sizeof line, f)) {
        std::string l = line;
        const size_t colon = l.rfind(" : ");
        if (colon == std::string::npos || l.substr(0, colon) != key) continue;
        ph_launch_shape sh{};
        if (sscanf(l.c_str() + colon + 3, "%d %d %d", &sh.envs_per_wave, &sh.specialised, &sh.schedule) == 3 &&
            shape_valid(s, sh)) {
          found = sh;
          have = true;
        }
      }
      fclose(f);
      if (have) return pippin_hip_set_launch_shape(s, &found);
    }
  s->tune_cand.clear();
  for (int spec : {1, 0})
    for (int K : {4, 2, 1})
      for (int sched : {PH_SCHED_GRID, PH_SCHED_WS, PH_SCHED_WS_BUCKET, PH_SCHED_PGS_SPLIT}) {
        ph_launch_shape sh{};
        sh.envs_per_wave = K;
        sh.specialised = spec;
        sh.schedule = sched;
        if (shape_valid(s, sh)) s->tune_cand.push_back(sh);
      }
  s->tune_ms.assign(s->tune_cand.size(), {});
  s->tune_next = 0;
  s->tune_round = 0;
  s->tune_rounds = rounds + 1;  // the first round loads each kernel and is not timed
  return PH_OK;
}

extern "C" int pippin_hip_autotune_pending(const ph_sim* s) { return s && s->tune_rounds > 0 ? 1 : 0; }

extern "C" int pippin_hip_use_graphs(ph_sim* s, int32_t on) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  s->use_graphs = on != 0;
  return PH_OK;
}

// The launch sequence of a plain step, from the graph cache (captured on first use).
static int launch_plain_graph(ph_sim* s, const ph_launch_shape& sh, int64_t begin, int64_t end, int64_t nstep) {
  for (auto& g : s->graphs)
    if (g.begin == begin && g.end == end && g.nstep == nstep && same_shape(g.shape, sh)) {
      HIPCHK(hipGraphLaunch(g.exec, s->stream));
      return PH_OK;
    }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  hipError_t launch_err = launch_plain(s, sh, begin, end, nstep);
  if (launch_err == hipSuccess) launch_err = hipGetLastError();
  hipError_t cap_err = hipStreamEndCapture(s->stream, &graph);
  HIPCHK(launch_err);
  HIPCHK(cap_err);
  StepGraph g;
  g.begin = begin;
  g.end = end;
  g.nstep = nstep;
  g.shape = sh;
  hipError_t inst = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(inst);
  if (s->graphs.size() >= 8) {  // keep the most recent few
    (void)hipGraphExecDestroy(s->graphs.front().exec);
    s->graphs.erase(s->graphs.begin());
  }
  s->graphs.push_back(g);
  HIPCHK(hipGraphLaunch(g.exec, s->stream));
  return PH_OK;
}

// One timed launch of the step over [begin, end) for pippin_hip_step and
// pippin_hip_rollout_device: an autotuning trial shape while trials remain,
// else the current shape (from the graph cache for a plain step).
static int timed_launch(ph_sim* s, int64_t begin, int64_t end, int64_t nstep, const Real* ctrl, Real* qpos_out) {
  HIPCHK(hipSetDevice(s->device));
  const bool plain = !s->M.mlp.mode;
  const bool trial = plain && s->tune_rounds > 0 && nstep > 0;
  const ph_launch_shape sh = trial ? s->tune_cand[s->tune_next] : current_shape(s);
  HIPCHK(hipEventRecord(s->ev0, s->stream));
  if (!plain) HIPCHK(launch_step(s, begin, end, nstep, ctrl, qpos_out));
  else if (s->use_graphs && !trial && !ctrl && !qpos_out) {
    if (int rc = launch_plain_graph(s, sh, begin, end, nstep)) return rc;
  } else
    HIPCHK(launch_plain(s, sh, begin, end, nstep, ctrl, qpos_out));
  s->poses_stale = true;
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(s->ev1, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->kernel_ms += ms;
  s->kernel_launches++;
  if (trial) {
    if (s->tune_round > 0) s->tune_ms[s->tune_next].push_back(ms / ((double)(end - begin) * (double)nstep));
    if (++s->tune_next == (int)s->tune_cand.size()) {
      s->tune_next = 0;
      if (++s->tune_round == s->tune_rounds) tune_finish(s);
    }
  }
  return status_of(s, begin, end);
}

extern "C" int pippin_hip_step(ph_sim* s, int64_t begin, int64_t end, int64_t nstep) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (nstep < 0) return fail(PH_ERR_ARG, "negative step count");
  if (end == begin) return PH_OK;
  return timed_launch(s, begin, end, nstep, nullptr, nullptr);
}

extern "C" int pippin_hip_kernel_timing(ph_sim* s, int32_t reset, double* total_ms, int64_t* launches) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  if (total_ms) *total_ms = s->kernel_ms;
  if (launches) *launches = s->kernel_launches;
  if (reset) {
    s->kernel_ms = 0.0;
    s->kernel_launches = 0;
  }
  return PH_OK;
}

extern "C" int pippin_hip_set_device(ph_sim* s, int32_t field, int64_t begin, int64_t end, const double* src) {
  if (int rc = check_range(s, begin, end)) return rc;
  double* d;
  int w = field_width(s, field, &d);
  if (field != PH_FIELD_QVEL && field != PH_FIELD_CTRL) return fail(PH_ERR_ARG, "set_device takes qvel or ctrl");
  if (w == 0 || end == begin) return PH_OK;
  HIPCHK(hipMemcpyAsync(d + begin * w, src, (end - begin) * w * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  return PH_OK;
}

extern "C" int pippin_hip_rollout_device(ph_sim* s, int64_t begin, int64_t end, int64_t nstep, const double* ctrl,
                                         double* qpos_out) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (nstep < 0 || !ctrl) return fail(PH_ERR_ARG, "bad rollout arguments");
  if (end == begin) return PH_OK;
  return timed_launch(s, begin, end, nstep, ctrl, qpos_out);
}

extern "C" int pippin_hip_state_device(ph_sim* s, int32_t field, double** ptr) {
  if (!s || !ptr) return fail(PH_ERR_ARG, "null argument");
  *ptr = field == PH_FIELD_QPOS ? s->s_qpos.p : field == PH_FIELD_QVEL ? s->s_qvel.p : field == PH_FIELD_CTRL ? s->s_ctrl.p : nullptr;
  // the caller may write qpos through the pointer at any time from now on,
  // so poses are recomputed on every poses() call
  if (field == PH_FIELD_QPOS) s->qpos_shared = true;
  return *ptr ? PH_OK : fail(PH_ERR_ARG, "unknown field");
}

extern "C" void* pippin_hip_stream(ph_sim* s) { return s ? (void*)s->stream : nullptr; }


extern "C" int pippin_hip_get(const ph_sim* s, int32_t field, int64_t begin, int64_t end, double* out, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  double* d;
  int w = field_width(s, field, &d);
  if (w < 0) return fail(PH_ERR_ARG, "unknown field");
  if (int rc = check_len("get", len, end - begin, w)) return rc;
  if (w == 0 || end == begin) return PH_OK;
  // copies queued by set_device run on the library stream, which the null
  // stream does not wait for
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(out, d + begin * w, (end - begin) * w * sizeof(double), hipMemcpyDeviceToHost));
  return PH_OK;
}

extern "C" int pippin_hip_set(ph_sim* s, int32_t field, int64_t begin, int64_t end, const double* src, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  double* d;
  int w = field_width(s, field, &d);
  if (w < 0) return fail(PH_ERR_ARG, "unknown field");
  if (int rc = check_len("set", len, end - begin, w)) return rc;
  if (w == 0 || end == begin) return PH_OK;
  HIPCHK(hipStreamSynchronize(s->stream));  // a pending set_device copy must not overwrite this one
  HIPCHK(hipMemcpy(d + begin * w, src, (end - begin) * w * sizeof(double), hipMemcpyHostToDevice));
  if (field == PH_FIELD_QPOS) s->poses_stale = true;
  return PH_OK;
}

extern "C" int pippin_hip_reset(ph_sim* s, const int64_t* envs, int64_t n) {
  if (!s || (n > 0 && !envs)) return fail(PH_ERR_ARG, "null argument");
  if (n <= 0) return PH_OK;
  for (int64_t i = 0; i < n; i++)
    if (envs[i] < 0 || envs[i] >= s->n) return fail(PH_ERR_ARG, "env id out of range");
  std::vector<int64_t> ids(envs, envs + n);
  HIPCHK(s->d_ids.upload(ids));
  hipLaunchKernelGGL(k_reset, dim3((unsigned)n), dim3(64), 0, s->stream, s->d_model.p, s->d_state.p, s->d_ids.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->poses_stale = true;
  return PH_OK;
}

// pippin_hip_reset with the env ids staged in the file-scope `scratch` instead
// of a buffer allocated per call.
extern "C" int pippin_hip_reset_noalloc(ph_sim* s, const int64_t* envs, int64_t n) {
  if (!s || (n > 0 && !envs)) return fail(PH_ERR_ARG, "null argument");
  if (n <= 0) return PH_OK;
  if (n > RESET_SCRATCH_MAX)
    return fail(PH_ERR_ARG, "reset_noalloc takes at most " + std::to_string(RESET_SCRATCH_MAX) + " envs per call, got " +
                                std::to_string(n));
  for (int64_t i = 0; i < n; i++)
    if (envs[i] < 0 || envs[i] >= s->n) return fail(PH_ERR_ARG, "env id out of range");
  HIPCHK(hipSetDevice(s->device));
  int64_t* ids = s->reset_scratch;
  HIPCHK(hipMemcpyAsync(ids, envs, n * sizeof(int64_t), hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(k_reset, dim3((unsigned)n), dim3(64), 0, s->stream, s->d_model.p, s->d_state.p,
                     (const int64_t*)ids);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->poses_stale = true;
  return PH_OK;
}

// Rerun kinematics for every env when qpos changed since the last poses (or
// a device qpos pointer is out): PippinCpu's kinematics after step and set.
static int refresh_poses(const ph_sim* s) {
  if (!s->poses_stale && !s->qpos_shared) return PH_OK;
  HIPCHK(hipSetDevice(s->device));
  hipLaunchKernelGGL(k_kin, dim3((unsigned)s->n), dim3(64), s->M.L.bytes, s->stream, s->d_model.p, s->d_state.p,
                     (int64_t)0, (const int64_t*)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->poses_stale = false;
  return PH_OK;
}

extern "C" int pippin_hip_poses_device(ph_sim* s, ph_pose** geoms, ph_pose** cams) {
  if (!s || !geoms || !cams) return fail(PH_ERR_ARG, "null argument");
  const ph_model_desc& d = s->hm.d;
  HIPCHK(hipSetDevice(s->device));
  if (!s->d_pose_geom.p) {
    HIPCHK(s->d_pose_geom.alloc((size_t)std::max<int64_t>(s->n * d.ngeom, 1)));
    HIPCHK(s->d_pose_cam.alloc((size_t)std::max<int64_t>(s->n * d.ncam, 1)));
    HIPCHK(s->d_cam_body.upload(s->hm.cam_body));
    HIPCHK(s->d_cam_pos.upload(s->hm.cam_pos));
    HIPCHK(s->d_cam_quat.upload(s->hm.cam_quat));
  }
  if (int rc = refresh_poses(s)) return rc;
  const int64_t total = s->n * (d.ngeom + d.ncam);
  if (total > 0) {
    hipLaunchKernelGGL(k_pose_f32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s->stream, s->d_model.p,
                       s->d_state.p, s->d_cam_body.p, s->d_cam_pos.p, s->d_cam_quat.p, s->d_pose_geom.p,
                       s->d_pose_cam.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  *geoms = s->d_pose_geom.p;
  *cams = s->d_pose_cam.p;
  return PH_OK;
}

extern "C" int pippin_hip_poses(const ph_sim* s, int64_t begin, int64_t end, ph_pose* geoms, int64_t ngl, ph_pose* cams,
                                int64_t ncl) {
  if (int rc = check_range(s, begin, end)) return rc;
  const ph_model_desc& d = s->hm.d;
  if (int rc = check_len("poses(geoms)", ngl, end - begin, d.ngeom)) return rc;
  if (int rc = check_len("poses(cams)", ncl, end - begin, d.ncam)) return rc;
  int64_t n = end - begin;
  if (n == 0) return PH_OK;
  if (int rc = refresh_poses(s)) return rc;
  std::vector<double> gp(n * 3 * d.ngeom), gm(n * 9 * d.ngeom), bp(n * 3 * d.nbody), bm(n * 9 * d.nbody);
  if (d.ngeom) {
    HIPCHK(hipMemcpy(gp.data(), s->s_gxpos.p + begin * 3 * d.ngeom, gp.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gm.data(), s->s_gxmat.p + begin * 9 * d.ngeom, gm.size() * 8, hipMemcpyDeviceToHost));
  }
  if (d.ncam) {
    HIPCHK(hipMemcpy(bp.data(), s->s_xpos.p + begin * 3 * d.nbody, bp.size() * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(bm.data(), s->s_xmat.p + begin * 9 * d.nbody, bm.size() * 8, hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < n; i++) {
    for (int g = 0; g < d.ngeom; g++) {
      ph_pose& p = geoms[i * d.ngeom + g];
      for (int k = 0; k < 3; k++) p.pos[k] = (float)gp[(i * d.ngeom + g) * 3 + k];
      for (int k = 0; k < 9; k++) p.mat[k] = (float)gm[(i * d.ngeom + g) * 9 + k];
    }
    // camera pose from its body (pippin_cpu.rs:441-446)
    for (int c = 0; c < d.ncam; c++) {
      int b = s->hm.cam_body[c];
      M3 xm = ldm(&bm[(i * d.nbody + b) * 9]);
      V3 pos = ld3(&bp[(i * d.nbody + b) * 3]) + mul_vec(xm, ld3(&s->hm.cam_pos[3 * c]));
      M3 mat = mul_mat(xm, to_mat(ldq(&s->hm.cam_quat[4 * c])));
      ph_pose& p = cams[i * d.ncam + c];
      p.pos[0] = (float)pos.x;
      p.pos[1] = (float)pos.y;
      p.pos[2] = (float)pos.z;
      for (int k = 0; k < 9; k++) p.mat[k] = (float)mat.m[k];
    }
  }
  return PH_OK;
}

// ---------------------------------------------------------------- parameters (params.rs)
static int param_width(const ph_sim* s, int p) {
  switch (p) {
    case PH_PARAM_GEOM_SIZE:
    case PH_PARAM_ACTUATOR_BIAS:
    case PH_PARAM_BODY_INERTIA:
    case PH_PARAM_BODY_IPOS: return 3;
    case PH_PARAM_GEOM_RGBA: return 4;
    case PH_PARAM_QPOS0: return s->hm.d.nq;
    default: return 1;
  }
}
static int64_t param_count(const ph_sim* s, int p) {
  const ph_model_desc& d = s->hm.d;
  switch (p) {
    case PH_PARAM_GEOM_FRICTION:
    case PH_PARAM_GEOM_SIZE:
    case PH_PARAM_GEOM_RGBA:
    case PH_PARAM_GEOM_CONTYPE:
    case PH_PARAM_GEOM_CONAFFINITY: return d.ngeom;
    case PH_PARAM_BODY_MASS:
    case PH_PARAM_BODY_INERTIA:
    case PH_PARAM_BODY_IPOS: return d.nbody;
    case PH_PARAM_QPOS0: return 1;
    case PH_PARAM_DOF_DAMPING: return d.nv;
    case PH_PARAM_ACTUATOR_GAIN:
    case PH_PARAM_ACTUATOR_BIAS: return d.nu;
    default: return -1;
  }
}
static const char* param_name(int p) {
  static const char* names[] = {"GeomFriction", "GeomSize", "GeomRgba", "GeomContype", "GeomConaffinity", "BodyMass",
                                "DofDamping", "ActuatorGain", "ActuatorBias", "BodyInertia", "BodyIpos", "Qpos0"};
  return (p >= 0 && p < 12) ? names[p] : "?";
}
static uint32_t sat_u32(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 4294967295.0) return 4294967295u;
  return (uint32_t)x;
}

// Validate one element edit against variant `v` (params.rs:111-177) and,
// when `apply`, perform it. Returns PH_OK or PH_ERR_ARG with the message.
static int param_apply(ph_sim* s, int v, int p, int64_t id, const double* x, bool apply, bool* need_setconst) {
  Variants& V = s->V;
  const ph_model_desc& d = s->hm.d;
  int w = param_width(s, p);
  for (int i = 0; i < w; i++)
    if (!std::isfinite(x[i])) return fail(PH_ERR_ARG, std::string(param_name(p)) + ": values must be finite");
  switch (p) {
    case PH_PARAM_GEOM_FRICTION:
      if (apply) V.friction[(int64_t)v * d.ngeom + id] = std::fmax(x[0], 0.0);
      break;
    case PH_PARAM_GEOM_SIZE: {
      int t = s->hm.geom_type[id];
      if (t == PH_GEOM_MESH || t == PH_GEOM_PLANE)
        return fail(PH_ERR_ARG, "geom_size applies to sphere, capsule, box, and cylinder geoms");
      for (int i = 0; i < 3; i++)
        if (x[i] < 0.0) return fail(PH_ERR_ARG, "geom_size must be non-negative");
      if (apply) {
        double* sz = &V.geom_size[((int64_t)v * d.ngeom + id) * 3];
        for (int i = 0; i < 3; i++) sz[i] = x[i];
        phh::local_aabb(t, sz, &V.geom_aabb[((int64_t)v * d.ngeom + id) * 6]);
      }
      break;
    }
    case PH_PARAM_GEOM_RGBA:
      if (apply)
        for (int i = 0; i < 4; i++) V.rgba[((int64_t)v * d.ngeom + id) * 4 + i] = (float)x[i];
      break;
    case PH_PARAM_GEOM_CONTYPE:
    case PH_PARAM_GEOM_CONAFFINITY: {
      uint32_t bits = sat_u32(std::fmax(x[0], 0.0));
      if (apply) (p == PH_PARAM_GEOM_CONTYPE ? V.contype : V.conaffinity)[(int64_t)v * d.ngeom + id] = bits;
      break;
    }
    case PH_PARAM_BODY_MASS: {
      if (id == 0) return fail(PH_ERR_ARG, "the world body has no mass");
      if (x[0] <= 0.0 && s->hm.body_jntnum[id] > 0) return fail(PH_ERR_ARG, "moving bodies need positive mass");
      if (apply) {
        double& m = V.body_mass[(int64_t)v * d.nbody + id];
        double scale = m > 0.0 ? x[0] / m : 1.0;
        m = x[0];
        double* I = &V.body_inertia[((int64_t)v * d.nbody + id) * 9];
        for (int i = 0; i < 9; i++) I[i] = I[i] * scale;
        *need_setconst = true;
      }
      break;
    }
    case PH_PARAM_DOF_DAMPING:
      if (apply) V.dof_damping[(int64_t)v * d.nv + id] = std::fmax(x[0], 0.0);
      break;
    case PH_PARAM_ACTUATOR_GAIN:
      if (apply) V.act_gain[(int64_t)v * d.nu + id] = x[0];
      break;
    case PH_PARAM_ACTUATOR_BIAS:
      if (apply)
        for (int i = 0; i < 3; i++) V.act_bias[((int64_t)v * d.nu + id) * 3 + i] = x[i];
      break;
    case PH_PARAM_BODY_INERTIA:
      if (id == 0 || x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0)
        return fail(PH_ERR_ARG, "body_inertia needs a non-world body and non-negative moments");
      if (apply) {
        double* I = &V.body_inertia[((int64_t)v * d.nbody + id) * 9];
        for (int i = 0; i < 9; i++) I[i] = 0.0;
        I[0] = x[0];
        I[4] = x[1];
        I[8] = x[2];
        *need_setconst = true;
      }
      break;
    case PH_PARAM_BODY_IPOS:
      if (id == 0) return fail(PH_ERR_ARG, "the world body has no center of mass");
      if (apply) {
        for (int i = 0; i < 3; i++) V.body_ipos[((int64_t)v * d.nbody + id) * 3 + i] = x[i];
        *need_setconst = true;
      }
      break;
    case PH_PARAM_QPOS0:
      if (apply)
        for (int i = 0; i < d.nq; i++) V.qpos_reset[(int64_t)v * d.nq + i] = x[i];
      break;
    default: return fail(PH_ERR_ARG, "unknown parameter");
  }
  if (apply) {
    V.damped[v] = phh::has_implicit_terms(s->hm, &V.dof_damping[(int64_t)v * d.nv], &V.act_bias[(int64_t)v * d.nu * 3])
                      ? 1 : 0;
  }
  return PH_OK;
}

// pippin_cpu.rs:365-374: validate on the base values first (no partial update), then copy on write per env.
extern "C" int pippin_hip_set_param(ph_sim* s, int32_t p, int64_t id, int64_t begin, int64_t end,
                                    const double* values, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  int64_t cnt = param_count(s, p);
  if (cnt < 0) return fail(PH_ERR_ARG, "unknown parameter");
  int w = param_width(s, p);
  if (int rc = check_len("set_param", len, end - begin, w)) return rc;
  if (id < 0 || id >= cnt)
    return fail(PH_ERR_ARG, std::string(param_name(p)) + ": id " + std::to_string(id) + " out of range (have " +
                                std::to_string(cnt) + ")");
  bool dummy = false;
  if (end > begin)
    if (int rc = param_apply(s, 0, p, id, values, false, &dummy)) return rc;
  for (int64_t i = 0; i < end - begin; i++)
    if (int rc = param_apply(s, s->env_var[begin + i], p, id, values + i * w, false, &dummy)) return rc;
  std::vector<int> recompute;
  for (int64_t i = 0; i < end - begin; i++) {
    int64_t e = begin + i;
    int v = s->env_var[e];
    if (s->var_refs[v] > 1) {
      int nv = variant_clone(s, v);
      s->var_refs[v]--;
      s->var_refs.push_back(1);
      s->env_var[e] = nv;
      v = nv;
    }
    bool sc = false;
    param_apply(s, v, p, id, values + i * w, true, &sc);
    if (sc) recompute.push_back(v);
  }
  HIPCHK(up(s->d_env_var, s->env_var, &s->M.env_var));
  if (int rc = upload_variants(s)) return rc;
  if (!recompute.empty()) {
    if (int rc = run_setconst(s, recompute)) return rc;
    if (int rc = upload_variants(s)) return rc;
  }
  return PH_OK;
}

extern "C" int pippin_hip_get_param(const ph_sim* s, int32_t p, int64_t id, int64_t begin, int64_t end, double* out,
                                    int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  int64_t cnt = param_count(s, p);
  if (cnt < 0) return fail(PH_ERR_ARG, "unknown parameter");
  int w = param_width(s, p);
  if (int rc = check_len("get_param", len, end - begin, w)) return rc;
  if (id < 0 || id >= cnt) return fail(PH_ERR_ARG, std::string(param_name(p)) + ": id out of range");
  const Variants& V = s->V;
  const ph_model_desc& d = s->hm.d;
  for (int64_t i = 0; i < end - begin; i++) {
    int64_t v = s->env_var[begin + i];
    double* o = out + i * w;
    switch (p) {
      case PH_PARAM_GEOM_FRICTION: o[0] = V.friction[v * d.ngeom + id]; break;
      case PH_PARAM_GEOM_SIZE:
        for (int k = 0; k < 3; k++) o[k] = V.geom_size[(v * d.ngeom + id) * 3 + k];
        break;
      case PH_PARAM_GEOM_RGBA:
        for (int k = 0; k < 4; k++) o[k] = V.rgba[(v * d.ngeom + id) * 4 + k];
        break;
      case PH_PARAM_GEOM_CONTYPE: o[0] = V.contype[v * d.ngeom + id]; break;
      case PH_PARAM_GEOM_CONAFFINITY: o[0] = V.conaffinity[v * d.ngeom + id]; break;
      case PH_PARAM_BODY_MASS: o[0] = V.body_mass[v * d.nbody + id]; break;
      case PH_PARAM_DOF_DAMPING: o[0] = V.dof_damping[v * d.nv + id]; break;
      case PH_PARAM_ACTUATOR_GAIN: o[0] = V.act_gain[v * d.nu + id]; break;
      case PH_PARAM_ACTUATOR_BIAS:
        for (int k = 0; k < 3; k++) o[k] = V.act_bias[(v * d.nu + id) * 3 + k];
        break;
      case PH_PARAM_BODY_INERTIA:
        for (int k = 0; k < 3; k++) o[k] = V.body_inertia[(v * d.nbody + id) * 9 + 4 * k];
        break;
      case PH_PARAM_BODY_IPOS:
        for (int k = 0; k < 3; k++) o[k] = V.body_ipos[(v * d.nbody + id) * 3 + k];
        break;
      case PH_PARAM_QPOS0:
        for (int k = 0; k < d.nq; k++) o[k] = V.qpos_reset[v * d.nq + k];
        break;
    }
  }
  return PH_OK;
}

// pippin_cpu.rs:388-400
extern "C" int pippin_hip_appearance(const ph_sim* s, int64_t begin, int64_t end, ph_appearance* out, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  const ph_model_desc& d = s->hm.d;
  if (int rc = check_len("appearance", len, end - begin, d.ngeom)) return rc;
  for (int64_t i = 0; i < end - begin; i++) {
    int64_t v = s->env_var[begin + i];
    for (int g = 0; g < d.ngeom; g++) {
      ph_appearance& a = out[i * d.ngeom + g];
      for (int k = 0; k < 4; k++) a.rgba[k] = s->V.rgba[(v * d.ngeom + g) * 4 + k];
      for (int k = 0; k < 3; k++) {
        double base = s->hm.geom_size[3 * g + k];
        a.scale[k] = base > 0.0 ? (float)(s->V.geom_size[(v * d.ngeom + g) * 3 + k] / base) : 1.0f;
      }
      a.pad = 0.0f;
    }
  }
  return PH_OK;
}

// ---------------------------------------------------------------- diagnostics
extern "C" int pippin_hip_ncon(const ph_sim* s, int64_t begin, int64_t end, int32_t* out, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (int rc = check_len("ncon", len, end - begin, 1)) return rc;
  if (end > begin) HIPCHK(hipMemcpy(out, s->s_ncon.p + begin, (end - begin) * sizeof(int), hipMemcpyDeviceToHost));
  return PH_OK;
}

extern "C" int pippin_hip_solver_iterations(const ph_sim* s, int64_t begin, int64_t end, int32_t* out, int64_t len) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (int rc = check_len("solver_iterations", len, end - begin, 1)) return rc;
  if (end > begin) HIPCHK(hipMemcpy(out, s->s_iters.p + begin, (end - begin) * sizeof(int), hipMemcpyDeviceToHost));
  return PH_OK;
}

extern "C" int pippin_hip_contacts(const ph_sim* s, int64_t env, ph_contact* out, int32_t cap, int32_t* n) {
  if (int rc = check_range(s, env, env + 1)) return rc;
  int nc = 0;
  HIPCHK(hipMemcpy(&nc, s->s_ncon.p + env, sizeof(int), hipMemcpyDeviceToHost));
  *n = nc;
  int m = std::min(nc, (int)cap);
  if (m <= 0) return PH_OK;
  std::vector<double> c(m * CW);
  std::vector<int> g(m * 3);
  HIPCHK(hipMemcpy(c.data(), s->s_con.p + env * s->M.ncmax * CW, c.size() * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(g.data(), s->s_cgeom.p + env * s->M.ncmax * 3, g.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < m; i++) {
    for (int k = 0; k < 3; k++) {
      out[i].pos[k] = c[i * CW + k];
      out[i].normal[k] = c[i * CW + 3 + k];
      out[i].tangent[k] = c[i * CW + 6 + k];
    }
    out[i].depth = c[i * CW + 9];
    out[i].geom1 = g[3 * i];
    out[i].geom2 = g[3 * i + 1];
    out[i].feature = g[3 * i + 2];
    out[i].pad = 0;
  }
  return PH_OK;
}

extern "C" int pippin_hip_history_capacity(const ph_sim* s, int32_t* cap) {
  if (!s || !cap) return fail(PH_ERR_ARG, "null argument");
  *cap = s->M.ncmax;
  return PH_OK;
}

extern "C" int pippin_hip_get_history(const ph_sim* s, int64_t begin, int64_t end, int32_t cap, int32_t* counts,
                                      ph_contact* out, double* lambda) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (cap < 0 || !counts || (cap > 0 && (!out || !lambda))) return fail(PH_ERR_ARG, "null argument or negative capacity");
  const int64_t n = end - begin;
  if (n == 0) return PH_OK;
  const int64_t nc = s->M.ncmax;
  HIPCHK(hipStreamSynchronize(s->stream));
  std::vector<int> cnt(n);
  HIPCHK(hipMemcpy(cnt.data(), s->s_npcon.p + begin, n * sizeof(int), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++)
    if (cnt[i] > cap)
      return fail(PH_ERR_SIZE, "get_history: env " + std::to_string(begin + i) + " holds " + std::to_string(cnt[i]) +
                                   " contacts, capacity " + std::to_string(cap));
  std::vector<double> c(n * nc * CW), l(n * nc * 3);
  std::vector<int> g(n * nc * 3);
  HIPCHK(hipMemcpy(c.data(), s->s_pcon.p + begin * nc * CW, c.size() * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(g.data(), s->s_pcgeom.p + begin * nc * 3, g.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(l.data(), s->s_plambda.p + begin * nc * 3, l.size() * 8, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++) {
    counts[i] = cnt[i];
    for (int64_t k = 0; k < cap; k++) {
      ph_contact& o = out[i * cap + k];
      double* lo = lambda + (i * cap + k) * 3;
      if (k >= cnt[i]) {
        o = ph_contact{};
        lo[0] = lo[1] = lo[2] = 0.0;
        continue;
      }
      const double* ci = &c[(i * nc + k) * CW];
      for (int t = 0; t < 3; t++) {
        o.pos[t] = ci[t];
        o.normal[t] = ci[3 + t];
        o.tangent[t] = ci[6 + t];
        lo[t] = l[(i * nc + k) * 3 + t];
      }
      o.depth = ci[9];
      o.geom1 = g[(i * nc + k) * 3];
      o.geom2 = g[(i * nc + k) * 3 + 1];
      o.feature = g[(i * nc + k) * 3 + 2];
      o.pad = 0;
    }
  }
  return PH_OK;
}

extern "C" int pippin_hip_set_history(ph_sim* s, int64_t begin, int64_t end, int32_t cap, const int32_t* counts,
                                      const ph_contact* in, const double* lambda) {
  if (int rc = check_range(s, begin, end)) return rc;
  if (cap < 0 || !counts || (cap > 0 && (!in || !lambda))) return fail(PH_ERR_ARG, "null argument or negative capacity");
  const int64_t n = end - begin;
  if (n == 0) return PH_OK;
  const int64_t nc = s->M.ncmax;
  const int ngeom = s->hm.d.ngeom;
  // validate everything before touching any env
  for (int64_t i = 0; i < n; i++) {
    const std::string env = "set_history: env " + std::to_string(begin + i);
    if (counts[i] < 0 || counts[i] > cap || counts[i] > nc)
      return fail(PH_ERR_ARG, env + ": count " + std::to_string(counts[i]) + " outside [0, min(cap, " +
                                  std::to_string(nc) + ")]");
    for (int k = 0; k < counts[i]; k++) {
      const ph_contact& c = in[i * cap + k];
      const double* lk = lambda + (i * (int64_t)cap + k) * 3;
      if (c.geom1 < 0 || c.geom1 >= ngeom || c.geom2 < 0 || c.geom2 >= ngeom)
        return fail(PH_ERR_ARG, env + ": geom id out of range");
      bool finite = std::isfinite(c.depth) && std::isfinite(lk[0]) && std::isfinite(lk[1]) && std::isfinite(lk[2]);
      for (int t = 0; t < 3; t++)
        finite = finite && std::isfinite(c.pos[t]) && std::isfinite(c.normal[t]) && std::isfinite(c.tangent[t]);
      if (!finite) return fail(PH_ERR_ARG, env + ": non-finite contact or impulse");
    }
  }
  std::vector<double> c(n * nc * CW, 0.0), l(n * nc * 3, 0.0);
  std::vector<int> g(n * nc * 3, 0), cnt(counts, counts + n);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < counts[i]; k++) {
      const ph_contact& ci = in[i * cap + k];
      double* co = &c[(i * nc + k) * CW];
      for (int t = 0; t < 3; t++) {
        co[t] = ci.pos[t];
        co[3 + t] = ci.normal[t];
        co[6 + t] = ci.tangent[t];
        l[(i * nc + k) * 3 + t] = lambda[(i * (int64_t)cap + k) * 3 + t];
      }
      co[9] = ci.depth;
      g[(i * nc + k) * 3] = ci.geom1;
      g[(i * nc + k) * 3 + 1] = ci.geom2;
      g[(i * nc + k) * 3 + 2] = ci.feature;
    }
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->s_pcon.p + begin * nc * CW, c.data(), c.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->s_pcgeom.p + begin * nc * 3, g.data(), g.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->s_plambda.p + begin * nc * 3, l.data(), l.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->s_npcon.p + begin, cnt.data(), n * sizeof(int), hipMemcpyHostToDevice));
  return PH_OK;
}

extern "C" int pippin_hip_mass_bias(ph_sim* s, int64_t begin, int64_t end, double* qm, int64_t qm_len, double* bias,
                                    int64_t bias_len) {
  if (int rc = check_range(s, begin, end)) return rc;
  int nv = s->hm.d.nv;
  if (int rc = check_len("mass_bias(qm)", qm_len, end - begin, (int64_t)nv * nv)) return rc;
  if (int rc = check_len("mass_bias(bias)", bias_len, end - begin, nv)) return rc;
  int64_t n = end - begin;
  if (n == 0) return PH_OK;
  DBuf<double> dq, db;
  HIPCHK(dq.alloc(n * nv * nv));
  HIPCHK(db.alloc(n * nv));
  hipLaunchKernelGGL(k_mass_bias, dim3((unsigned)n), dim3(64), s->M.L.bytes, s->stream, s->d_model.p, s->d_state.p, begin, dq.p, db.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(qm, dq.p, n * nv * nv * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(bias, db.p, n * nv * 8, hipMemcpyDeviceToHost));
  return PH_OK;
}

extern "C" int pippin_hip_invweight(const ph_sim* s, int64_t begin, int64_t end, double* dof, int64_t dof_len,
                                    double* body, int64_t body_len) {
  if (int rc = check_range(s, begin, end)) return rc;
  const ph_model_desc& d = s->hm.d;
  if (int rc = check_len("invweight(dof)", dof_len, end - begin, d.nv)) return rc;
  if (int rc = check_len("invweight(body)", body_len, end - begin, 2 * d.nbody)) return rc;
  for (int64_t i = 0; i < end - begin; i++) {
    int64_t v = s->env_var[begin + i];
    for (int k = 0; k < d.nv; k++) dof[i * d.nv + k] = s->V.dof_invweight[v * d.nv + k];
    for (int k = 0; k < 2 * d.nbody; k++) body[i * 2 * d.nbody + k] = s->V.body_invweight[v * 2 * d.nbody + k];
  }
  return PH_OK;
}

extern "C" int pippin_hip_collision_pairs(const ph_sim* s, int64_t env, int32_t* pairs, int32_t cap, int32_t* n) {
  if (int rc = check_range(s, env, env + 1)) return rc;
  int64_t v = s->env_var[env];
  const ph_model_desc& d = s->hm.d;
  int cnt = 0;
  for (size_t p = 0; p < s->hm.pair_g1.size(); p++) {
    int g1 = s->hm.pair_g1[p], g2 = s->hm.pair_g2[p];
    uint32_t t1 = s->V.contype[v * d.ngeom + g1], a1 = s->V.conaffinity[v * d.ngeom + g1];
    uint32_t t2 = s->V.contype[v * d.ngeom + g2], a2 = s->V.conaffinity[v * d.ngeom + g2];
    if ((t1 & a2) == 0 && (t2 & a1) == 0) continue;
    if (cnt < cap) {
      pairs[2 * cnt] = g1;
      pairs[2 * cnt + 1] = g2;
    }
    cnt++;
  }
  *n = cnt;
  return PH_OK;
}

extern "C" int pippin_hip_selftest_math(const double* x, int64_t n, double* out, int32_t device) {
  if (n <= 0) return PH_OK;
  if (!x || !out) return fail(PH_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(device));
  DBuf<double> dx, dout;
  HIPCHK(dx.alloc(n));
  HIPCHK(dout.alloc(4 * n));
  HIPCHK(hipMemcpy(dx.p, x, n * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_selftest_math, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, dx.p, n, dout.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(out, dout.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost));
  return PH_OK;
}

// ---------------------------------------------------------------- learned warm start
// (no counterpart in the reference; the network and its use are described in mlp.hpp)

static int mlp_check(ph_sim* s) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  if (!s->mlp_ready) return fail(PH_ERR_ARG, "learned warm start not initialised (pippin_hip_mlp_init)");
  return PH_OK;
}

extern "C" int pippin_hip_mlp_init(ph_sim* s, int32_t hidden, uint64_t seed, int64_t capacity, int64_t max_batch) {
  if (!s) return fail(PH_ERR_ARG, "null sim");
  if (hidden < 16 || hidden > MLP_HMAX || hidden % 16)
    return fail(PH_ERR_ARG, "mlp_init: hidden width must be a multiple of 16 in [16, 256]");
  if (max_batch < 16 || max_batch % 16) return fail(PH_ERR_ARG, "mlp_init: max_batch must be a positive multiple of 16");
  if (capacity < std::max<int64_t>(s->n * MLP_CL, max_batch))
    return fail(PH_ERR_ARG, "mlp_init: capacity must hold one collect (envs x 16) and one batch");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int h = hidden;
  const int64_t np = mlp_nparam(h), nw = mlp_nweight(h), n = std::max<int64_t>(s->n, 1);
  // weights ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)) over the real fan-in, padding and biases 0
  std::vector<float> P(np, 0.0f);
  for (int l = 0; l < 4; l++) {
    const int IN = mlp_in(l, h), OUT = mlp_out(l, h);
    const int fin = l == 0 ? MLP_NF : h;
    const double bound = 1.0 / std::sqrt((double)fin);
    for (int i = 0; i < IN; i++)
      for (int o = 0; o < OUT; o++) {
        if ((l == 0 && i >= MLP_NF) || (l == 3 && o >= MLP_NO)) continue;
        const int64_t idx = mlp_woff(l, h) + (int64_t)i * OUT + o;
        const double u = (double)(mlp_mix(seed + (uint64_t)idx) >> 11) * (1.0 / 9007199254740992.0);
        P[idx] = (float)((2.0 * u - 1.0) * bound);
      }
  }
  const int splits = 64;
  HIPCHK(s->m_P.upload(P));
  for (auto* b : {&s->m_m, &s->m_v, &s->m_G}) {
    HIPCHK(b->alloc(np));
    HIPCHK(b->zero());
  }
  HIPCHK(s->m_wbf.alloc(nw));
  HIPCHK(s->m_wtbf.alloc(nw));
  HIPCHK(s->m_x0.alloc(max_batch * MLP_KIN));
  HIPCHK(s->m_z.alloc(3 * max_batch * h));
  HIPCHK(s->m_a.alloc(3 * max_batch * h));
  HIPCHK(s->m_dz.alloc(3 * max_batch * h + max_batch * MLP_NOP));
  HIPCHK(s->m_y.alloc(max_batch * MLP_NOP));
  HIPCHK(s->m_loss.alloc(max_batch / 16));
  HIPCHK(s->m_gp.alloc(splits * np));
  HIPCHK(s->m_xin.alloc(max_batch * MLP_NF));
  HIPCHK(s->m_tin.alloc(max_batch * MLP_NO));
  HIPCHK(s->m_rfeat.alloc(capacity * MLP_NF));
  HIPCHK(s->m_rtgt.alloc(capacity * MLP_NO));
  HIPCHK(s->m_feat.alloc(n * MLP_CL * MLP_NF));
  HIPCHK(s->m_feat.zero());
  HIPCHK(s->m_pred.alloc(n * MLP_CL * MLP_NO));
  HIPCHK(s->m_pred.zero());
  HIPCHK(s->m_tgt.alloc(n * MLP_CL * MLP_NO));
  HIPCHK(s->m_tgt.zero());
  HIPCHK(s->m_cnt.alloc(n));
  HIPCHK(s->m_cnt.zero());
  HIPCHK(s->m_total.alloc(1 + n));  // appended count, then per-env ring offsets
  s->net = phm::Net{h, s->m_P.p, s->m_m.p, s->m_v.p, s->m_G.p, s->m_wbf.p, s->m_wtbf.p};
  s->batch = phm::Batch{max_batch, splits, s->m_x0.p, s->m_z.p, s->m_a.p, s->m_dz.p, s->m_y.p, s->m_loss.p, s->m_gp.p};
  phm::to_bf16(s->net, s->stream);
  HIPCHK(hipGetLastError());
  {
    const int bytes = s->envs_per_wg * s->M.L.bytes;
    if (bytes > 64 * 1024)
      HIPCHK(hipFuncSetAttribute(phm::step_kernel(s->envs_per_wg, s->spec), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 bytes));
  }
  s->S.mlp_feat = s->m_feat.p;
  s->S.mlp_pred = s->m_pred.p;
  s->S.mlp_tgt = s->m_tgt.p;
  s->S.mlp_cnt = s->m_cnt.p;
  s->M.mlp = phm::view(s->net, 0);
  s->ring_cap = capacity;
  s->ring_head = s->ring_size = s->adam_t = 0;
  s->log_begin = s->log_end = 0;
  s->mlp_ready = true;
  if (int rc = sync_desc(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return PH_OK;
}

extern "C" int pippin_hip_mlp_mode(ph_sim* s, int32_t mode) {
  if (int rc = mlp_check(s)) return rc;
  if (mode & ~(MLP_LOG | MLP_WARM)) return fail(PH_ERR_ARG, "mlp_mode: unknown flag");
  if ((mode & MLP_WARM) && s->net.h > MLP_HMAX_STEP)
    return fail(PH_ERR_ARG, "mlp_mode: the step kernel runs networks up to 64 wide");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->M.mlp.mode = mode;
  s->log_begin = s->log_end = 0;
  return sync_desc(s);
}

extern "C" int pippin_hip_mlp_collect(ph_sim* s, int64_t* stored) {
  if (int rc = mlp_check(s)) return rc;
  HIPCHK(hipSetDevice(s->device));
  int64_t total = 0;
  if (s->log_end > s->log_begin) {
    phm::collect(s->m_cnt.p, s->m_feat.p, s->m_tgt.p, s->log_begin, s->log_end, s->m_rfeat.p, s->m_rtgt.p,
                 s->ring_cap, s->ring_head, s->m_total.p, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&total, s->m_total.p, sizeof(int64_t), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->ring_head = (s->ring_head + total) % s->ring_cap;
    s->ring_size = std::min(s->ring_cap, s->ring_size + total);
    s->log_begin = s->log_end = 0;  // each launch's records are appended once
  }
  if (stored) *stored = s->ring_size;
  return PH_OK;
}

extern "C" int pippin_hip_mlp_train(ph_sim* s, int32_t iters, int64_t batch, float lr, uint64_t seed,
                                    float* loss_out) {
  if (int rc = mlp_check(s)) return rc;
  if (batch < 16 || batch % 16 || batch > s->batch.bmax)
    return fail(PH_ERR_ARG, "mlp_train: batch must be a multiple of 16 up to max_batch");
  if (s->ring_size == 0) return fail(PH_ERR_ARG, "mlp_train: no recorded contacts (log, step, collect first)");
  if (!(lr > 0.0f)) return fail(PH_ERR_ARG, "mlp_train: learning rate must be positive");
  HIPCHK(hipSetDevice(s->device));
  std::vector<float> part(batch / 16);
  for (int32_t it = 0; it < iters; it++) {
    phm::forward_backward(s->net, s->batch, s->m_rfeat.p, s->m_rtgt.p, s->ring_size, batch, false, seed, s->adam_t,
                          true, s->stream);
    phm::adam(s->net, ++s->adam_t, lr, s->stream);
    HIPCHK(hipGetLastError());
    if (loss_out) {
      HIPCHK(hipMemcpyAsync(part.data(), s->m_loss.p, part.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      double sum = 0.0;
      for (float x : part) sum += x;
      loss_out[it] = (float)(sum / (4.0 * (double)batch));
    }
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return PH_OK;
}

extern "C" int64_t pippin_hip_mlp_num_params(const ph_sim* s) {
  if (!s || !s->mlp_ready) return -1;
  return mlp_nparam(s->net.h);
}

extern "C" int pippin_hip_mlp_get_params(ph_sim* s, float* out, int64_t n) {
  if (int rc = mlp_check(s)) return rc;
  if (int rc = check_len("mlp_get_params", n, 1, mlp_nparam(s->net.h))) return rc;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(out, s->m_P.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return PH_OK;
}

// Replaces the parameters (padding entries must be 0) and restarts Adam.
extern "C" int pippin_hip_mlp_set_params(ph_sim* s, const float* in, int64_t n) {
  if (int rc = mlp_check(s)) return rc;
  const int h = s->net.h;
  if (int rc = check_len("mlp_set_params", n, 1, mlp_nparam(h))) return rc;
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(in[i])) return fail(PH_ERR_ARG, "mlp_set_params: non-finite parameter");
  for (int i = MLP_NF; i < MLP_KIN; i++)
    for (int o = 0; o < h; o++)
      if (in[mlp_woff(0, h) + (int64_t)i * h + o] != 0.0f) return fail(PH_ERR_ARG, "mlp_set_params: nonzero input padding");
  for (int i = 0; i < h; i++)
    for (int o = MLP_NO; o < MLP_NOP; o++)
      if (in[mlp_woff(3, h) + (int64_t)i * MLP_NOP + o] != 0.0f)
        return fail(PH_ERR_ARG, "mlp_set_params: nonzero output padding");
  for (int o = MLP_NO; o < MLP_NOP; o++)
    if (in[mlp_boff(3, h) + o] != 0.0f) return fail(PH_ERR_ARG, "mlp_set_params: nonzero output padding");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->m_P.p, in, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(s->m_m.zero());
  HIPCHK(s->m_v.zero());
  s->adam_t = 0;
  phm::to_bf16(s->net, s->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return PH_OK;
}

static int mlp_batch_in(ph_sim* s, const char* what, const float* x, const float* t, int64_t n) {
  if (n < 16 || n % 16 || n > s->batch.bmax)
    return fail(PH_ERR_ARG, std::string(what) + ": sample count must be a multiple of 16 up to max_batch");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->m_xin.p, x, n * MLP_NF * sizeof(float), hipMemcpyHostToDevice));
  if (t) HIPCHK(hipMemcpy(s->m_tin.p, t, n * MLP_NO * sizeof(float), hipMemcpyHostToDevice));
  return PH_OK;
}

extern "C" int pippin_hip_mlp_forward(ph_sim* s, const float* x, int64_t n, float* y) {
  if (int rc = mlp_check(s)) return rc;
  if (int rc = mlp_batch_in(s, "mlp_forward", x, nullptr, n)) return rc;
  phm::forward_backward(s->net, s->batch, s->m_xin.p, nullptr, n, n, true, 0, 0, false, s->stream);
  HIPCHK(hipGetLastError());
  std::vector<float> yp(n * MLP_NOP);
  HIPCHK(hipMemcpyAsync(yp.data(), s->m_y.p, yp.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int64_t i = 0; i < n; i++)
    for (int o = 0; o < MLP_NO; o++) y[i * MLP_NO + o] = yp[i * MLP_NOP + o];
  return PH_OK;
}

extern "C" int pippin_hip_mlp_loss_grad(ph_sim* s, const float* x, const float* t, int64_t n, float* grad,
                                        float* loss) {
  if (int rc = mlp_check(s)) return rc;
  if (int rc = mlp_batch_in(s, "mlp_loss_grad", x, t, n)) return rc;
  phm::forward_backward(s->net, s->batch, s->m_xin.p, s->m_tin.p, n, n, true, 0, 0, true, s->stream);
  HIPCHK(hipGetLastError());
  std::vector<float> part(n / 16);
  HIPCHK(hipMemcpyAsync(part.data(), s->m_loss.p, part.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  if (grad)
    HIPCHK(hipMemcpyAsync(grad, s->m_G.p, mlp_nparam(s->net.h) * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  double sum = 0.0;
  for (float v : part) sum += v;
  if (loss) *loss = (float)(sum / (4.0 * (double)n));
  return PH_OK;
}

extern "C" int pippin_hip_mlp_records(ph_sim* s, int64_t env, float* feat, float* pred, float* tgt, int32_t* count) {
  if (int rc = mlp_check(s)) return rc;
  if (env < 0 || env >= s->n) return fail(PH_ERR_ARG, "mlp_records: env out of range");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  int c = 0;
  HIPCHK(hipMemcpy(&c, s->m_cnt.p + env, sizeof(int), hipMemcpyDeviceToHost));
  if (count) *count = c;
  if (feat)
    HIPCHK(hipMemcpy(feat, s->m_feat.p + env * MLP_CL * MLP_NF, MLP_CL * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
  if (pred)
    HIPCHK(hipMemcpy(pred, s->m_pred.p + env * MLP_CL * MLP_NO, MLP_CL * MLP_NO * sizeof(float), hipMemcpyDeviceToHost));
  if (tgt)
    HIPCHK(hipMemcpy(tgt, s->m_tgt.p + env * MLP_CL * MLP_NO, MLP_CL * MLP_NO * sizeof(float), hipMemcpyDeviceToHost));
  return PH_OK;
}
