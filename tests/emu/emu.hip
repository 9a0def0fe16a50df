// Test-only serial emulation of the device kernels' per-env code.
//
// Runs the exact kernel body (pippin_amd/csrc/kbody.hpp) with a serial
// execution policy on the host, over host copies of the device model and
// state. The tests use it on machines without a GPU to check the kernels'
// indexing and arithmetic against the oracle before anything runs on the
// device. It is never loaded by the product library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <algorithm>
#include <cstdlib>
#include <limits>
#include <vector>

#include "../../include/pippin_hip.h"
#include "../../pippin_amd/csrc/host_model.hpp"
#include "../../pippin_amd/csrc/kbody.hpp"

using namespace phk;

struct HostX : HostLanes {
  template <class F>
  static void par(int n, F&& f) {
    for (int i = 0; i < n; i++) f(i);
  }
  static void sync() {}
};

// The network of the learned warm start for one sample, in FP32 with BF16
// rounding at the device's points (inputs, weights, hidden activations); the
// sums run sequentially from the bias, where the device's matrix cores
// accumulate in their own order.
static void host_mlp_forward(const MlpW& W, const float* f, float* y) {
  const int h = W.h;
  float x[MLP_HMAX], a[MLP_HMAX];
  for (int i = 0; i < MLP_KIN; i++) x[i] = i < MLP_NF ? bf2f(f2bf(f[i])) : 0.0f;
  int nin = MLP_KIN;
  for (int l = 0; l < 3; l++) {
    for (int o = 0; o < h; o++) {
      float acc = W.b[l][o];
      for (int i = 0; i < nin; i++) acc += bf2f(W.wt[l][(int64_t)o * nin + i]) * x[i];
      a[o] = bf2f(f2bf(mlp_gelu(acc)));
    }
    for (int o = 0; o < h; o++) x[o] = a[o];
    nin = h;
  }
  for (int o = 0; o < MLP_NO; o++) {
    float acc = W.b[3][o];
    for (int i = 0; i < h; i++) acc += bf2f(W.wt[3][(int64_t)o * h + i]) * x[i];
    y[o] = acc;
  }
}

// The policy of the learned-warm-start kernel with one env at a time: the
// network runs over this env's recorded contacts.
struct HostXM : HostX {
  template <class MT>
  static void mlp_tiles(const MT& M, const DState& S, int64_t e, int cnt) {
    for (int k = 0; k < cnt; k++)
      host_mlp_forward(M.mlp, S.mlp_feat.p + (e * MLP_CL + k) * MLP_NF, S.mlp_pred.p + (e * MLP_CL + k) * MLP_NO);
  }
};

struct emu_sim {
  phh::HostModel hm;
  DModel M;
  DState S;
  int64_t n;
  std::vector<int> env_var;
  std::vector<double> v_size, v_aabb, v_fric, v_mass, v_ipos, v_inertia, v_damp, v_gain, v_bias, v_diw, v_biw, v_qres;
  std::vector<uint32_t> v_ct, v_ca;
  std::vector<int> v_damped;
  std::vector<double> qpos, qvel, ctrl, qacc, applied, time, gxpos, gxmat, xpos, xmat, con, pcon, plambda, stage, gjac,
      grow, islf, gmj;
  std::vector<int> ncon, iters, nrows, status, cgeom, pcgeom, npcon, cost;
  std::vector<double> lds;
  // learned warm start
  std::vector<float> mlp_P, mlp_feat, mlp_pred, mlp_tgt;
  std::vector<uint16_t> mlp_wbf, mlp_wtbf;
  std::vector<int> mlp_cnt;
};

static thread_local std::string g_err;

extern "C" {

const char* emu_last_error(void) { return g_err.c_str(); }

emu_sim* emu_create(const ph_model_desc* desc, int64_t n) {
  emu_sim* s = new emu_sim();
  std::string err;
  if (!phh::build_host_model(*desc, s->hm, &err)) {
    g_err = err;
    delete s;
    return nullptr;
  }
  phh::HostModel& hm = s->hm;
  const ph_model_desc& d = hm.d;
  DModel& M = s->M;
  phh::init_dmodel_scalars(hm, d.solver_kind, 0, M);
  M.body_parent = hm.body_parent.data();
  M.body_jntadr = hm.body_jntadr.data();
  M.body_jntnum = hm.body_jntnum.data();
  M.body_dofadr = hm.body_dofadr.data();
  M.body_dofnum = hm.body_dofnum.data();
  M.body_lastdof = hm.body_lastdof.data();
  M.body_chain = hm.body_chain.data();
  M.body_pos = hm.body_pos.data();
  M.body_quat = hm.body_quat.data();
  M.level_adr = hm.level_adr.data();
  M.level_body = hm.level_body.data();
  M.acc_level_adr = hm.acc_level_adr.data();
  M.acc_parent = hm.acc_parent.data();
  M.acc_child_adr = hm.acc_child_adr.data();
  M.acc_child = hm.acc_child.data();
  M.jnt_type = hm.jnt_type.data();
  M.jnt_qposadr = hm.jnt_qposadr.data();
  M.jnt_dofadr = hm.jnt_dofadr.data();
  M.jnt_limited = hm.jnt_limited.data();
  M.jnt_pos = hm.jnt_pos.data();
  M.jnt_axis = hm.jnt_axis.data();
  M.jnt_range = hm.jnt_range.data();
  M.jnt_stiffness = hm.jnt_stiffness.data();
  M.jnt_springref = hm.jnt_springref.data();
  M.dof_body = hm.dof_body.data();
  M.dof_parent = hm.dof_parent.data();
  M.dof_jnt = hm.dof_jnt.data();
  M.dof_anc = hm.dof_anc.data();
  M.dof_armature = hm.dof_armature.data();
  M.block_adr = hm.block_adr.data();
  M.dof_block = hm.dof_block.data();
  M.geom_type = hm.geom_type.data();
  M.geom_body = hm.geom_body.data();
  M.geom_pos = hm.geom_pos.data();
  M.geom_quat = hm.geom_quat.data();
  M.pair_g1 = hm.pair_g1.data();
  M.pair_g2 = hm.pair_g2.data();
  M.act_gear = hm.actuator_gear.data();
  M.act_ctrllimited = hm.actuator_ctrllimited.data();
  M.act_forcelimited = hm.actuator_forcelimited.data();
  M.act_ctrlrange = hm.actuator_ctrlrange.data();
  M.act_forcerange = hm.actuator_forcerange.data();
  M.act_madr = hm.actuator_moment_adr.data();
  M.act_mnum = hm.actuator_moment_num.data();
  M.mom_jnt = hm.moment_jnt.data();
  M.mom_coef = hm.moment_coef.data();
  M.dofact_adr = hm.dofact_adr.data();
  M.dofact_num = hm.dofact_num.data();
  M.dofact_u = hm.dofact_u.data();
  M.dofact_coef = hm.dofact_coef.data();
  M.eq_j1 = hm.eq_joint1.data();
  M.eq_j2 = hm.eq_joint2.data();
  M.eq_poly = hm.eq_polycoef.data();
  M.eq_solref = hm.eq_solref.data();
  M.eq_solimp = hm.eq_solimp.data();
  M.qpos0 = hm.qpos0.data();
  M.geom_dataid = hm.geom_dataid.data();
  M.mesh_vertadr = hm.mesh_vertadr.data();
  M.mesh_vertnum = hm.mesh_vertnum.data();
  M.mesh_vert = hm.mesh_vert.data();
  M.face_normal = hm.face_normal.data();
  M.face_vertadr = hm.face_vertadr.data();
  M.face_vertnum = hm.face_vertnum.data();
  M.face_vert = hm.face_vert.data();
  M.vert_faceadr = hm.vert_faceadr.data();
  M.vert_facenum = hm.vert_facenum.data();
  M.vert_face = hm.vert_face.data();
  s->env_var.assign(n, 0);
  M.env_var = s->env_var.data();
  s->v_size = hm.geom_size;
  s->v_aabb = hm.geom_aabb;
  s->v_fric = hm.geom_friction;
  s->v_ct = hm.geom_contype;
  s->v_ca = hm.geom_conaffinity;
  s->v_mass = hm.body_mass;
  s->v_ipos = hm.body_ipos;
  s->v_inertia = hm.body_inertia;
  s->v_damp = hm.dof_damping;
  s->v_gain = hm.actuator_gain;
  s->v_bias = hm.actuator_bias;
  s->v_diw.assign(d.nv + 1, 0.0);
  s->v_biw.assign(2 * d.nbody, 0.0);
  s->v_qres = hm.qpos_reset;
  s->v_damped.assign(1, hm.damped ? 1 : 0);
  M.v_geom_size = s->v_size.data();
  M.v_geom_aabb = s->v_aabb.data();
  M.v_friction = s->v_fric.data();
  M.v_contype = s->v_ct.data();
  M.v_conaffinity = s->v_ca.data();
  M.v_body_mass = s->v_mass.data();
  M.v_body_ipos = s->v_ipos.data();
  M.v_body_inertia = s->v_inertia.data();
  M.v_dof_damping = s->v_damp.data();
  M.v_act_gain = s->v_gain.data();
  M.v_act_bias = s->v_bias.data();
  M.v_dof_invweight = s->v_diw.data();
  M.v_body_invweight = s->v_biw.data();
  M.v_qpos_reset = s->v_qres.data();
  M.v_damped = s->v_damped.data();

  s->n = n;
  int64_t nn = n > 0 ? n : 1;
  s->qpos.resize(nn * d.nq);
  for (int64_t e = 0; e < nn; e++)
    for (int k = 0; k < d.nq; k++) s->qpos[e * d.nq + k] = hm.qpos_reset[k];
  s->qvel.assign(nn * d.nv, 0.0);
  s->ctrl.assign(nn * d.nu + 1, 0.0);
  s->qacc.assign(nn * d.nv, 0.0);
  s->applied.assign(nn * d.nv, 0.0);
  s->time.assign(nn, 0.0);
  s->gxpos.assign(nn * 3 * d.ngeom, 0.0);
  s->gxmat.assign(nn * 9 * d.ngeom, 0.0);
  s->xpos.assign(nn * 3 * d.nbody, 0.0);
  s->xmat.assign(nn * 9 * d.nbody, 0.0);
  s->con.assign(nn * M.ncmax * CW, 0.0);
  s->cgeom.assign(nn * M.ncmax * 3, 0);
  s->pcon.assign(nn * M.ncmax * CW, 0.0);
  s->pcgeom.assign(nn * M.ncmax * 3, 0);
  s->plambda.assign(nn * M.ncmax * 3, 0.0);
  s->stage.assign(nn * (M.npair > 0 ? M.npair : 1) * MAXHIT * CW, 0.0);
  s->gjac.assign(nn * (int64_t)M.nrmax * d.nv, 0.0);
  s->grow.assign(nn * (int64_t)M.nrmax * 10, 0.0);
  s->islf.assign(std::max<int64_t>(nn * (int64_t)M.isl_fstride, 1), 0.0);
  s->gmj.assign(M.solver_kind == 1 ? nn * (int64_t)M.nrmax * d.nv : 1, 0.0);
  // EMU_POISON_SCRATCH=1 fills the buffers the device library leaves
  // uninitialised at create (contacts, staging, row spill, island scratch)
  // with NaN / -1, so that a read before write differs from the oracle
  if (const char* ps = std::getenv("EMU_POISON_SCRATCH")) {
    // "1": NaN and -1; any other number x: doubles x, ints (int)x
    const double x = std::atof(ps);
    const double dv = std::string(ps) == "1" ? std::numeric_limits<double>::quiet_NaN() : x;
    const int iv = std::string(ps) == "1" ? -1 : (int)x;
    for (auto* v : {&s->gxpos, &s->gxmat, &s->xpos, &s->xmat, &s->con, &s->pcon, &s->plambda, &s->stage, &s->gjac,
                    &s->grow, &s->islf, &s->gmj})
      std::fill(v->begin(), v->end(), dv);
    for (auto* v : {&s->cgeom, &s->pcgeom}) std::fill(v->begin(), v->end(), iv);
  }
  s->ncon.assign(nn, 0);
  s->iters.assign(nn, 0);
  s->nrows.assign(nn, 0);
  s->status.assign(nn, 0);
  s->cost.assign(nn, 0);
  s->npcon.assign(nn, 0);
  DState& S = s->S;
  S.n = n;
  S.qpos = s->qpos.data();
  S.qvel = s->qvel.data();
  S.ctrl = s->ctrl.data();
  S.qacc = s->qacc.data();
  S.applied = s->applied.data();
  S.time = s->time.data();
  S.geom_xpos = s->gxpos.data();
  S.geom_xmat = s->gxmat.data();
  S.xpos = s->xpos.data();
  S.xmat = s->xmat.data();
  S.ncon = s->ncon.data();
  S.iters = s->iters.data();
  S.nrows = s->nrows.data();
  S.status = s->status.data();
  S.cost = s->cost.data();
  S.con = s->con.data();
  S.cgeom = s->cgeom.data();
  S.pcon = s->pcon.data();
  S.pcgeom = s->pcgeom.data();
  S.plambda = s->plambda.data();
  S.npcon = s->npcon.data();
  S.stage = s->stage.data();
  S.gjac = s->gjac.data();
  S.grow = s->grow.data();
  S.isl_f = s->islf.data();
  S.gmj = s->gmj.data();
  s->lds.assign(M.L.bytes / sizeof(double) + 16, 0.0);
  if (std::getenv("EMU_POISON_LDS")) std::fill(s->lds.begin(), s->lds.end(), std::numeric_limits<double>::quiet_NaN());
  // set_const for the base variant
  Env E;
  E.s = s->lds.data();
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  E.e = 0;
  E.var = 0;
  for (int k = 0; k < d.nq; k++) E.s[M.L.qpos + k] = hm.qpos_reset[k];
  set_const_env<HostX>(M, E, s->v_diw.data(), s->v_biw.data());
  for (int64_t e = 0; e < n; e++) {
    E.e = e;
    load_state<HostX>(M, S, E);
    kinematics<HostX>(M, E);
    store_poses<HostX>(M, S, E);
  }
  return s;
}

void emu_destroy(emu_sim* s) { delete s; }

int emu_step(emu_sim* s, int64_t begin, int64_t end, int64_t nstep) {
  Env E;
  E.s = s->lds.data();
  E.si = reinterpret_cast<int*>(E.s + s->M.L.ndouble);
  // EMU_POISON_LDS=1 fills the working set with NaN before each env, so that a
  // read of LDS the step has not written shows up as a difference from the oracle
  const bool poison = std::getenv("EMU_POISON_LDS") != nullptr;
  for (int64_t e = begin; e < end; e++) {
    E.e = e;
    E.var = s->env_var[e];
    if (poison) std::fill(s->lds.begin(), s->lds.end(), std::numeric_limits<double>::quiet_NaN());
    if (s->M.mlp.mode)
      run_env_mlp<HostXM>(s->M, s->S, E, nstep, nullptr, nullptr, e - begin, true);
    else
      run_env<HostX>(s->M, s->S, E, nstep, nullptr, nullptr, e - begin);
  }
  return 0;
}

// Learned warm start: parameters in the device layout (mlp.hpp), hidden width h, mode.
int emu_mlp_set(emu_sim* s, int h, int mode, const float* P) {
  const int64_t np = mlp_nparam(h), nw = mlp_nweight(h), n = std::max<int64_t>(s->n, 1);
  s->mlp_P.assign(P, P + np);
  s->mlp_wbf.assign(nw, 0);
  s->mlp_wtbf.assign(nw, 0);
  for (int l = 0; l < 4; l++) {
    const int IN = mlp_in(l, h), OUT = mlp_out(l, h);
    for (int i = 0; i < IN; i++)
      for (int o = 0; o < OUT; o++) {
        const uint16_t b = f2bf(P[mlp_woff(l, h) + (int64_t)i * OUT + o]);
        s->mlp_wbf[mlp_wbf_off(l, h) + (int64_t)i * OUT + o] = b;
        s->mlp_wtbf[mlp_wbf_off(l, h) + (int64_t)o * IN + i] = b;
      }
  }
  if (s->mlp_feat.empty()) {
    s->mlp_feat.assign(n * MLP_CL * MLP_NF, 0.0f);
    s->mlp_pred.assign(n * MLP_CL * MLP_NO, 0.0f);
    s->mlp_tgt.assign(n * MLP_CL * MLP_NO, 0.0f);
    s->mlp_cnt.assign(n, 0);
  }
  MlpW& W = s->M.mlp;
  W.mode = mode;
  W.h = h;
  for (int l = 0; l < 4; l++) {
    W.wt[l] = s->mlp_wtbf.data() + mlp_wbf_off(l, h);
    W.w[l] = s->mlp_wbf.data() + mlp_wbf_off(l, h);
    W.b[l] = s->mlp_P.data() + mlp_boff(l, h);
  }
  s->S.mlp_feat = s->mlp_feat.data();
  s->S.mlp_pred = s->mlp_pred.data();
  s->S.mlp_tgt = s->mlp_tgt.data();
  s->S.mlp_cnt = s->mlp_cnt.data();
  return 0;
}

int emu_mlp_records(emu_sim* s, int64_t env, float* feat, float* pred, float* tgt, int32_t* count) {
  *count = s->mlp_cnt[env];
  memcpy(feat, s->mlp_feat.data() + env * MLP_CL * MLP_NF, MLP_CL * MLP_NF * sizeof(float));
  memcpy(pred, s->mlp_pred.data() + env * MLP_CL * MLP_NO, MLP_CL * MLP_NO * sizeof(float));
  memcpy(tgt, s->mlp_tgt.data() + env * MLP_CL * MLP_NO, MLP_CL * MLP_NO * sizeof(float));
  return 0;
}

int emu_get(emu_sim* s, int field, int64_t begin, int64_t end, double* out) {
  const ph_model_desc& d = s->hm.d;
  int w = field == 0 ? d.nq : field == 1 ? d.nv : d.nu;
  const std::vector<double>& v = field == 0 ? s->qpos : field == 1 ? s->qvel : s->ctrl;
  memcpy(out, v.data() + begin * w, (end - begin) * w * sizeof(double));
  return 0;
}

int emu_set(emu_sim* s, int field, int64_t begin, int64_t end, const double* src) {
  const ph_model_desc& d = s->hm.d;
  int w = field == 0 ? d.nq : field == 1 ? d.nv : d.nu;
  std::vector<double>& v = field == 0 ? s->qpos : field == 1 ? s->qvel : s->ctrl;
  memcpy(v.data() + begin * w, src, (end - begin) * w * sizeof(double));
  return 0;
}

int emu_ncon(emu_sim* s, int64_t begin, int64_t end, int32_t* out) {
  for (int64_t e = begin; e < end; e++) out[e - begin] = s->ncon[e];
  return 0;
}

int emu_iters(emu_sim* s, int64_t begin, int64_t end, int32_t* out) {
  for (int64_t e = begin; e < end; e++) out[e - begin] = s->iters[e];
  return 0;
}

int emu_status(emu_sim* s, int64_t begin, int64_t end, int32_t* out) {
  for (int64_t e = begin; e < end; e++) out[e - begin] = s->status[e];
  return 0;
}

int emu_contacts(emu_sim* s, int64_t env, ph_contact* out, int32_t cap, int32_t* n) {
  int nc = s->ncon[env];
  *n = nc;
  for (int i = 0; i < nc && i < cap; i++) {
    const double* c = &s->con[(env * s->M.ncmax + i) * CW];
    for (int k = 0; k < 3; k++) {
      out[i].pos[k] = c[k];
      out[i].normal[k] = c[3 + k];
      out[i].tangent[k] = c[6 + k];
    }
    out[i].depth = c[9];
    const int* g = &s->cgeom[(env * s->M.ncmax + i) * 3];
    out[i].geom1 = g[0];
    out[i].geom2 = g[1];
    out[i].feature = g[2];
    out[i].pad = 0;
  }
  return 0;
}

int emu_mass_bias(emu_sim* s, int64_t begin, int64_t end, double* qm, double* bias) {
  const DModel& M = s->M;
  Env E;
  E.s = s->lds.data();
  E.si = reinterpret_cast<int*>(E.s + M.L.ndouble);
  for (int64_t e = begin; e < end; e++) {
    E.e = e;
    E.var = s->env_var[e];
    load_state<HostX>(M, s->S, E);
    kinematics<HostX>(M, E);
    crba<HostX>(M, E);
    velocity_rne<HostX>(M, E);
    memcpy(qm + (e - begin) * M.nv * M.nv, E.s + M.L.qm, M.nv * M.nv * sizeof(double));
    memcpy(bias + (e - begin) * M.nv, E.s + M.L.bias, M.nv * sizeof(double));
  }
  return 0;
}

int emu_invweight(emu_sim* s, double* dof, double* body) {
  memcpy(dof, s->v_diw.data(), s->M.nv * sizeof(double));
  memcpy(body, s->v_biw.data(), 2 * s->M.nbody * sizeof(double));
  return 0;
}

int emu_poses(emu_sim* s, int64_t begin, int64_t end, double* gxpos, double* gxmat) {
  int ng = s->M.ngeom;
  for (int64_t e = begin; e < end; e++) {  // poses on demand, as the device library computes them
    Env E;
    E.s = s->lds.data();
    E.si = reinterpret_cast<int*>(E.s + s->M.L.ndouble);
    E.e = e;
    E.var = s->env_var[e];
    load_state<HostX>(s->M, s->S, E);
    kinematics<HostX>(s->M, E);
    store_poses<HostX>(s->M, s->S, E);
  }
  memcpy(gxpos, s->gxpos.data() + begin * 3 * ng, (end - begin) * 3 * ng * sizeof(double));
  memcpy(gxmat, s->gxmat.data() + begin * 9 * ng, (end - begin) * 9 * ng * sizeof(double));
  return 0;
}

}  // extern "C"
