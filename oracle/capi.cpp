// ORACLE — test infrastructure only (see math.hpp).
// C entry points the tests (ctypes) and the CPU-baseline leg of bench.py call.
// They restate PippinCpu (crates/pippin-env/src/pippin_cpu.rs) and Batch
// (crates/pippin/src/batch.rs) over the restated engine, plus known-answer
// hooks into the math and narrow phase.
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>

#include "../include/pippin_hip.h"
#include "engine.hpp"

using namespace orc;

struct orc_sim {
  Model base;
  std::vector<std::shared_ptr<Model>> models;  // copy-on-write per env, pippin_cpu.rs:14-21
  std::vector<Data> envs;
  std::string err;
  bool failed = false;  // a non-PD mass matrix (the reference panics)
};

static thread_local std::string g_err;

// Parallel-for over env chunks of `chunk` envs with `nthreads` workers that
// take chunks from a shared counter (rayon's par_chunks with work stealing).
template <class F>
static void par_chunks(size_t begin, size_t end, size_t chunk, int nthreads, F f) {
  size_t nchunks = (end - begin + chunk - 1) / chunk;
  std::atomic<size_t> next(0);
  auto worker = [&]() {
    for (;;) {
      size_t c = next.fetch_add(1);
      if (c >= nchunks) return;
      size_t b = begin + c * chunk, e = b + chunk < end ? b + chunk : end;
      f(b, e);
    }
  };
  if (nthreads <= 1) {
    worker();
    return;
  }
  std::vector<std::thread> ts;
  for (int t = 0; t < nthreads; t++) ts.emplace_back(worker);
  for (auto& t : ts) t.join();
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// pippin_cpu.rs:267-273
orc_sim* orc_create(const ph_model_desc* desc, int64_t n) {
  orc_sim* s = new orc_sim();
  s->base = model_from_desc(*desc);
  Data proto = data_new(s->base);
  kinematics(s->base, proto);
  auto shared = std::make_shared<Model>(s->base);
  s->models.assign((size_t)n, shared);
  s->envs.assign((size_t)n, proto);
  return s;
}

void orc_destroy(orc_sim* s) { delete s; }

int64_t orc_num_envs(orc_sim* s) { return (int64_t)s->envs.size(); }

// pippin_cpu.rs:98-112: step-synchronous over the range, then kinematics.
int orc_step(orc_sim* s, int64_t begin, int64_t end, int64_t nstep, int nthreads) {
  if (begin < 0 || end > (int64_t)s->envs.size() || begin > end) return PH_ERR_ARG;
  size_t len = (size_t)(end - begin);
  if (len == 0) return PH_OK;
  int T = nthreads > 0 ? nthreads : 1;
  size_t chunk = len / (4 * (size_t)T);
  if (chunk < 1) chunk = 1;
  std::atomic<bool> bad(false);
  par_chunks((size_t)begin, (size_t)end, chunk, T, [&](size_t b, size_t e) {
    for (int64_t t = 0; t < nstep; t++)
      for (size_t i = b; i < e; i++)
        if (!step(*s->models[i], s->envs[i])) bad = true;
    for (size_t i = b; i < e; i++) kinematics(*s->models[i], s->envs[i]);
  });
  if (bad) {
    g_err = "mass matrix is not positive definite";
    return PH_ERR_NOTPD;
  }
  return PH_OK;
}

// batch.rs:46-59: each env runs all its steps; ctrl (n, nstep, nu), qpos_out (n, nstep, nq).
int orc_rollout(orc_sim* s, int64_t begin, int64_t end, int64_t nstep, const double* ctrl, double* qpos_out,
                int nthreads) {
  size_t nu = s->base.nu, nq = s->base.nq;
  std::atomic<bool> bad(false);
  int T = nthreads > 0 ? nthreads : 1;
  par_chunks((size_t)begin, (size_t)end, 1, T, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; i++) {
      size_t k = i - (size_t)begin;
      Data& d = s->envs[i];
      for (int64_t t = 0; t < nstep; t++) {
        if (nu) std::memcpy(d.ctrl.data(), ctrl + (k * nstep + t) * nu, nu * sizeof(double));
        if (!step(*s->models[i], d)) bad = true;
        if (qpos_out) std::memcpy(qpos_out + (k * nstep + t) * nq, d.qpos.data(), nq * sizeof(double));
      }
    }
  });
  return bad ? PH_ERR_NOTPD : PH_OK;
}

// pippin_cpu.rs:357-363
int orc_reset(orc_sim* s, const int64_t* ids, int64_t n) {
  for (int64_t k = 0; k < n; k++) {
    size_t i = (size_t)ids[k];
    data_reset(s->envs[i], *s->models[i]);
    kinematics(*s->models[i], s->envs[i]);
  }
  return PH_OK;
}

// Data fields by PH_FIELD_* (data.rs:8-80); time is a one-element view.
// Returns false for an unknown field.
static bool field_ptr(Data& d, int f, Real** p, size_t* width) {
  std::vector<Real>* v = nullptr;
  switch (f) {
    case PH_FIELD_QPOS: v = &d.qpos; break;
    case PH_FIELD_QVEL: v = &d.qvel; break;
    case PH_FIELD_CTRL: v = &d.ctrl; break;
    case PH_FIELD_QACC: v = &d.qacc; break;
    case PH_FIELD_QFRC_APPLIED: v = &d.qfrc_applied; break;
    case PH_FIELD_TIME: *width = 1; *p = &d.time; return true;
    default: return false;
  }
  *width = v->size();
  *p = v->data();
  return true;
}

int orc_get(orc_sim* s, int field, int64_t begin, int64_t end, double* out) {
  for (int64_t i = begin; i < end; i++) {
    size_t w = 0;
    Real* v = nullptr;
    if (!field_ptr(s->envs[(size_t)i], field, &v, &w)) return PH_ERR_ARG;
    if (w) std::memcpy(out + (i - begin) * w, v, w * sizeof(double));
  }
  return PH_OK;
}

// pippin_cpu.rs:413-430: setting qpos reruns kinematics.
int orc_set(orc_sim* s, int field, int64_t begin, int64_t end, const double* src) {
  for (int64_t i = begin; i < end; i++) {
    Data& d = s->envs[(size_t)i];
    size_t w = 0;
    Real* v = nullptr;
    if (!field_ptr(d, field, &v, &w)) return PH_ERR_ARG;
    if (w) std::memcpy(v, src + (i - begin) * w, w * sizeof(double));
    if (field == PH_FIELD_QPOS) kinematics(*s->models[(size_t)i], d);
  }
  return PH_OK;
}

// Scratch::prev_contacts / prev_lambda (data.rs:60-80) of envs [begin, end):
// counts (n), contacts (n x cap), lambda (n x cap x 3). Returns PH_ERR_SIZE
// when an env holds more than cap.
int orc_get_history(orc_sim* s, int64_t begin, int64_t end, int32_t cap, int32_t* counts, ph_contact* out,
                    double* lambda) {
  for (int64_t i = begin; i < end; i++) {
    const Data& d = s->envs[(size_t)i];
    if ((int64_t)d.prev_contacts.size() > cap) return PH_ERR_SIZE;
    counts[i - begin] = (int32_t)d.prev_contacts.size();
    for (int32_t k = 0; k < cap; k++) {
      ph_contact& o = out[(i - begin) * cap + k];
      double* lo = lambda + ((i - begin) * cap + k) * 3;
      o = ph_contact{};
      lo[0] = lo[1] = lo[2] = 0.0;
      if (k >= counts[i - begin]) continue;
      const Contact& c = d.prev_contacts[(size_t)k];
      for (int t = 0; t < 3; t++) {
        o.pos[t] = c.pos[t];
        o.normal[t] = c.normal[t];
        o.tangent[t] = c.tangent[t];
        lo[t] = d.prev_lambda[(size_t)k][(size_t)t];
      }
      o.depth = c.depth;
      o.geom1 = (int32_t)c.geom[0];
      o.geom2 = (int32_t)c.geom[1];
      o.feature = (int32_t)c.feature;
    }
  }
  return PH_OK;
}

int orc_set_history(orc_sim* s, int64_t begin, int64_t end, int32_t cap, const int32_t* counts, const ph_contact* in,
                    const double* lambda) {
  for (int64_t i = begin; i < end; i++)
    if (counts[i - begin] < 0 || counts[i - begin] > cap) return PH_ERR_ARG;
  for (int64_t i = begin; i < end; i++) {
    Data& d = s->envs[(size_t)i];
    d.prev_contacts.clear();
    d.prev_lambda.clear();
    for (int32_t k = 0; k < counts[i - begin]; k++) {
      const ph_contact& c = in[(i - begin) * cap + k];
      Contact o;
      for (int t = 0; t < 3; t++) {
        o.pos[t] = c.pos[t];
        o.normal[t] = c.normal[t];
        o.tangent[t] = c.tangent[t];
      }
      o.depth = c.depth;
      o.geom[0] = (size_t)c.geom1;
      o.geom[1] = (size_t)c.geom2;
      o.feature = (uint32_t)c.feature;
      d.prev_contacts.push_back(o);
      const double* lk = lambda + ((i - begin) * cap + k) * 3;
      d.prev_lambda.push_back({lk[0], lk[1], lk[2]});
    }
  }
  return PH_OK;
}

int orc_set_applied(orc_sim* s, int64_t env, const double* f) {
  Data& d = s->envs[(size_t)env];
  std::memcpy(d.qfrc_applied.data(), f, d.qfrc_applied.size() * sizeof(double));
  return PH_OK;
}

// pippin_cpu.rs:432-448
int orc_poses(orc_sim* s, int64_t begin, int64_t end, ph_pose* geoms, ph_pose* cams) {
  const Model& m = s->base;
  size_t ng = m.ngeom(), nc = m.cam_body.size();
  auto pose = [](Vec3 p, const Mat3& r, ph_pose* o) {
    for (int k = 0; k < 3; k++) o->pos[k] = (float)p[k];
    for (int k = 0; k < 9; k++) o->mat[k] = (float)r[k];
  };
  for (int64_t i = begin; i < end; i++) {
    const Data& d = s->envs[(size_t)i];
    size_t k = (size_t)(i - begin);
    for (size_t g = 0; g < ng; g++) pose(d.geom_xpos[g], d.geom_xmat[g], geoms + k * ng + g);
    for (size_t c = 0; c < nc; c++) {
      size_t b = m.cam_body[c];
      Vec3 pos = d.xpos[b] + mul_vec(d.xmat[b], m.cam_pos[c]);
      Mat3 mat = mul_mat(d.xmat[b], to_mat(m.cam_quat[c]));
      pose(pos, mat, cams + k * nc + c);
    }
  }
  return PH_OK;
}

// pippin_cpu.rs:365-374: validate on a clone of the base model, then per env.
int orc_set_param(orc_sim* s, int p, int64_t id, int64_t begin, int64_t end, const double* values) {
  size_t w = (size_t)param_width(p, s->base);
  Model probe = s->base;
  std::string err;
  if (id < 0 || !set_param(probe, p, (size_t)id, values, w, &err)) {
    g_err = id < 0 ? "negative id" : err;
    return PH_ERR_ARG;
  }
  for (int64_t i = begin; i < end; i++) {
    auto& mp = s->models[(size_t)i];
    if (mp.use_count() > 1) mp = std::make_shared<Model>(*mp);
    if (!set_param(*mp, p, (size_t)id, values + (i - begin) * w, w, &err)) {
      g_err = err;
      return PH_ERR_ARG;
    }
  }
  return PH_OK;
}

int orc_get_param(orc_sim* s, int p, int64_t id, int64_t begin, int64_t end, double* out) {
  size_t w = (size_t)param_width(p, s->base);
  if (id < 0 || (size_t)id >= param_count(p, s->base)) {
    g_err = "id out of range";
    return PH_ERR_ARG;
  }
  for (int64_t i = begin; i < end; i++) {
    std::vector<Real> v = get_param(*s->models[(size_t)i], p, (size_t)id);
    std::memcpy(out + (i - begin) * w, v.data(), w * sizeof(double));
  }
  return PH_OK;
}

// pippin_cpu.rs:388-400
int orc_appearance(orc_sim* s, int64_t begin, int64_t end, ph_appearance* out) {
  size_t ng = s->base.ngeom();
  for (int64_t i = begin; i < end; i++) {
    const Model& m = *s->models[(size_t)i];
    for (size_t g = 0; g < ng; g++) {
      ph_appearance& a = out[(i - begin) * ng + g];
      for (int k = 0; k < 4; k++) a.rgba[k] = m.geom_rgba[4 * g + k];
      for (int k = 0; k < 3; k++) {
        Real base = s->base.geom_size[g][k];
        a.scale[k] = base > 0.0 ? (float)(m.geom_size[g][k] / base) : 1.0f;
      }
      a.pad = 0.0f;
    }
  }
  return PH_OK;
}

// ---- diagnostics ----
int orc_ncon(orc_sim* s, int64_t begin, int64_t end, int32_t* out) {
  for (int64_t i = begin; i < end; i++) out[i - begin] = (int32_t)s->envs[(size_t)i].contacts.size();
  return PH_OK;
}

int orc_contacts(orc_sim* s, int64_t env, ph_contact* out, int32_t cap, int32_t* n) {
  const auto& cs = s->envs[(size_t)env].contacts;
  *n = (int32_t)cs.size();
  for (size_t i = 0; i < cs.size() && (int32_t)i < cap; i++) {
    for (int k = 0; k < 3; k++) {
      out[i].pos[k] = cs[i].pos[k];
      out[i].normal[k] = cs[i].normal[k];
      out[i].tangent[k] = cs[i].tangent[k];
    }
    out[i].depth = cs[i].depth;
    out[i].geom1 = (int32_t)cs[i].geom[0];
    out[i].geom2 = (int32_t)cs[i].geom[1];
    out[i].feature = (int32_t)cs[i].feature;
    out[i].pad = 0;
  }
  return PH_OK;
}

// Per env: newton iterations, rows of the last solve, line-search steps of the last step.
int orc_step_stats(orc_sim* s, int64_t begin, int64_t end, int32_t* iters, int32_t* nrows, int32_t* ls) {
  for (int64_t i = begin; i < end; i++) {
    const Data& d = s->envs[(size_t)i];
    if (iters) iters[i - begin] = (int32_t)d.newton_iters;
    if (nrows) nrows[i - begin] = (int32_t)d.nrows;
    if (ls) ls[i - begin] = (int32_t)d.ls_steps;
  }
  return PH_OK;
}

// kinematics + crba + velocity + rne at the current state; M (nv x nv) and qfrc_bias (nv).
int orc_mass_bias(orc_sim* s, int64_t begin, int64_t end, double* qm, double* bias) {
  size_t nv = s->base.nv;
  for (int64_t i = begin; i < end; i++) {
    const Model& m = *s->models[(size_t)i];
    Data& d = s->envs[(size_t)i];
    kinematics(m, d);
    crba(m, d);
    velocity(m, d);
    rne(m, d);
    if (qm) std::memcpy(qm + (i - begin) * nv * nv, d.qm.data(), nv * nv * sizeof(double));
    if (bias) std::memcpy(bias + (i - begin) * nv, d.qfrc_bias.data(), nv * sizeof(double));
  }
  return PH_OK;
}

int orc_invweight(orc_sim* s, int64_t begin, int64_t end, double* dof, double* body) {
  size_t nv = s->base.nv, nb = s->base.nbody();
  for (int64_t i = begin; i < end; i++) {
    const Model& m = *s->models[(size_t)i];
    if (dof) std::memcpy(dof + (i - begin) * nv, m.dof_invweight.data(), nv * sizeof(double));
    if (body)
      for (size_t b = 0; b < nb; b++) {
        body[((i - begin) * nb + b) * 2] = m.body_invweight[b].first;
        body[((i - begin) * nb + b) * 2 + 1] = m.body_invweight[b].second;
      }
  }
  return PH_OK;
}

int orc_collision_pairs(orc_sim* s, int64_t env, int32_t* pairs, int32_t cap, int32_t* n) {
  const auto& ps = s->models[(size_t)env]->collision_pairs;
  *n = (int32_t)ps.size();
  for (size_t i = 0; i < ps.size() && (int32_t)i < cap; i++) {
    pairs[2 * i] = (int32_t)ps[i].first;
    pairs[2 * i + 1] = (int32_t)ps[i].second;
  }
  return PH_OK;
}

// Compiled-model facts the tests check against closed forms.
int orc_model_info(orc_sim* s, int64_t env, double* body_mass, double* body_ipos, double* body_inertia,
                   int32_t* dof_parent, int32_t* body_weldid, int32_t* blocks, int32_t* nblocks) {
  const Model& m = *s->models[(size_t)env];
  for (size_t b = 0; b < m.nbody(); b++) {
    if (body_mass) body_mass[b] = m.body_mass[b];
    for (int k = 0; k < 3 && body_ipos; k++) body_ipos[3 * b + k] = m.body_ipos[b][k];
    for (int k = 0; k < 9 && body_inertia; k++) body_inertia[9 * b + k] = m.body_inertia[b][k];
    if (body_weldid) body_weldid[b] = (int32_t)m.body_weldid[b];
  }
  for (size_t k = 0; k < m.nv && dof_parent; k++)
    dof_parent[k] = m.dof_parent[k] == NONE ? -1 : (int32_t)m.dof_parent[k];
  if (nblocks) *nblocks = (int32_t)m.dof_blocks.size();
  for (size_t b = 0; b < m.dof_blocks.size() && blocks; b++) {
    blocks[2 * b] = (int32_t)m.dof_blocks[b].start;
    blocks[2 * b + 1] = (int32_t)m.dof_blocks[b].end;
  }
  return PH_OK;
}

// ---- known-answer hooks ----
void orc_cr_sincos(const double* x, int64_t n, double* s, double* c) {
  for (int64_t i = 0; i < n; i++) cr_sincos(x[i], s + i, c + i);
}
double orc_impedance(const double* solimp, double pos) { return impedance(solimp, pos); }
double orc_dot(const double* a, const double* b, int64_t n) { return sdot_n(a, b, (size_t)n); }
int orc_cholesky_block(double* a, int64_t stride, int64_t s0, int64_t e) {
  return cholesky_block(a, (size_t)stride, (size_t)s0, (size_t)e) ? 1 : 0;
}
void orc_cholesky_solve_block(const double* l, int64_t stride, int64_t s0, int64_t e, double* x) {
  cholesky_solve_block(l, (size_t)stride, (size_t)s0, (size_t)e, x);
}
void orc_quat_from_axis_angle(const double* axis, double angle, double* q) {
  Quat r = from_axis_angle(V3(axis[0], axis[1], axis[2]), angle);
  for (int k = 0; k < 4; k++) q[k] = r[k];
}
void orc_quat_to_mat(const double* q, double* m) {
  Mat3 r = to_mat(Q4(q[0], q[1], q[2], q[3]));
  for (int k = 0; k < 9; k++) m[k] = r[k];
}
void orc_quat_from_mat(const double* m, double* q) {
  Mat3 r;
  for (int k = 0; k < 9; k++) r[k] = m[k];
  Quat o = from_mat(r);
  for (int k = 0; k < 4; k++) q[k] = o[k];
}
void orc_quat_integrate(const double* q, const double* w, double h, double* out) {
  Quat r = integrate_local(Q4(q[0], q[1], q[2], q[3]), V3(w[0], w[1], w[2]), h);
  for (int k = 0; k < 4; k++) out[k] = r[k];
}
// Narrow phase between two posed primitives; hits as ph_contact (geom ids = -1).
int orc_collide(int ta, const double* pa, const double* ma, const double* sa, int tb, const double* pb,
                const double* mb, const double* sb, double margin, ph_contact* out, int32_t cap) {
  GeomPose a, b;
  a.typ = ta;
  b.typ = tb;
  a.hull = b.hull = nullptr;
  for (int k = 0; k < 3; k++) {
    a.pos[k] = pa[k];
    a.size[k] = sa[k];
    b.pos[k] = pb[k];
    b.size[k] = sb[k];
  }
  for (int k = 0; k < 9; k++) {
    a.mat[k] = ma[k];
    b.mat[k] = mb[k];
  }
  std::vector<Hit> hits;
  if (needs_general(ta, tb)) collide_general(a, b, margin, hits);
  else if (pair_supported(ta, tb)) collide_pair(a, b, margin, hits);
  for (size_t i = 0; i < hits.size() && (int32_t)i < cap; i++) {
    for (int k = 0; k < 3; k++) {
      out[i].pos[k] = hits[i].pos[k];
      out[i].normal[k] = hits[i].normal[k];
      out[i].tangent[k] = hits[i].tangent[k];
    }
    out[i].depth = hits[i].depth;
    out[i].geom1 = out[i].geom2 = -1;
    out[i].feature = (int32_t)i;
    out[i].pad = 0;
  }
  return (int)hits.size();
}

}  // extern "C"
