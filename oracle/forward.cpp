// ORACLE — test infrastructure only (see math.hpp).
// Smooth dynamics and time stepping: crates/pippin/src/forward.rs.
#include "engine.hpp"

namespace orc {

// forward.rs:14-90
void kinematics(const Model& m, Data& d) {
  d.xpos[0] = v3zero();
  d.xquat[0] = qidentity();
  d.xmat[0] = m3identity();
  d.xipos[0] = v3zero();
  for (size_t b = 1; b < m.nbody(); b++) {
    size_t p = m.body_parent[b];
    Vec3 pos = d.xpos[p] + mul_vec(d.xmat[p], m.body_pos[b]);
    Quat quat = qmul(d.xquat[p], m.body_quat[b]);
    for (size_t j = m.body_jntadr[b]; j < m.body_jntadr[b] + m.body_jntnum[b]; j++) {
      size_t qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j];
      const Real* q = d.qpos.data();
      switch (m.jnt_type[j]) {
        case Free: {
          pos = V3(q[qa], q[qa + 1], q[qa + 2]);
          quat = qnormalized(Q4(q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]));
          d.xanchor[j] = pos;
          d.xaxis[j] = V3(0, 0, 1);
          Mat3 r = to_mat(quat);
          for (int k = 0; k < 3; k++) {
            Vec3 e = v3zero();
            e[k] = 1.0;
            d.cdof[da + k] = SP(v3zero(), e);
            Vec3 a = col(r, k);
            d.cdof[da + 3 + k] = SP(a, cross(pos, a));
          }
          break;
        }
        case Ball: {
          Vec3 anchor = pos + qrotate(quat, m.jnt_pos[j]);
          Quat ql = qnormalized(Q4(q[qa], q[qa + 1], q[qa + 2], q[qa + 3]));
          quat = qmul(quat, ql);
          pos = anchor - qrotate(quat, m.jnt_pos[j]);
          d.xanchor[j] = anchor;
          d.xaxis[j] = V3(0, 0, 1);
          Mat3 r = to_mat(quat);
          for (int k = 0; k < 3; k++) {
            Vec3 a = col(r, k);
            d.cdof[da + k] = SP(a, cross(anchor, a));
          }
          break;
        }
        case Hinge: {
          Vec3 anchor = pos + qrotate(quat, m.jnt_pos[j]);
          Vec3 axis = qrotate(quat, m.jnt_axis[j]);
          quat = qmul(quat, from_axis_angle(m.jnt_axis[j], q[qa] - m.qpos0[qa]));
          pos = anchor - qrotate(quat, m.jnt_pos[j]);
          d.xanchor[j] = anchor;
          d.xaxis[j] = axis;
          d.cdof[da] = SP(axis, cross(anchor, axis));
          break;
        }
        case Slide: {
          Vec3 anchor = pos + qrotate(quat, m.jnt_pos[j]);
          Vec3 axis = qrotate(quat, m.jnt_axis[j]);
          pos += axis * (q[qa] - m.qpos0[qa]);
          d.xanchor[j] = anchor;
          d.xaxis[j] = axis;
          d.cdof[da] = SP(v3zero(), axis);
          break;
        }
      }
    }
    quat = qnormalized(quat);
    d.xpos[b] = pos;
    d.xquat[b] = quat;
    d.xmat[b] = to_mat(quat);
    d.xipos[b] = pos + mul_vec(d.xmat[b], m.body_ipos[b]);
    d.cinert[b] = from_com(m.body_mass[b], d.xipos[b], mrotate(m.body_inertia[b], d.xmat[b]));
  }
  for (size_t g = 0; g < m.ngeom(); g++) {
    size_t b = m.geom_body[g];
    d.geom_xpos[g] = d.xpos[b] + mul_vec(d.xmat[b], m.geom_pos[g]);
    d.geom_xmat[g] = mul_mat(d.xmat[b], to_mat(m.geom_quat[g]));
  }
}

// forward.rs:93-112
void crba(const Model& m, Data& d) {
  size_t nv = m.nv;
  d.crb = d.cinert;
  for (size_t b = m.nbody() - 1; b >= 1; b--) {
    size_t p = m.body_parent[b];
    d.crb[p] = siadd(d.crb[p], d.crb[b]);
  }
  std::fill(d.qm.begin(), d.qm.end(), 0.0);
  for (size_t i = 0; i < nv; i++) {
    Spatial f = mul_motion(d.crb[m.dof_body[i]], d.cdof[i]);
    size_t j = i;
    while (j != NONE) {
      Real v = sdot(d.cdof[j], f);
      d.qm[i * nv + j] = v;
      d.qm[j * nv + i] = v;
      j = m.dof_parent[j];
    }
    d.qm[i * nv + i] += m.dof_armature[i];
  }
}

// forward.rs:115-144
void velocity(const Model& m, Data& d) {
  d.cvel[0] = spzero();
  for (size_t b = 1; b < m.nbody(); b++) {
    Spatial v = d.cvel[m.body_parent[b]];
    for (size_t j = m.body_jntadr[b]; j < m.body_jntadr[b] + m.body_jntnum[b]; j++) {
      size_t da = m.jnt_dofadr[j];
      int t = m.jnt_type[j];
      if (t == Free || t == Ball) {
        size_t lin = t == Free ? 3 : 0;
        size_t rot = t == Free ? da + 3 : da;
        for (size_t k = 0; k < lin; k++) {
          d.cdof_dot[da + k] = spzero();
          v += sscale(d.cdof[da + k], d.qvel[da + k]);
        }
        for (size_t k = 0; k < 3; k++) d.cdof_dot[rot + k] = cross_motion(v, d.cdof[rot + k]);
        for (size_t k = 0; k < 3; k++) v += sscale(d.cdof[rot + k], d.qvel[rot + k]);
      } else {
        d.cdof_dot[da] = cross_motion(v, d.cdof[da]);
        v += sscale(d.cdof[da], d.qvel[da]);
      }
    }
    d.cvel[b] = v;
  }
}

// forward.rs:147-171
void rne(const Model& m, Data& d) {
  size_t nb = m.nbody();
  d.cacc.resize(nb);
  d.cfrc.resize(nb);
  d.cacc[0] = SP(v3zero(), -m.gravity);
  d.cfrc[0] = spzero();
  for (size_t b = 1; b < nb; b++) {
    Spatial a = d.cacc[m.body_parent[b]];
    for (size_t k = m.body_dofadr[b]; k < m.body_dofadr[b] + m.body_dofnum[b]; k++) a += sscale(d.cdof_dot[k], d.qvel[k]);
    d.cacc[b] = a;
    Spatial h = mul_motion(d.cinert[b], d.cvel[b]);
    d.cfrc[b] = mul_motion(d.cinert[b], a) + cross_force(d.cvel[b], h);
  }
  for (size_t b = nb - 1; b >= 1; b--) {
    size_t p = m.body_parent[b];
    Spatial f = d.cfrc[b];
    d.cfrc[p] += f;
  }
  for (size_t k = 0; k < m.nv; k++) d.qfrc_bias[k] = sdot(d.cdof[k], d.cfrc[m.dof_body[k]]);
}

// forward.rs:173-183
void passive(const Model& m, Data& d) {
  for (size_t k = 0; k < m.nv; k++) d.qfrc_passive[k] = -m.dof_damping[k] * d.qvel[k];
  for (size_t j = 0; j < m.njnt(); j++) {
    Real kp = m.jnt_stiffness[j];
    if (kp != 0.0 && (m.jnt_type[j] == Hinge || m.jnt_type[j] == Slide))
      d.qfrc_passive[m.jnt_dofadr[j]] -= kp * (d.qpos[m.jnt_qposadr[j]] - m.jnt_springref[j]);
  }
}

// forward.rs:185-211
void actuation(const Model& m, Data& d) {
  std::fill(d.qfrc_actuator.begin(), d.qfrc_actuator.end(), 0.0);
  for (size_t u = 0; u < m.nu; u++) {
    Real gear = m.actuator_gear[u];
    const auto& moment = m.actuator_moment[u];
    Real sl = rsum_start(), sv = rsum_start();
    for (const auto& e : moment) sl += e.coef * d.qpos[m.jnt_qposadr[e.jnt]];
    for (const auto& e : moment) sv += e.coef * d.qvel[m.jnt_dofadr[e.jnt]];
    Real length = gear * sl;
    Real vel = gear * sv;
    Real ctrl = d.ctrl[u];
    if (m.actuator_ctrllimited[u]) ctrl = rclamp(ctrl, m.actuator_ctrlrange_lo[u], m.actuator_ctrlrange_hi[u]);
    Vec3 b = m.actuator_bias[u];
    Real f = m.actuator_gain[u] * ctrl + b[0] + b[1] * length + b[2] * vel;
    bool clamped = false;
    if (m.actuator_forcelimited[u]) {
      Real lo = m.actuator_forcerange_lo[u], hi = m.actuator_forcerange_hi[u];
      clamped = f < lo || f > hi;
      f = rclamp(f, lo, hi);
    }
    d.actuator_force[u] = f;
    d.actuator_clamped[u] = clamped;
    for (const auto& e : moment) d.qfrc_actuator[m.jnt_dofadr[e.jnt]] += gear * e.coef * f;
  }
}

// forward.rs:214-260
void collide(const Model& m, Data& d) {
  d.contacts.clear();
  Real margin = m.solver.contact_margin;
  std::vector<std::pair<Vec3, Vec3>> aabb(m.ngeom());
  for (size_t g = 0; g < m.ngeom(); g++) {
    Vec3 c = m.geom_aabb[g].first, h = m.geom_aabb[g].second;
    const Mat3& r = d.geom_xmat[g];
    if (!r_isfinite(h[0])) {
      aabb[g] = {v3zero(), h};
      continue;
    }
    Vec3 center = d.geom_xpos[g] + mul_vec(r, c);
    Vec3 half = V3(r_fabs(r[0]) * h[0] + r_fabs(r[1]) * h[1] + r_fabs(r[2]) * h[2],
                   r_fabs(r[3]) * h[0] + r_fabs(r[4]) * h[1] + r_fabs(r[5]) * h[2],
                   r_fabs(r[6]) * h[0] + r_fabs(r[7]) * h[1] + r_fabs(r[8]) * h[2]);
    aabb[g] = {center, half};
  }
  std::vector<Hit> hits;
  for (const auto& pr : m.collision_pairs) {
    size_t g1 = pr.first, g2 = pr.second;
    Vec3 c1 = aabb[g1].first, h1 = aabb[g1].second, c2 = aabb[g2].first, h2 = aabb[g2].second;
    bool skip = false;
    for (int k = 0; k < 3 && !skip; k++)
      if (r_fabs(c1[k] - c2[k]) > h1[k] + h2[k] + margin) skip = true;
    if (skip) continue;
    auto hull_of = [&](size_t g) { return m.geom_type[g] == Mesh ? &m.mesh_hull[m.geom_dataid[g]] : nullptr; };
    GeomPose p1 = {m.geom_type[g1], d.geom_xpos[g1], d.geom_xmat[g1], m.geom_size[g1], hull_of(g1)};
    GeomPose p2 = {m.geom_type[g2], d.geom_xpos[g2], d.geom_xmat[g2], m.geom_size[g2], hull_of(g2)};
    hits.clear();
    if (needs_general(m.geom_type[g1], m.geom_type[g2])) collide_general(p1, p2, margin, hits);
    else collide_pair(p1, p2, margin, hits);
    for (size_t i = 0; i < hits.size(); i++) {
      Contact c;
      c.pos = hits[i].pos;
      c.normal = hits[i].normal;
      c.tangent = hits[i].tangent;
      c.depth = hits[i].depth;
      c.geom[0] = g1;
      c.geom[1] = g2;
      c.feature = (uint32_t)i;
      d.contacts.push_back(c);
    }
  }
}

// forward.rs:263-271
void forward(const Model& m, Data& d) {
  kinematics(m, d);
  crba(m, d);
  velocity(m, d);
  rne(m, d);
  passive(m, d);
  actuation(m, d);
  collide(m, d);
}

// forward.rs:275-333 (Data::new starts from qpos_reset, data.rs:177)
void set_const(Model& m) {
  Data d = data_new(m);
  kinematics(m, d);
  crba(m, d);
  size_t nv = m.nv;
  m.dof_invweight.assign(nv, 0.0);
  m.body_invweight.assign(m.nbody(), {0.0, 0.0});
  if (nv == 0) return;
  std::vector<Real> l = d.qm;
  if (!cholesky(l.data(), nv)) return;
  std::vector<Real> x(nv), e(nv), row(nv);
  auto quad = [&](const std::vector<Real>& r) {
    x = r;
    cholesky_solve(l.data(), nv, x.data());
    Real s = rsum_start();
    for (size_t k = 0; k < nv; k++) s += r[k] * x[k];
    return s;
  };
  for (size_t k = 0; k < nv; k++) {
    std::fill(e.begin(), e.end(), 0.0);
    e[k] = 1.0;
    m.dof_invweight[k] = quad(e);
  }
  for (size_t j = 0; j < m.njnt(); j++) {
    size_t da = m.jnt_dofadr[j];
    std::vector<size_t> groups;
    if (m.jnt_type[j] == Free) groups = {0, 3};
    else if (m.jnt_type[j] == Ball) groups = {0};
    else continue;
    for (size_t g : groups) {
      Real s = rsum_start();
      for (size_t i = 0; i < 3; i++) s += m.dof_invweight[da + g + i];
      Real avg = s / 3.0;
      for (size_t i = 0; i < 3; i++) m.dof_invweight[da + g + i] = avg;
    }
  }
  for (size_t b = 1; b < m.nbody(); b++) {
    Real w[2] = {0.0, 0.0};
    for (int axis = 0; axis < 3; axis++) {
      Vec3 dir = v3zero();
      dir[axis] = 1.0;
      for (int kind = 0; kind < 2; kind++) {
        std::fill(row.begin(), row.end(), 0.0);
        size_t k = m.body_lastdof[b];
        while (k != NONE) {
          const Spatial& c = d.cdof[k];
          row[k] = kind == 0 ? dot(dir, point_velocity(c, d.xipos[b])) : dot(dir, c.ang);
          k = m.dof_parent[k];
        }
        w[kind] += quad(row) / 3.0;
      }
    }
    m.body_invweight[b] = {w[0], w[1]};
  }
}

// forward.rs:441-450
static bool is_ancestor(const Model& m, size_t a, size_t b) {
  size_t k = m.dof_parent[b];
  while (k != NONE) {
    if (k == a) return true;
    k = m.dof_parent[k];
  }
  return false;
}

// forward.rs:407-431
static void add_implicit_terms(const Model& m, const std::vector<char>& clamped, Real h, Real* mat) {
  size_t nv = m.nv;
  for (size_t k = 0; k < nv; k++) mat[k * nv + k] += h * m.dof_damping[k];
  if (m.integrator == ImplicitFast) {
    for (size_t u = 0; u < m.nu; u++) {
      Real kv = -m.actuator_bias[u][2] * m.actuator_gear[u] * m.actuator_gear[u];
      if (kv == 0.0 || clamped[u]) continue;
      for (const auto& ea : m.actuator_moment[u])
        for (const auto& eb : m.actuator_moment[u]) {
          size_t a = m.jnt_dofadr[ea.jnt], b = m.jnt_dofadr[eb.jnt];
          if (a == b || is_ancestor(m, a, b) || is_ancestor(m, b, a)) mat[a * nv + b] += h * kv * ea.coef * eb.coef;
        }
    }
  }
}

// forward.rs:392-401
static bool factor_mass(const Model& m, Data& d, Real h) {
  d.qm_chol = d.qm;
  add_implicit_terms(m, d.actuator_clamped, h, d.qm_chol.data());
  for (const auto& r : m.dof_blocks)
    if (!cholesky_block(d.qm_chol.data(), m.nv, r.start, r.end)) return false;
  return true;
}

// forward.rs:434-438
void mass_solve(const Model& m, const std::vector<Real>& l, Real* x) {
  for (const auto& r : m.dof_blocks) cholesky_solve_block(l.data(), m.nv, r.start, r.end, x);
}

// forward.rs:453-456
static bool has_implicit_terms(const Model& m) {
  for (Real b : m.dof_damping)
    if (b != 0.0) return true;
  if (m.integrator == ImplicitFast)
    for (const auto& b : m.actuator_bias)
      if (b[2] != 0.0) return true;
  return false;
}

// forward.rs:459-480
void integrate_pos(const Model& m, Real* qpos, const Real* qvel, Real h) {
  for (size_t j = 0; j < m.njnt(); j++) {
    size_t qa = m.jnt_qposadr[j], da = m.jnt_dofadr[j];
    switch (m.jnt_type[j]) {
      case Free: {
        for (int k = 0; k < 3; k++) qpos[qa + k] += h * qvel[da + k];
        Quat q = Q4(qpos[qa + 3], qpos[qa + 4], qpos[qa + 5], qpos[qa + 6]);
        Vec3 w = V3(qvel[da + 3], qvel[da + 4], qvel[da + 5]);
        Quat r = integrate_local(q, w, h);
        for (int k = 0; k < 4; k++) qpos[qa + 3 + k] = r[k];
        break;
      }
      case Ball: {
        Quat q = Q4(qpos[qa], qpos[qa + 1], qpos[qa + 2], qpos[qa + 3]);
        Vec3 w = V3(qvel[da], qvel[da + 1], qvel[da + 2]);
        Quat r = integrate_local(q, w, h);
        for (int k = 0; k < 4; k++) qpos[qa + k] = r[k];
        break;
      }
      default: qpos[qa] += h * qvel[da];
    }
  }
}

// forward.rs:336-389
bool step(const Model& m, Data& d) {
  forward(m, d);
  size_t nv = m.nv;
  Real h = m.timestep;
  bool damped = has_implicit_terms(m);
  std::vector<Real> smooth(nv), acc(nv), v(nv), v_pos;
  for (size_t k = 0; k < nv; k++) smooth[k] = d.qfrc_passive[k] + d.qfrc_actuator[k] + d.qfrc_applied[k] - d.qfrc_bias[k];
  acc = smooth;
  if (m.solver.kind == Newton) {
    if (!factor_mass(m, d, 0.0)) return false;
    mass_solve(m, d.qm_chol, acc.data());
    newton_solve(m, d, acc.data());
    if (damped) {
      if (!factor_mass(m, d, h)) return false;
      for (size_t k = 0; k < nv; k++) acc[k] = smooth[k] + d.qfrc_constraint[k];
      mass_solve(m, d.qm_chol, acc.data());
    }
    for (size_t k = 0; k < nv; k++) v[k] = d.qvel[k] + h * acc[k];
    v_pos = v;
  } else {
    if (!factor_mass(m, d, h)) return false;
    mass_solve(m, d.qm_chol, acc.data());
    for (size_t k = 0; k < nv; k++) v[k] = d.qvel[k] + h * acc[k];
    pgs_solve(m, d, v.data(), v_pos);
  }
  for (size_t k = 0; k < nv; k++) d.qacc[k] = (v[k] - d.qvel[k]) / h;
  integrate_pos(m, d.qpos.data(), v_pos.data(), h);
  d.qvel = v;
  d.time += h;
  return true;
}

}  // namespace orc
