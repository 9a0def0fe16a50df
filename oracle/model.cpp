// ORACLE — test infrastructure only (see math.hpp).
// Model compilation and per-env data: crates/pippin/src/model.rs and data.rs,
// plus the flat-descriptor import used by the tests.
#include <set>

#include "../include/pippin_hip.h"
#include "engine.hpp"

namespace orc {

// model.rs:318-363
static std::vector<Range> compute_dof_blocks(const Model& m) {
  size_t nv = m.nv;
  std::vector<size_t> root(nv);
  for (size_t k = 0; k < nv; k++) {
    size_t r = k;
    while (m.dof_parent[r] != NONE) r = m.dof_parent[r];
    root[k] = r;
  }
  std::vector<size_t> group = root;
  auto find = [](const std::vector<size_t>& g, size_t x) {
    while (g[x] != x) x = g[x];
    return x;
  };
  for (const auto& moment : m.actuator_moment) {
    std::vector<size_t> dofs;
    for (const auto& e : moment) dofs.push_back(m.jnt_dofadr[e.jnt]);
    for (size_t w = 0; w + 1 < dofs.size(); w++) {
      size_t a = find(group, root[dofs[w]]), b = find(group, root[dofs[w + 1]]);
      if (a != b) group[a > b ? a : b] = a < b ? a : b;
    }
  }
  std::vector<size_t> gid(nv);
  for (size_t k = 0; k < nv; k++) gid[k] = find(group, root[k]);
  std::vector<Range> blocks;
  std::set<size_t> seen;
  size_t k = 0;
  while (k < nv) {
    size_t e = k + 1;
    while (e < nv && gid[e] == gid[k]) e++;
    if (!seen.insert(gid[k]).second) {
      blocks.clear();
      if (nv > 0) blocks.push_back(Range{0, nv});
      return blocks;
    }
    blocks.push_back(Range{k, e});
    k = e;
  }
  return blocks;
}

// model.rs:366-382
std::pair<Vec3, Vec3> local_aabb(const Model& m, size_t g) {
  Vec3 s = m.geom_size[g];
  Real inf = INFINITY;
  switch (m.geom_type[g]) {
    case Plane: return {v3zero(), V3(inf, inf, inf)};
    case Sphere: return {v3zero(), V3(s[0], s[0], s[0])};
    case Capsule: return {v3zero(), V3(s[0], s[0], s[0] + s[1])};
    case Box: return {v3zero(), s};
    case Cylinder: return {v3zero(), V3(s[0], s[0], s[1])};
    case Mesh: {
      const HullShape& h = m.mesh_hull[m.geom_dataid[g]];
      Vec3 lo = V3(inf, inf, inf), hi = V3(-inf, -inf, -inf);
      for (const Vec3& p : h.vert)
        for (int k = 0; k < 3; k++) {
          lo[k] = p[k] < lo[k] ? p[k] : lo[k];
          hi[k] = p[k] > hi[k] ? p[k] : hi[k];
        }
      return {(lo + hi) * 0.5, (hi - lo) * 0.5};
    }
    default: return {v3zero(), v3zero()};
  }
}

// model.rs:385-395
Real rbound(const Model& m, size_t g) {
  Vec3 s = m.geom_size[g];
  switch (m.geom_type[g]) {
    case Plane: return INFINITY;
    case Sphere: return s[0];
    case Capsule: return s[0] + s[1];
    case Box: return norm(s);
    case Cylinder: return r_sqrt(s[0] * s[0] + s[1] * s[1]);
    case Mesh: {
      Real r = 0.0;
      for (const Vec3& p : m.mesh_hull[m.geom_dataid[g]].vert) {
        Real l = norm(p);
        r = l > r ? l : r;
      }
      return r;
    }
    default: return 0.0;
  }
}

// A hull from the descriptor's mesh arrays, with its face normals
// (cross(v1 - v0, v2 - v0) over its length) and the faces around each vertex.
static HullShape hull_from_desc(const ph_model_desc& d, int mesh) {
  HullShape h;
  const int va = d.mesh_vertadr[mesh], vn = d.mesh_vertnum[mesh];
  for (int i = 0; i < vn; i++)
    h.vert.push_back(V3(d.mesh_vert[3 * (va + i)], d.mesh_vert[3 * (va + i) + 1], d.mesh_vert[3 * (va + i) + 2]));
  h.vert_faces.resize((size_t)vn);
  const int fa = d.mesh_faceadr[mesh], fn = d.mesh_facenum[mesh];
  for (int f = 0; f < fn; f++) {
    std::vector<size_t> face;
    for (int k = 0; k < d.face_vertnum[fa + f]; k++) face.push_back((size_t)d.face_vert[d.face_vertadr[fa + f] + k]);
    for (size_t v : face) h.vert_faces[v].push_back((size_t)f);
    Vec3 nrm = cross(h.vert[face[1]] - h.vert[face[0]], h.vert[face[2]] - h.vert[face[0]]);
    Real l = r_sqrt(dot(nrm, nrm));
    h.normal.push_back(V3(nrm[0] / l, nrm[1] / l, nrm[2] / l));
    h.faces.push_back(face);
  }
  return h;
}

// model.rs:430-452
static bool can_collide(const Model& m, size_t g1, size_t g2) {
  size_t b1 = m.geom_body[g1], b2 = m.geom_body[g2];
  for (const auto& ex : m.exclude_pairs)
    if ((ex.first == b1 && ex.second == b2) || (ex.first == b2 && ex.second == b1)) return false;
  size_t w1 = m.body_weldid[b1], w2 = m.body_weldid[b2];
  if (w1 == w2) return false;
  bool masks = (m.geom_contype[g1] & m.geom_conaffinity[g2]) != 0 || (m.geom_contype[g2] & m.geom_conaffinity[g1]) != 0;
  if (!masks) return false;
  size_t p1 = m.body_weldid[m.body_parent[w1]];
  size_t p2 = m.body_weldid[m.body_parent[w2]];
  if ((p1 == w2 && w2 != 0) || (p2 == w1 && w1 != 0)) return false;
  return pair_supported(m.geom_type[g1], m.geom_type[g2]);
}

// model.rs:398-407
void rebuild_collision_pairs(Model& m) {
  m.collision_pairs.clear();
  for (size_t g1 = 0; g1 < m.ngeom(); g1++)
    for (size_t g2 = g1 + 1; g2 < m.ngeom(); g2++)
      if (can_collide(m, g1, g2)) m.collision_pairs.push_back({g1, g2});
}

// model.rs:279-316
void compile(Model& m) {
  if (m.qpos_reset.size() != m.qpos0.size()) m.qpos_reset = m.qpos0;
  size_t nbody = m.nbody();
  m.body_lastdof.assign(nbody, NONE);
  m.dof_parent.assign(m.nv, NONE);
  for (size_t b = 1; b < nbody; b++) {
    size_t last = m.body_lastdof[m.body_parent[b]];
    for (size_t k = 0; k < m.body_dofnum[b]; k++) {
      size_t d = m.body_dofadr[b] + k;
      m.dof_parent[d] = last;
      last = d;
    }
    m.body_lastdof[b] = last;
  }
  m.dof_blocks = compute_dof_blocks(m);
  m.body_weldid.resize(nbody);
  for (size_t b = 0; b < nbody; b++) m.body_weldid[b] = b;
  for (size_t b = 1; b < nbody; b++)
    if (m.body_jntnum[b] == 0) m.body_weldid[b] = m.body_weldid[m.body_parent[b]];
  m.geom_rbound.resize(m.ngeom());
  m.geom_aabb.resize(m.ngeom());
  for (size_t g = 0; g < m.ngeom(); g++) {
    m.geom_rbound[g] = rbound(m, g);
    m.geom_aabb[g] = local_aabb(m, g);
  }
  rebuild_collision_pairs(m);
  if (m.solver.contact_hertz <= 0.0) m.solver.contact_hertz = 0.25 / m.timestep;
}

static size_t idx(int32_t v) { return v < 0 ? NONE : (size_t)v; }

// Import the raw fields of a flat descriptor, then run compile() and
// set_const() as mjcf.rs:389-393 does after parsing.
Model model_from_desc(const ph_model_desc& d) {
  Model m;
  m.timestep = d.timestep;
  m.gravity = V3(d.gravity[0], d.gravity[1], d.gravity[2]);
  m.integrator = d.integrator;
  m.solver.kind = d.solver_kind;
  m.solver.iterations = (size_t)d.solver_iterations;
  m.solver.tolerance = d.tolerance;
  m.solver.contact_margin = d.contact_margin;
  for (int i = 0; i < 2; i++) m.solver.solref[i] = d.solref[i];
  for (int i = 0; i < 5; i++) m.solver.solimp[i] = d.solimp[i];
  m.solver.impratio = d.impratio;
  m.solver.contact_hertz = d.contact_hertz;
  m.solver.contact_damping_ratio = d.contact_damping_ratio;
  m.solver.contact_slop = d.contact_slop;
  m.nq = d.nq;
  m.nv = d.nv;
  m.nu = d.nu;
  for (int b = 0; b < d.nbody; b++) {
    m.body_parent.push_back(idx(d.body_parent[b]));
    m.body_pos.push_back(V3(d.body_pos[3 * b], d.body_pos[3 * b + 1], d.body_pos[3 * b + 2]));
    m.body_quat.push_back(Q4(d.body_quat[4 * b], d.body_quat[4 * b + 1], d.body_quat[4 * b + 2], d.body_quat[4 * b + 3]));
    m.body_jntadr.push_back((size_t)d.body_jntadr[b]);
    m.body_jntnum.push_back((size_t)d.body_jntnum[b]);
    m.body_dofadr.push_back((size_t)d.body_dofadr[b]);
    m.body_dofnum.push_back((size_t)d.body_dofnum[b]);
    m.body_mass.push_back(d.body_mass[b]);
    m.body_ipos.push_back(V3(d.body_ipos[3 * b], d.body_ipos[3 * b + 1], d.body_ipos[3 * b + 2]));
    Mat3 I;
    for (int k = 0; k < 9; k++) I[k] = d.body_inertia[9 * b + k];
    m.body_inertia.push_back(I);
  }
  for (int j = 0; j < d.njnt; j++) {
    m.jnt_type.push_back(d.jnt_type[j]);
    m.jnt_body.push_back((size_t)d.jnt_body[j]);
    m.jnt_qposadr.push_back((size_t)d.jnt_qposadr[j]);
    m.jnt_dofadr.push_back((size_t)d.jnt_dofadr[j]);
    m.jnt_pos.push_back(V3(d.jnt_pos[3 * j], d.jnt_pos[3 * j + 1], d.jnt_pos[3 * j + 2]));
    m.jnt_axis.push_back(V3(d.jnt_axis[3 * j], d.jnt_axis[3 * j + 1], d.jnt_axis[3 * j + 2]));
    m.jnt_limited.push_back(d.jnt_limited[j] != 0);
    m.jnt_range_lo.push_back(d.jnt_range[2 * j]);
    m.jnt_range_hi.push_back(d.jnt_range[2 * j + 1]);
    m.jnt_stiffness.push_back(d.jnt_stiffness[j]);
    m.jnt_springref.push_back(d.jnt_springref[j]);
  }
  for (int k = 0; k < d.nv; k++) {
    m.dof_body.push_back((size_t)d.dof_body[k]);
    m.dof_jnt.push_back((size_t)d.dof_jnt[k]);
    m.dof_damping.push_back(d.dof_damping[k]);
    m.dof_armature.push_back(d.dof_armature[k]);
  }
  for (int k = 0; k < d.nmesh; k++) m.mesh_hull.push_back(hull_from_desc(d, k));
  for (int g = 0; g < d.ngeom; g++) {
    m.geom_type.push_back(d.geom_type[g]);
    m.geom_body.push_back((size_t)d.geom_body[g]);
    m.geom_dataid.push_back(idx(d.geom_dataid[g]));
    m.geom_pos.push_back(V3(d.geom_pos[3 * g], d.geom_pos[3 * g + 1], d.geom_pos[3 * g + 2]));
    m.geom_quat.push_back(Q4(d.geom_quat[4 * g], d.geom_quat[4 * g + 1], d.geom_quat[4 * g + 2], d.geom_quat[4 * g + 3]));
    m.geom_size.push_back(V3(d.geom_size[3 * g], d.geom_size[3 * g + 1], d.geom_size[3 * g + 2]));
    m.geom_friction.push_back(d.geom_friction[g]);
    m.geom_contype.push_back(d.geom_contype[g]);
    m.geom_conaffinity.push_back(d.geom_conaffinity[g]);
    for (int k = 0; k < 4; k++) m.geom_rgba.push_back(d.geom_rgba[4 * g + k]);
  }
  for (int c = 0; c < d.ncam; c++) {
    m.cam_body.push_back((size_t)d.cam_body[c]);
    m.cam_pos.push_back(V3(d.cam_pos[3 * c], d.cam_pos[3 * c + 1], d.cam_pos[3 * c + 2]));
    m.cam_quat.push_back(Q4(d.cam_quat[4 * c], d.cam_quat[4 * c + 1], d.cam_quat[4 * c + 2], d.cam_quat[4 * c + 3]));
  }
  for (int u = 0; u < d.nu; u++) {
    m.actuator_gear.push_back(d.actuator_gear[u]);
    m.actuator_gain.push_back(d.actuator_gain[u]);
    m.actuator_bias.push_back(V3(d.actuator_bias[3 * u], d.actuator_bias[3 * u + 1], d.actuator_bias[3 * u + 2]));
    m.actuator_ctrllimited.push_back(d.actuator_ctrllimited[u] != 0);
    m.actuator_ctrlrange_lo.push_back(d.actuator_ctrlrange[2 * u]);
    m.actuator_ctrlrange_hi.push_back(d.actuator_ctrlrange[2 * u + 1]);
    m.actuator_forcelimited.push_back(d.actuator_forcelimited[u] != 0);
    m.actuator_forcerange_lo.push_back(d.actuator_forcerange[2 * u]);
    m.actuator_forcerange_hi.push_back(d.actuator_forcerange[2 * u + 1]);
    std::vector<MomentEntry> mom;
    for (int e = 0; e < d.actuator_moment_num[u]; e++) {
      int i = d.actuator_moment_adr[u] + e;
      mom.push_back(MomentEntry{(size_t)d.moment_jnt[i], d.moment_coef[i]});
    }
    m.actuator_moment.push_back(mom);
  }
  for (int e = 0; e < d.neq; e++) {
    m.eq_joint1.push_back((size_t)d.eq_joint1[e]);
    m.eq_joint2.push_back(idx(d.eq_joint2[e]));
    m.eq_polycoef.push_back(std::vector<Real>(d.eq_polycoef + 5 * e, d.eq_polycoef + 5 * e + 5));
    m.eq_solref.push_back(std::vector<Real>(d.eq_solref + 2 * e, d.eq_solref + 2 * e + 2));
    m.eq_solimp.push_back(std::vector<Real>(d.eq_solimp + 5 * e, d.eq_solimp + 5 * e + 5));
  }
  for (int e = 0; e < d.nexclude; e++)
    m.exclude_pairs.push_back({(size_t)d.exclude_pairs[2 * e], (size_t)d.exclude_pairs[2 * e + 1]});
  m.qpos0.assign(d.qpos0, d.qpos0 + d.nq);
  m.qpos_reset.assign(d.qpos_reset, d.qpos_reset + d.nq);
  compile(m);
  set_const(m);
  return m;
}

// data.rs:83-120
Data data_new(const Model& m) {
  Data d;
  size_t nb = m.nbody(), ng = m.ngeom(), nj = m.njnt(), nv = m.nv;
  d.time = 0.0;
  d.qpos = m.qpos_reset;
  d.qvel.assign(nv, 0.0);
  d.ctrl.assign(m.nu, 0.0);
  d.qfrc_applied.assign(nv, 0.0);
  d.xpos.assign(nb, v3zero());
  d.xquat.assign(nb, qidentity());
  d.xmat.assign(nb, m3identity());
  d.xipos.assign(nb, v3zero());
  d.xanchor.assign(nj, v3zero());
  d.xaxis.assign(nj, v3zero());
  d.geom_xpos.assign(ng, v3zero());
  d.geom_xmat.assign(ng, m3identity());
  d.cdof.assign(nv, spzero());
  d.cdof_dot.assign(nv, spzero());
  d.cinert.assign(nb, sizero());
  d.crb.assign(nb, sizero());
  d.cvel.assign(nb, spzero());
  d.qm.assign(nv * nv, 0.0);
  d.qm_chol.assign(nv * nv, 0.0);
  d.qfrc_bias.assign(nv, 0.0);
  d.qfrc_passive.assign(nv, 0.0);
  d.qfrc_actuator.assign(nv, 0.0);
  d.qfrc_constraint.assign(nv, 0.0);
  d.qacc.assign(nv, 0.0);
  d.actuator_force.assign(m.nu, 0.0);
  d.actuator_clamped.assign(m.nu, 0);
  return d;
}

// data.rs:122-131: qacc (the Newton warm start) is deliberately kept.
void data_reset(Data& d, const Model& m) {
  d.time = 0.0;
  d.qpos = m.qpos_reset;
  std::fill(d.qvel.begin(), d.qvel.end(), 0.0);
  std::fill(d.ctrl.begin(), d.ctrl.end(), 0.0);
  std::fill(d.qfrc_applied.begin(), d.qfrc_applied.end(), 0.0);
  d.contacts.clear();
  d.prev_contacts.clear();
  d.prev_lambda.clear();
}

}  // namespace orc
