// Host model build: copy of the raw descriptor and the derived data of
// crates/pippin/src/model.rs::compile (279-316), in the device's layout.
#include "host_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>

namespace phh {

template <class T>
static void cp(std::vector<T>& v, const T* p, int64_t n) {
  v.assign(p, p + (n > 0 ? n : 0));
}

void retarget(HostModel& m) {
  ph_model_desc& d = m.d;
  d.body_parent = m.body_parent.data();
  d.body_pos = m.body_pos.data();
  d.body_quat = m.body_quat.data();
  d.body_jntadr = m.body_jntadr.data();
  d.body_jntnum = m.body_jntnum.data();
  d.body_dofadr = m.body_dofadr.data();
  d.body_dofnum = m.body_dofnum.data();
  d.body_mass = m.body_mass.data();
  d.body_ipos = m.body_ipos.data();
  d.body_inertia = m.body_inertia.data();
  d.jnt_type = m.jnt_type.data();
  d.jnt_body = m.jnt_body.data();
  d.jnt_qposadr = m.jnt_qposadr.data();
  d.jnt_dofadr = m.jnt_dofadr.data();
  d.jnt_pos = m.jnt_pos.data();
  d.jnt_axis = m.jnt_axis.data();
  d.jnt_limited = m.jnt_limited.data();
  d.jnt_range = m.jnt_range.data();
  d.jnt_stiffness = m.jnt_stiffness.data();
  d.jnt_springref = m.jnt_springref.data();
  d.dof_body = m.dof_body.data();
  d.dof_jnt = m.dof_jnt.data();
  d.dof_damping = m.dof_damping.data();
  d.dof_armature = m.dof_armature.data();
  d.geom_type = m.geom_type.data();
  d.geom_body = m.geom_body.data();
  d.geom_pos = m.geom_pos.data();
  d.geom_quat = m.geom_quat.data();
  d.geom_size = m.geom_size.data();
  d.geom_friction = m.geom_friction.data();
  d.geom_contype = m.geom_contype.data();
  d.geom_conaffinity = m.geom_conaffinity.data();
  d.geom_rgba = m.geom_rgba.data();
  d.geom_group = m.geom_group.data();
  d.geom_dataid = m.geom_dataid.data();
  d.cam_body = m.cam_body.data();
  d.cam_pos = m.cam_pos.data();
  d.cam_quat = m.cam_quat.data();
  d.cam_fovy = m.cam_fovy.data();
  d.actuator_gear = m.actuator_gear.data();
  d.actuator_gain = m.actuator_gain.data();
  d.actuator_bias = m.actuator_bias.data();
  d.actuator_ctrllimited = m.actuator_ctrllimited.data();
  d.actuator_ctrlrange = m.actuator_ctrlrange.data();
  d.actuator_forcelimited = m.actuator_forcelimited.data();
  d.actuator_forcerange = m.actuator_forcerange.data();
  d.actuator_moment_adr = m.actuator_moment_adr.data();
  d.actuator_moment_num = m.actuator_moment_num.data();
  d.moment_jnt = m.moment_jnt.data();
  d.moment_coef = m.moment_coef.data();
  d.eq_joint1 = m.eq_joint1.data();
  d.eq_joint2 = m.eq_joint2.data();
  d.eq_polycoef = m.eq_polycoef.data();
  d.eq_solref = m.eq_solref.data();
  d.eq_solimp = m.eq_solimp.data();
  d.exclude_pairs = m.exclude_pairs.data();
  d.qpos0 = m.qpos0.data();
  d.qpos_reset = m.qpos_reset.data();
  d.mesh_vertadr = m.mesh_vertadr.data();
  d.mesh_vertnum = m.mesh_vertnum.data();
  d.mesh_faceadr = m.mesh_faceadr.data();
  d.mesh_facenum = m.mesh_facenum.data();
  d.mesh_vert = m.mesh_vert.data();
  d.face_vertadr = m.face_vertadr.data();
  d.face_vertnum = m.face_vertnum.data();
  d.face_vert = m.face_vert.data();
}

// Copy the hull arrays of `s`, check their indices, and derive face normals and
// the faces around each vertex. False (err set) for malformed hulls.
static bool import_hulls(const ph_model_desc& s, HostModel& m, std::string* err) {
  if (s.nmesh < 0 || s.nmeshvert < 0 || s.nmeshface < 0 || s.nmeshfacevert < 0) {
    *err = "negative mesh counts";
    return false;
  }
  cp(m.mesh_vertadr, s.mesh_vertadr, s.nmesh);
  cp(m.mesh_vertnum, s.mesh_vertnum, s.nmesh);
  cp(m.mesh_faceadr, s.mesh_faceadr, s.nmesh);
  cp(m.mesh_facenum, s.mesh_facenum, s.nmesh);
  cp(m.mesh_vert, s.mesh_vert, 3 * (int64_t)s.nmeshvert);
  cp(m.face_vertadr, s.face_vertadr, s.nmeshface);
  cp(m.face_vertnum, s.face_vertnum, s.nmeshface);
  cp(m.face_vert, s.face_vert, s.nmeshfacevert);
  m.face_normal.assign(3 * (size_t)s.nmeshface, 0.0);
  m.vert_faceadr.assign(s.nmeshvert, 0);
  m.vert_facenum.assign(s.nmeshvert, 0);
  std::vector<std::vector<int32_t>> around(s.nmeshvert);
  for (int k = 0; k < s.nmesh; k++) {
    const int va = m.mesh_vertadr[k], vn = m.mesh_vertnum[k], fa = m.mesh_faceadr[k], fn = m.mesh_facenum[k];
    if (va < 0 || vn < 4 || va + vn > s.nmeshvert || vn > 64 || fa < 0 || fn < 4 || fa + fn > s.nmeshface) {
      *err = "mesh " + std::to_string(k) + ": hull ranges out of bounds (or more than 64 vertices)";
      return false;
    }
    for (int f = fa; f < fa + fn; f++) {
      const int a = m.face_vertadr[f], c = m.face_vertnum[f];
      if (a < 0 || c < 3 || a + c > s.nmeshfacevert) {
        *err = "mesh " + std::to_string(k) + ": face ranges out of bounds";
        return false;
      }
      for (int i = 0; i < c; i++) {
        const int v = m.face_vert[a + i];
        if (v < 0 || v >= vn) {
          *err = "mesh " + std::to_string(k) + ": face vertex out of range";
          return false;
        }
        around[va + v].push_back(f);
      }
      auto P = [&](int i) {
        const double* p = &m.mesh_vert[3 * (size_t)(va + m.face_vert[a + i])];
        return phk::v3(p[0], p[1], p[2]);
      };
      const phk::V3 n = phk::cross(P(1) - P(0), P(2) - P(0));
      const double l = std::sqrt(phk::dot(n, n));
      m.face_normal[3 * f] = n.x / l;
      m.face_normal[3 * f + 1] = n.y / l;
      m.face_normal[3 * f + 2] = n.z / l;
    }
  }
  m.vert_face.clear();
  for (int v = 0; v < s.nmeshvert; v++) {
    m.vert_faceadr[v] = (int)m.vert_face.size();
    m.vert_facenum[v] = (int)around[v].size();
    if (around[v].empty()) {
      *err = "hull vertex " + std::to_string(v) + " lies on no face";
      return false;
    }
    m.vert_face.insert(m.vert_face.end(), around[v].begin(), around[v].end());
  }
  return true;
}

void local_aabb(int t, const double* s, double* o) {
  const double inf = INFINITY;
  o[0] = o[1] = o[2] = 0.0;
  switch (t) {
    case PH_GEOM_PLANE: o[3] = o[4] = o[5] = inf; break;
    case PH_GEOM_SPHERE: o[3] = o[4] = o[5] = s[0]; break;
    case PH_GEOM_CAPSULE: o[3] = s[0]; o[4] = s[0]; o[5] = s[0] + s[1]; break;
    case PH_GEOM_BOX: o[3] = s[0]; o[4] = s[1]; o[5] = s[2]; break;
    case PH_GEOM_CYLINDER: o[3] = s[0]; o[4] = s[0]; o[5] = s[1]; break;
    default: o[3] = o[4] = o[5] = 0.0;
  }
}

bool has_implicit_terms(const HostModel& m, const double* damping, const double* bias) {
  for (int k = 0; k < m.d.nv; k++)
    if (damping[k] != 0.0) return true;
  if (m.d.integrator == PH_INTEGRATOR_IMPLICITFAST)
    for (int u = 0; u < m.d.nu; u++)
      if (bias[3 * u + 2] != 0.0) return true;
  return false;
}

// model.rs:430-452 without the contype/conaffinity test (applied per env on the device).
bool pair_static_ok(const HostModel& m, int g1, int g2) {
  int b1 = m.geom_body[g1], b2 = m.geom_body[g2];
  for (int e = 0; e < m.d.nexclude; e++) {
    int a = m.exclude_pairs[2 * e], b = m.exclude_pairs[2 * e + 1];
    if ((a == b1 && b == b2) || (a == b2 && b == b1)) return false;
  }
  int w1 = m.body_weldid[b1], w2 = m.body_weldid[b2];
  if (w1 == w2) return false;
  int p1 = m.body_weldid[m.body_parent[w1]], p2 = m.body_weldid[m.body_parent[w2]];
  if ((p1 == w2 && w2 != 0) || (p2 == w1 && w1 != 0)) return false;
  return !(m.geom_type[g1] == PH_GEOM_PLANE && m.geom_type[g2] == PH_GEOM_PLANE);
}

static bool needs_general(int a, int b) {
  return a == PH_GEOM_MESH || b == PH_GEOM_MESH || (a == PH_GEOM_CYLINDER && b != PH_GEOM_PLANE) ||
         (b == PH_GEOM_CYLINDER && a != PH_GEOM_PLANE);
}

static int pair_max_hits(int a, int b) {
  if (needs_general(a, b)) return 4;  // manifolds above four points are reduced to four
  if (a > b) std::swap(a, b);
  if (a == PH_GEOM_PLANE && b == PH_GEOM_CAPSULE) return 2;
  if ((a == PH_GEOM_PLANE && (b == PH_GEOM_BOX || b == PH_GEOM_CYLINDER)) || (a == PH_GEOM_BOX && b == PH_GEOM_BOX))
    return 4;
  if (a == PH_GEOM_CAPSULE && (b == PH_GEOM_CAPSULE || b == PH_GEOM_BOX)) return 3;
  return 1;
}

// model.rs:318-363
static void compute_dof_blocks(HostModel& m) {
  int nv = m.d.nv;
  std::vector<int> root(nv), group;
  for (int k = 0; k < nv; k++) {
    int r = k;
    while (m.dof_parent[r] >= 0) r = m.dof_parent[r];
    root[k] = r;
  }
  group = root;
  auto find = [&](int x) {
    while (group[x] != x) x = group[x];
    return x;
  };
  for (int u = 0; u < m.d.nu; u++) {
    int a0 = m.actuator_moment_adr[u], n = m.actuator_moment_num[u];
    for (int w = 0; w + 1 < n; w++) {
      int a = find(root[m.jnt_dofadr[m.moment_jnt[a0 + w]]]);
      int b = find(root[m.jnt_dofadr[m.moment_jnt[a0 + w + 1]]]);
      if (a != b) group[std::max(a, b)] = std::min(a, b);
    }
  }
  std::vector<int> gid(nv);
  for (int k = 0; k < nv; k++) gid[k] = find(root[k]);
  m.block_adr.clear();
  std::set<int> seen;
  int k = 0;
  m.block_adr.push_back(0);
  while (k < nv) {
    int e = k + 1;
    while (e < nv && gid[e] == gid[k]) e++;
    if (!seen.insert(gid[k]).second) {
      m.block_adr.assign({0, nv});
      return;
    }
    m.block_adr.push_back(e);
    k = e;
  }
  if (nv == 0) m.block_adr.clear();
}

static void compute_dof_block(HostModel& m) {
  m.dof_block.assign(m.d.nv, 0);
  for (size_t b = 0; b + 1 < m.block_adr.size(); b++)
    for (int k = m.block_adr[b]; k < m.block_adr[b + 1]; k++) m.dof_block[k] = (int)b;
}

bool build_host_model(const ph_model_desc& s, HostModel& m, std::string* err) {
  m.d = s;
  int nb = s.nbody, nj = s.njnt, nv = s.nv, ng = s.ngeom, nu = s.nu;
  if (nb < 1) {
    *err = "model needs a world body";
    return false;
  }
  if (nv > 64) {
    *err = "models with more than 64 dofs are not supported by the device path";
    return false;
  }
  cp(m.body_parent, s.body_parent, nb);
  cp(m.body_pos, s.body_pos, 3 * nb);
  cp(m.body_quat, s.body_quat, 4 * nb);
  cp(m.body_jntadr, s.body_jntadr, nb);
  cp(m.body_jntnum, s.body_jntnum, nb);
  cp(m.body_dofadr, s.body_dofadr, nb);
  cp(m.body_dofnum, s.body_dofnum, nb);
  cp(m.body_mass, s.body_mass, nb);
  cp(m.body_ipos, s.body_ipos, 3 * nb);
  cp(m.body_inertia, s.body_inertia, 9 * nb);
  cp(m.jnt_type, s.jnt_type, nj);
  cp(m.jnt_body, s.jnt_body, nj);
  cp(m.jnt_qposadr, s.jnt_qposadr, nj);
  cp(m.jnt_dofadr, s.jnt_dofadr, nj);
  cp(m.jnt_pos, s.jnt_pos, 3 * nj);
  cp(m.jnt_axis, s.jnt_axis, 3 * nj);
  cp(m.jnt_limited, s.jnt_limited, nj);
  cp(m.jnt_range, s.jnt_range, 2 * nj);
  cp(m.jnt_stiffness, s.jnt_stiffness, nj);
  cp(m.jnt_springref, s.jnt_springref, nj);
  cp(m.dof_body, s.dof_body, nv);
  cp(m.dof_jnt, s.dof_jnt, nv);
  cp(m.dof_damping, s.dof_damping, nv);
  cp(m.dof_armature, s.dof_armature, nv);
  cp(m.geom_type, s.geom_type, ng);
  cp(m.geom_body, s.geom_body, ng);
  cp(m.geom_pos, s.geom_pos, 3 * ng);
  cp(m.geom_quat, s.geom_quat, 4 * ng);
  cp(m.geom_size, s.geom_size, 3 * ng);
  cp(m.geom_friction, s.geom_friction, ng);
  cp(m.geom_contype, s.geom_contype, ng);
  cp(m.geom_conaffinity, s.geom_conaffinity, ng);
  cp(m.geom_rgba, s.geom_rgba, 4 * ng);
  cp(m.geom_group, s.geom_group, ng);
  cp(m.geom_dataid, s.geom_dataid, ng);
  cp(m.cam_body, s.cam_body, s.ncam);
  cp(m.cam_pos, s.cam_pos, 3 * s.ncam);
  cp(m.cam_quat, s.cam_quat, 4 * s.ncam);
  cp(m.cam_fovy, s.cam_fovy, s.ncam);
  cp(m.actuator_gear, s.actuator_gear, nu);
  cp(m.actuator_gain, s.actuator_gain, nu);
  cp(m.actuator_bias, s.actuator_bias, 3 * nu);
  cp(m.actuator_ctrllimited, s.actuator_ctrllimited, nu);
  cp(m.actuator_ctrlrange, s.actuator_ctrlrange, 2 * nu);
  cp(m.actuator_forcelimited, s.actuator_forcelimited, nu);
  cp(m.actuator_forcerange, s.actuator_forcerange, 2 * nu);
  cp(m.actuator_moment_adr, s.actuator_moment_adr, nu);
  cp(m.actuator_moment_num, s.actuator_moment_num, nu);
  cp(m.moment_jnt, s.moment_jnt, s.nmoment);
  cp(m.moment_coef, s.moment_coef, s.nmoment);
  cp(m.eq_joint1, s.eq_joint1, s.neq);
  cp(m.eq_joint2, s.eq_joint2, s.neq);
  cp(m.eq_polycoef, s.eq_polycoef, 5 * s.neq);
  cp(m.eq_solref, s.eq_solref, 2 * s.neq);
  cp(m.eq_solimp, s.eq_solimp, 5 * s.neq);
  cp(m.exclude_pairs, s.exclude_pairs, 2 * s.nexclude);
  cp(m.qpos0, s.qpos0, s.nq);
  cp(m.qpos_reset, s.qpos_reset, s.nq);
  if ((int)m.qpos_reset.size() != s.nq) m.qpos_reset = m.qpos0;
  if (!import_hulls(s, m, err)) return false;
  for (int g = 0; g < ng; g++)
    if (m.geom_type[g] == PH_GEOM_MESH && (m.geom_dataid[g] < 0 || m.geom_dataid[g] >= s.nmesh)) {
      *err = "mesh geom " + std::to_string(g) + " names no mesh";
      return false;
    }
  retarget(m);

  // dof tree (model.rs:285-296)
  m.body_lastdof.assign(nb, -1);
  m.dof_parent.assign(nv, -1);
  for (int b = 1; b < nb; b++) {
    int last = m.body_lastdof[m.body_parent[b]];
    for (int k = 0; k < m.body_dofnum[b]; k++) {
      int d = m.body_dofadr[b] + k;
      m.dof_parent[d] = last;
      last = d;
    }
    m.body_lastdof[b] = last;
  }
  compute_dof_blocks(m);
  compute_dof_block(m);
  // weld ids (model.rs:301-306)
  m.body_weldid.resize(nb);
  for (int b = 0; b < nb; b++) m.body_weldid[b] = b;
  for (int b = 1; b < nb; b++)
    if (m.body_jntnum[b] == 0) m.body_weldid[b] = m.body_weldid[m.body_parent[b]];
  // dof chains
  m.body_chain.assign(nb, 0);
  for (int b = 1; b < nb; b++)
    for (int k = m.body_lastdof[b]; k >= 0; k = m.dof_parent[k]) m.body_chain[b] |= (uint64_t)1 << k;
  m.dof_anc.assign(nv, 0);
  for (int k = 0; k < nv; k++)
    for (int a = m.dof_parent[k]; a >= 0; a = m.dof_parent[a]) m.dof_anc[k] |= (uint64_t)1 << a;
  // bodies grouped by depth, parents before children
  std::vector<int> depth(nb, 0);
  int maxd = 0;
  for (int b = 1; b < nb; b++) {
    if (m.body_parent[b] >= b) {
      *err = "bodies must be listed after their parents";
      return false;
    }
    depth[b] = depth[m.body_parent[b]] + 1;
    maxd = std::max(maxd, depth[b]);
  }
  m.level_adr.clear();
  m.level_body.clear();
  for (int l = 1; l <= maxd; l++) {
    m.level_adr.push_back((int)m.level_body.size());
    for (int b = 1; b < nb; b++)
      if (depth[b] == l) m.level_body.push_back(b);
  }
  m.level_adr.push_back((int)m.level_body.size());
  // Leaf-to-root sums (forward.rs:100-103, 166-170: for b descending,
  // x[parent(b)] += x[b]). A parent's sum only depends on the order of its own
  // children, which all sit one level below it; so levels deepest first, and
  // within a level each parent adds its children in descending index order.
  m.acc_level_adr.clear();
  m.acc_parent.clear();
  m.acc_child_adr.clear();
  m.acc_child.clear();
  for (int l = maxd; l >= 1; l--) {
    m.acc_level_adr.push_back((int)m.acc_parent.size());
    std::vector<int> parents;
    for (int b = 1; b < nb; b++)
      if (depth[b] == l && std::find(parents.begin(), parents.end(), m.body_parent[b]) == parents.end())
        parents.push_back(m.body_parent[b]);
    for (int p : parents) {
      m.acc_parent.push_back(p);
      m.acc_child_adr.push_back((int)m.acc_child.size());
      for (int b = nb - 1; b >= 1; b--)
        if (depth[b] == l && m.body_parent[b] == p) m.acc_child.push_back(b);
    }
  }
  m.acc_level_adr.push_back((int)m.acc_parent.size());
  m.acc_child_adr.push_back((int)m.acc_child.size());
  // local boxes
  m.geom_aabb.assign(6 * ng, 0.0);
  for (int g = 0; g < ng; g++) {
    local_aabb(m.geom_type[g], &m.geom_size[3 * g], &m.geom_aabb[6 * g]);
    if (m.geom_type[g] == PH_GEOM_MESH) {  // model.rs:375-380: the hull's box, centre and half-extents
      const int k = m.geom_dataid[g];
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int v = m.mesh_vertadr[k]; v < m.mesh_vertadr[k] + m.mesh_vertnum[k]; v++)
        for (int c = 0; c < 3; c++) {
          const double x = m.mesh_vert[3 * v + c];
          lo[c] = x < lo[c] ? x : lo[c];
          hi[c] = x > hi[c] ? x : hi[c];
        }
      for (int c = 0; c < 3; c++) {
        m.geom_aabb[6 * g + c] = (lo[c] + hi[c]) * 0.5;
        m.geom_aabb[6 * g + 3 + c] = (hi[c] - lo[c]) * 0.5;
      }
    }
  }
  // candidate pairs in lexicographic order (model.rs:398-407)
  m.pair_g1.clear();
  m.pair_g2.clear();
  m.ngeneral = 0;
  m.ncmax = 0;
  for (int g1 = 0; g1 < ng; g1++)
    for (int g2 = g1 + 1; g2 < ng; g2++) {
      if (!pair_static_ok(m, g1, g2)) continue;
      if (needs_general(m.geom_type[g1], m.geom_type[g2])) {
        // cylinder-vs-non-plane and mesh pairs take the general convex
        // narrow phase (dgeneral.hpp)
        m.ngeneral++;
      }
      m.pair_g1.push_back(g1);
      m.pair_g2.push_back(g2);
      m.ncmax += pair_max_hits(m.geom_type[g1], m.geom_type[g2]);
    }
  // actuator contributions per dof, in (actuator, moment entry) order (forward.rs:207-209)
  m.dofact_adr.assign(nv, 0);
  m.dofact_num.assign(nv, 0);
  m.dofact_u.clear();
  m.dofact_coef.clear();
  for (int k = 0; k < nv; k++) {
    m.dofact_adr[k] = (int)m.dofact_u.size();
    for (int u = 0; u < nu; u++)
      for (int i = 0; i < m.actuator_moment_num[u]; i++) {
        int e = m.actuator_moment_adr[u] + i;
        if (m.jnt_dofadr[m.moment_jnt[e]] == k) {
          m.dofact_u.push_back(u);
          m.dofact_coef.push_back(m.moment_coef[e]);
        }
      }
    m.dofact_num[k] = (int)m.dofact_u.size() - m.dofact_adr[k];
  }
  m.nlimited = 0;
  for (int j = 0; j < nj; j++)
    if (m.jnt_limited[j] && (m.jnt_type[j] == PH_JNT_HINGE || m.jnt_type[j] == PH_JNT_SLIDE)) m.nlimited++;
  m.contact_hertz = s.contact_hertz > 0.0 ? s.contact_hertz : 0.25 / s.timestep;
  m.damped = has_implicit_terms(m, m.dof_damping.data(), m.actuator_bias.data());
  return true;
}

// Scalars, capacities and LDS layout of the device model (pointers are left null).
void init_dmodel_scalars(const HostModel& hm, int solver, int max_contacts, phk::DModel& M) {
  const ph_model_desc& d = hm.d;
  memset(&M, 0, sizeof(M));
  M.nq = d.nq;
  M.nv = d.nv;
  M.nu = d.nu;
  M.nbody = d.nbody;
  M.njnt = d.njnt;
  M.ngeom = d.ngeom;
  M.ncam = d.ncam;
  M.neq = d.neq;
  M.npair = (int)hm.pair_g1.size();
  M.ngeneral = hm.ngeneral;
  M.nlevel = (int)hm.level_adr.size() - 1;
  M.nacc = (int)hm.acc_level_adr.size() - 1;
  M.nblock = hm.block_adr.empty() ? 0 : (int)hm.block_adr.size() - 1;
  M.integrator = d.integrator;
  M.solver_kind = solver;
  M.iterations = d.solver_iterations;
  M.timestep = d.timestep;
  for (int k = 0; k < 3; k++) M.gravity[k] = d.gravity[k];
  M.tolerance = d.tolerance;
  M.margin = d.contact_margin;
  M.solref[0] = d.solref[0];
  M.solref[1] = d.solref[1];
  for (int k = 0; k < 5; k++) M.solimp[k] = d.solimp[k];
  M.impratio = d.impratio;
  M.hertz = hm.contact_hertz;
  M.zeta = d.contact_damping_ratio;
  M.slop = d.contact_slop;
  M.ncmax = std::max(hm.ncmax, 1);
  if (max_contacts > 0) M.ncmax = std::min(M.ncmax, max_contacts);
  M.nrmax = std::max(4 * M.ncmax + 2 * hm.nlimited + d.neq, 1);
  // island scratch (newton.rs:378-473), only when several dof blocks exist
  M.isl_fstride = M.nblock > 1 ? d.nv * d.nv + M.nrmax * (d.nv + 2) + 4 * d.nv : 0;
  M.isl_istride = M.nblock > 1 ? 2 * M.nrmax + d.nv + 2 * M.nblock + 1 : 0;  // ints, in LDS (Layout::i_isl)
  // Rows beyond nrl live in HBM. Keep up to 32 in LDS, fewer if that keeps
  // the per-env working set within 10 KiB (four 4-env workgroups per CU); one
  // row at a time, so that a step just past 16 rows still finds them in LDS
  // when the budget has room (Ant: 18).
  M.nrl = std::min(M.nrmax, 32);
  M.L = make_layout(hm, M.nrl, M.npair, M.ncmax, M.isl_istride);
  while (M.L.bytes > 10 * 1024 && M.nrl > 16) {
    M.nrl -= 1;
    M.L = make_layout(hm, M.nrl, M.npair, M.ncmax, M.isl_istride);
  }

}

// LDS layout of one env (offsets in doubles, then ints).
//
// Arrays live across a whole step come first. The rest share one span, by
// live range within a step (kinematics -> crba -> velocity/rne -> collision ->
// rows/Newton/integration):
//   S1 (kinematics .. collision): cinert, xpos, xmat
//   after S1, mutually exclusive in time:
//     kinematics temporaries: xquat, hinge quaternions
//     crba: composite inertias
//     velocity/rne: cvel, cacc, cfrc
//     collision: geom poses, world AABBs, per-pair hit counts and offsets
//   C (after collision, over S1 and all of the above): Newton vectors, the
//     mass factor (which the Newton Hessian reuses: the factor of M is consumed
//     by the unconstrained solve before Newton and rebuilt after it), the
//     touching-contact list, constraint rows (row positions and diagonals share
//     storage with J*a and J*p, which Newton fills only after the rows are built).
// Poses for export are recomputed by the final kinematics pass of a launch.
phk::Layout make_layout(const phh::HostModel& hm, int nrl, int npair, int ncmax, int isl_ints) {
  phk::Layout L;
  const ph_model_desc& d = hm.d;
  int nq = d.nq, nv = d.nv, nu = d.nu, nb = d.nbody, ng = d.ngeom;
  int o = 0;
  auto take = [&](int n) {
    int r = o;
    o += n;
    return r;
  };
  auto ints = [](int n) { return (n + 1) / 2; };  // int32 arrays in double slots
  L.qpos = take(nq);
  L.qvel = take(nv);
  L.ctrl = take(nu);
  L.qacc = take(nv);
  L.applied = take(nv);
  L.xipos = take(3 * nb);
  L.cdof = take(6 * nv);
  L.qm = take(nv * nv);
  L.bias = take(nv);
  L.passive = take(nv);
  L.qfrc_act = take(nv);
  L.actf = take(nu);
  L.actclamp = take(nu);
  L.smooth = take(nv);
  L.qfrc_c = take(nv);
  L.acc = take(nv);
  L.vel = take(nv);
  L.scal = take(16);
  const int span = o;
  L.cinert = take(13 * nb);
  L.xpos = take(3 * nb);
  L.xmat = take(9 * nb);
  const int r1 = o;
  int end = o;
  o = r1;  // kinematics temporaries
  L.xquat = take(4 * nb);
  L.jq = take(4 * d.njnt);
  end = std::max(end, o);
  o = r1;  // crba
  L.crb = take(13 * nb);
  end = std::max(end, o);
  o = r1;  // velocity / rne
  L.cvel = take(6 * nb);
  L.cacc = take(6 * nb);
  L.cfrc = take(6 * nb);
  end = std::max(end, o);
  o = r1;  // collision
  L.gxpos = take(3 * ng);
  L.gxmat = take(9 * ng);
  L.aabb = take(6 * ng);
  L.paircnt = take(ints(npair));
  L.pairoff = take(ints(npair));
  end = std::max(end, o);
  o = span;  // C
  L.a = take(nv);
  L.a0 = take(nv);
  L.grad = take(nv);
  L.p = take(nv);
  L.tmp = take(nv);
  L.tmp2 = take(nv);
  L.mp = take(nv);
  L.chol = take(nv * nv);
  L.H = L.chol;
  L.cact = take(ints(ncmax));
  L.J = take(nrl * nv);
  L.aref = take(nrl);
  L.dd = take(nrl);
  L.jar = take(nrl);
  L.jp = take(nrl);
  L.rpos = L.jar;
  L.rdiag = L.jp;
  L.ract = take(nrl);
  L.rprev = take(nrl);
  L.rf = take(nrl);
  o = std::max(end, o);
  L.ndouble = o;
  int oi = 0;
  L.i_iscal = oi;
  oi += 16;
  L.i_isl = oi;
  oi += isl_ints;
  L.nint = oi;
  // 16-byte aligned so that envs can be packed side by side in one workgroup
  L.bytes = (int)((L.ndouble * sizeof(double) + L.nint * sizeof(int) + 15) / 16 * 16);
  return L;
}

}  // namespace phh
