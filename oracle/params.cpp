// ORACLE — test infrastructure only (see math.hpp).
// Runtime model parameters: crates/pippin/src/params.rs.
#include <cstdio>

#include "../include/pippin_hip.h"
#include "engine.hpp"

namespace orc {

static const char* param_name(int p) {
  static const char* names[] = {"GeomFriction", "GeomSize", "GeomRgba", "GeomContype", "GeomConaffinity", "BodyMass",
                                "DofDamping", "ActuatorGain", "ActuatorBias", "BodyInertia", "BodyIpos", "Qpos0"};
  return (p >= 0 && p < 12) ? names[p] : "?";
}

// params.rs:61-72
int param_width(int p, const Model& m) {
  switch (p) {
    case PH_PARAM_GEOM_SIZE:
    case PH_PARAM_ACTUATOR_BIAS:
    case PH_PARAM_BODY_INERTIA:
    case PH_PARAM_BODY_IPOS: return 3;
    case PH_PARAM_GEOM_RGBA: return 4;
    case PH_PARAM_QPOS0: return (int)m.nq;
    default: return 1;
  }
}

// params.rs:75-83
size_t param_count(int p, const Model& m) {
  switch (p) {
    case PH_PARAM_GEOM_FRICTION:
    case PH_PARAM_GEOM_SIZE:
    case PH_PARAM_GEOM_RGBA:
    case PH_PARAM_GEOM_CONTYPE:
    case PH_PARAM_GEOM_CONAFFINITY: return m.ngeom();
    case PH_PARAM_BODY_MASS:
    case PH_PARAM_BODY_INERTIA:
    case PH_PARAM_BODY_IPOS: return m.nbody();
    case PH_PARAM_QPOS0: return 1;
    case PH_PARAM_DOF_DAMPING: return m.nv;
    case PH_PARAM_ACTUATOR_GAIN:
    case PH_PARAM_ACTUATOR_BIAS: return m.nu;
    default: return 0;
  }
}

// params.rs:93-108
std::vector<Real> get_param(const Model& m, int p, size_t id) {
  switch (p) {
    case PH_PARAM_GEOM_FRICTION: return {m.geom_friction[id]};
    case PH_PARAM_GEOM_SIZE: return {m.geom_size[id][0], m.geom_size[id][1], m.geom_size[id][2]};
    case PH_PARAM_GEOM_RGBA:
      return {(Real)m.geom_rgba[4 * id], (Real)m.geom_rgba[4 * id + 1], (Real)m.geom_rgba[4 * id + 2],
              (Real)m.geom_rgba[4 * id + 3]};
    case PH_PARAM_GEOM_CONTYPE: return {(Real)m.geom_contype[id]};
    case PH_PARAM_GEOM_CONAFFINITY: return {(Real)m.geom_conaffinity[id]};
    case PH_PARAM_BODY_MASS: return {m.body_mass[id]};
    case PH_PARAM_DOF_DAMPING: return {m.dof_damping[id]};
    case PH_PARAM_ACTUATOR_GAIN: return {m.actuator_gain[id]};
    case PH_PARAM_ACTUATOR_BIAS: return {m.actuator_bias[id][0], m.actuator_bias[id][1], m.actuator_bias[id][2]};
    case PH_PARAM_BODY_INERTIA: return {m.body_inertia[id][0], m.body_inertia[id][4], m.body_inertia[id][8]};
    case PH_PARAM_BODY_IPOS: return {m.body_ipos[id][0], m.body_ipos[id][1], m.body_ipos[id][2]};
    case PH_PARAM_QPOS0: return m.qpos_reset;
    default: return {};
  }
}

// Rust `f64 as u32` saturates: NaN -> 0, negatives -> 0, > u32::MAX -> u32::MAX.
static uint32_t sat_u32(Real xr) {
  double x = r_val(xr);
  if (!(x > 0.0)) return 0;
  if (x >= 4294967295.0) return 4294967295u;
  return (uint32_t)x;
}

// params.rs:111-177
bool set_param(Model& m, int p, size_t id, const Real* v, size_t n, std::string* err) {
  char buf[256];
  if (id >= param_count(p, m)) {
    snprintf(buf, sizeof buf, "%s: id %zu out of range (have %zu)", param_name(p), id, param_count(p, m));
    *err = buf;
    return false;
  }
  if (n != (size_t)param_width(p, m)) {
    snprintf(buf, sizeof buf, "%s: expected %d values, got %zu", param_name(p), param_width(p, m), n);
    *err = buf;
    return false;
  }
  for (size_t i = 0; i < n; i++)
    if (!r_isfinite(v[i])) {
      snprintf(buf, sizeof buf, "%s: values must be finite", param_name(p));
      *err = buf;
      return false;
    }
  switch (p) {
    case PH_PARAM_GEOM_FRICTION: m.geom_friction[id] = rmax(v[0], 0.0); break;
    case PH_PARAM_GEOM_SIZE:
      if (m.geom_type[id] == Mesh || m.geom_type[id] == Plane) {
        *err = "geom_size applies to sphere, capsule, box, and cylinder geoms";
        return false;
      }
      for (int i = 0; i < 3; i++)
        if (v[i] < 0.0) {
          *err = "geom_size must be non-negative";
          return false;
        }
      m.geom_size[id] = V3(v[0], v[1], v[2]);
      m.geom_rbound[id] = rbound(m, id);
      m.geom_aabb[id] = local_aabb(m, id);
      break;
    case PH_PARAM_GEOM_RGBA:
      for (int i = 0; i < 4; i++) m.geom_rgba[4 * id + i] = (float)v[i];
      break;
    case PH_PARAM_GEOM_CONTYPE:
    case PH_PARAM_GEOM_CONAFFINITY: {
      uint32_t bits = sat_u32(rmax(v[0], 0.0));
      if (p == PH_PARAM_GEOM_CONTYPE) m.geom_contype[id] = bits;
      else m.geom_conaffinity[id] = bits;
      rebuild_collision_pairs(m);
      break;
    }
    case PH_PARAM_BODY_MASS: {
      if (id == 0) {
        *err = "the world body has no mass";
        return false;
      }
      if (v[0] <= 0.0 && m.body_jntnum[id] > 0) {
        *err = "moving bodies need positive mass";
        return false;
      }
      Real scale = m.body_mass[id] > 0.0 ? v[0] / m.body_mass[id] : 1.0;
      m.body_mass[id] = v[0];
      m.body_inertia[id] = mscale(m.body_inertia[id], scale);
      set_const(m);
      break;
    }
    case PH_PARAM_DOF_DAMPING: m.dof_damping[id] = rmax(v[0], 0.0); break;
    case PH_PARAM_ACTUATOR_GAIN: m.actuator_gain[id] = v[0]; break;
    case PH_PARAM_ACTUATOR_BIAS: m.actuator_bias[id] = V3(v[0], v[1], v[2]); break;
    case PH_PARAM_BODY_INERTIA:
      if (id == 0 || v[0] < 0.0 || v[1] < 0.0 || v[2] < 0.0) {
        *err = "body_inertia needs a non-world body and non-negative moments";
        return false;
      }
      m.body_inertia[id] = m3diag(V3(v[0], v[1], v[2]));
      set_const(m);
      break;
    case PH_PARAM_BODY_IPOS:
      if (id == 0) {
        *err = "the world body has no center of mass";
        return false;
      }
      m.body_ipos[id] = V3(v[0], v[1], v[2]);
      set_const(m);
      break;
    case PH_PARAM_QPOS0: m.qpos_reset.assign(v, v + n); break;
    default: *err = "unknown parameter"; return false;
  }
  return true;
}

}  // namespace orc
