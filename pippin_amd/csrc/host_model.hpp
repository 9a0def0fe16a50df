// Host-side model: a private copy of the raw descriptor plus the derived data
// that crates/pippin/src/model.rs::compile computes (dof tree, dof blocks, weld
// ids, local bounding boxes, collision pairs), rearranged into the flat arrays
// the device kernels read.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pippin_hip.h"
#include "dmodel.hpp"

namespace phh {

struct HostModel {
  // raw fields (owned copies of ph_model_desc arrays)
  ph_model_desc d;  // scalar fields; pointers re-targeted at the vectors below
  std::vector<int32_t> body_parent, body_jntadr, body_jntnum, body_dofadr, body_dofnum, jnt_type, jnt_body,
      jnt_qposadr, jnt_dofadr, jnt_limited, dof_body, dof_jnt, geom_type, geom_body, geom_group, geom_dataid, cam_body,
      actuator_ctrllimited, actuator_forcelimited, actuator_moment_adr, actuator_moment_num, moment_jnt, eq_joint1,
      eq_joint2, exclude_pairs;
  std::vector<uint32_t> geom_contype, geom_conaffinity;
  std::vector<float> geom_rgba;
  std::vector<double> body_pos, body_quat, body_mass, body_ipos, body_inertia, jnt_pos, jnt_axis, jnt_range,
      jnt_stiffness, jnt_springref, dof_damping, dof_armature, geom_pos, geom_quat, geom_size, geom_friction, cam_pos,
      cam_quat, cam_fovy, actuator_gear, actuator_gain, actuator_bias, actuator_ctrlrange, actuator_forcerange,
      moment_coef, eq_polycoef, eq_solref, eq_solimp, qpos0, qpos_reset;

  // mesh hulls (raw), then what the narrow phase reads: a unit normal per face
  // (cross(v1 - v0, v2 - v0) over its length) and the faces around each hull
  // vertex (global vertex index), in increasing face order
  std::vector<int32_t> mesh_vertadr, mesh_vertnum, mesh_faceadr, mesh_facenum, face_vertadr, face_vertnum, face_vert;
  std::vector<double> mesh_vert, face_normal;
  std::vector<int32_t> vert_faceadr, vert_facenum, vert_face;

  // derived (model.rs:279-316)
  std::vector<int32_t> body_lastdof, dof_parent, body_weldid, block_adr, dof_block;
  std::vector<uint64_t> body_chain, dof_anc;
  std::vector<int32_t> level_adr, level_body;
  // leaf-to-root accumulation schedule (crba / rne): levels deepest first;
  // each level lists (parent, children in descending index order) groups
  std::vector<int32_t> acc_level_adr, acc_parent, acc_child_adr, acc_child;
  std::vector<double> geom_aabb;  // ngeom x 6 (center, half-extents)
  std::vector<int32_t> pair_g1, pair_g2;  // candidate pairs: every static filter except the masks
  int ngeneral = 0;  // candidate pairs that take the general convex narrow phase
  std::vector<int32_t> dofact_adr, dofact_num, dofact_u;
  std::vector<double> dofact_coef;
  double contact_hertz = 0;
  int ncmax = 0;   // sum of per-pair hit maxima over candidate pairs
  int nlimited = 0;
  bool damped = false;
};

// Copy `desc` and compute the derived data. Returns false with `err` set if
// the model uses something the device path does not implement.
bool build_host_model(const ph_model_desc& desc, HostModel& m, std::string* err);
void retarget(HostModel& m);  // point m.d's arrays at m's vectors
// Local bounding box (model.rs:366-382) of a primitive of type t and size s.
void local_aabb(int t, const double* s, double* out6);
bool has_implicit_terms(const HostModel& m, const double* dof_damping, const double* act_bias);
// Candidate-pair filter pieces (model.rs:430-452), masks excluded.
bool pair_static_ok(const HostModel& m, int g1, int g2);
// LDS layout for one env: nrl rows kept in LDS, npair candidate pairs, ncmax contacts.
phk::Layout make_layout(const HostModel& m, int nrl, int npair, int ncmax, int isl_ints);
void init_dmodel_scalars(const HostModel& hm, int solver, int max_contacts, phk::DModel& M);

}  // namespace phh
