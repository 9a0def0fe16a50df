// ORACLE — test infrastructure only (see math.hpp). A C++ restatement of
// Pippin's FP64 CPU engine: crates/pippin/src/{model,data,forward,collision,
// newton,solver,params,batch}.rs. Each function cites the lines it follows.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "math.hpp"

struct ph_model_desc;

namespace orc {

static const size_t NONE = (size_t)-1;  // usize::MAX

enum JointType { Free = 0, Ball = 1, Slide = 2, Hinge = 3 };
enum GeomType { Plane = 0, Sphere = 1, Capsule = 2, Box = 3, Cylinder = 4, Mesh = 5 };
enum Integrator { Euler = 0, ImplicitFast = 1 };
enum SolverKind { Newton = 0, Pgs = 1 };

static inline size_t jnt_nq(int t) { return t == Free ? 7 : (t == Ball ? 4 : 1); }
static inline size_t jnt_nv(int t) { return t == Free ? 6 : (t == Ball ? 3 : 1); }

// model.rs:59-100
struct SolverOptions {
  int kind = Newton;
  size_t iterations = 30;
  Real tolerance = 1e-8;
  Real contact_margin = 1e-3;
  Real solref[2] = {0.02, 1.0};
  Real solimp[5] = {0.9, 0.95, 0.001, 0.5, 2.0};
  Real impratio = 1.0;
  Real contact_hertz = 0.0;
  Real contact_damping_ratio = 1.0;
  Real contact_slop = 5e-4;
};

struct Range {
  size_t start, end;
};

struct MomentEntry {
  size_t jnt;
  Real coef;
};

// A mesh's convex hull as the narrow phase reads it (Model::mesh_hull,
// model.rs:176-182, and the support queries of fasthull.rs): vertices in the
// geom frame, faces (vertex lists, counter-clockwise from outside), the unit
// normal of each face, and the faces around each vertex in increasing order.
struct HullShape {
  std::vector<Vec3> vert;
  std::vector<std::vector<size_t>> faces;
  std::vector<Vec3> normal;
  std::vector<std::vector<size_t>> vert_faces;
};

// model.rs:102-249 (names and keyframes are not needed by the engine).
struct Model {
  Real timestep = 0.002;
  Vec3 gravity = {{0, 0, -9.81}};
  int integrator = Euler;
  SolverOptions solver;
  size_t nq = 0, nv = 0, nu = 0;

  std::vector<size_t> body_parent, body_jntadr, body_jntnum, body_dofadr, body_dofnum, body_lastdof;
  std::vector<Vec3> body_pos, body_ipos;
  std::vector<Quat> body_quat;
  std::vector<Real> body_mass;
  std::vector<Mat3> body_inertia;
  std::vector<size_t> body_weldid;

  std::vector<int> jnt_type;
  std::vector<size_t> jnt_body, jnt_qposadr, jnt_dofadr;
  std::vector<Vec3> jnt_pos, jnt_axis;
  std::vector<char> jnt_limited;
  std::vector<Real> jnt_range_lo, jnt_range_hi, jnt_stiffness, jnt_springref;

  std::vector<size_t> dof_body, dof_jnt, dof_parent;
  std::vector<Real> dof_damping, dof_armature;

  std::vector<int> geom_type;
  std::vector<size_t> geom_body, geom_dataid;
  std::vector<Vec3> geom_pos, geom_size;
  std::vector<Quat> geom_quat;
  std::vector<Real> geom_friction, geom_rbound;
  std::vector<uint32_t> geom_contype, geom_conaffinity;
  std::vector<float> geom_rgba;  // ngeom x 4
  std::vector<std::pair<Vec3, Vec3>> geom_aabb;
  std::vector<HullShape> mesh_hull;

  std::vector<size_t> cam_body;
  std::vector<Vec3> cam_pos;
  std::vector<Quat> cam_quat;

  std::vector<Real> actuator_gear, actuator_gain;
  std::vector<Vec3> actuator_bias;  // [b0, b1, b2]
  std::vector<char> actuator_ctrllimited, actuator_forcelimited;
  std::vector<Real> actuator_ctrlrange_lo, actuator_ctrlrange_hi, actuator_forcerange_lo, actuator_forcerange_hi;
  std::vector<std::vector<MomentEntry>> actuator_moment;

  std::vector<size_t> eq_joint1, eq_joint2;
  std::vector<std::vector<Real>> eq_polycoef, eq_solref, eq_solimp;

  std::vector<Range> dof_blocks;
  std::vector<std::pair<size_t, size_t>> exclude_pairs;

  std::vector<std::pair<Real, Real>> body_invweight;  // (trans, rot)
  std::vector<Real> dof_invweight;
  std::vector<Real> qpos0, qpos_reset;
  std::vector<std::pair<size_t, size_t>> collision_pairs;

  size_t nbody() const { return body_parent.size(); }
  size_t njnt() const { return jnt_type.size(); }
  size_t ngeom() const { return geom_type.size(); }
};

// collision.rs:9-43
struct Contact {
  Vec3 pos, normal, tangent;
  Real depth;
  size_t geom[2];
  uint32_t feature;
};
struct Hit {
  Vec3 pos, normal;
  Real depth;
  Vec3 tangent;
};
struct GeomPose {
  int typ;
  Vec3 pos;
  Mat3 mat;
  Vec3 size;
  const HullShape* hull;  // mesh geoms only
};

// newton.rs:34-49
struct NRow {
  Real aref = 0, d = 0;
  bool bilateral = false;
  bool active(Real r) const { return bilateral || r < 0.0; }
};

// solver.rs:16-32
enum RowKindTag { RNormal = 0, RFriction = 1, RBilateral = 2 };
struct Row {
  int kind;
  Real sep;      // Normal/Bilateral
  size_t normal;  // Friction
  Real mu;       // Friction
  Real inv_a;
  size_t contact;
};

// data.rs:8-167 (fields the engine reads or writes).
struct Data {
  Real time = 0;
  std::vector<Real> qpos, qvel, ctrl, qfrc_applied;
  std::vector<Vec3> xpos, xipos, xanchor, xaxis, geom_xpos;
  std::vector<Quat> xquat;
  std::vector<Mat3> xmat, geom_xmat;
  std::vector<Spatial> cdof, cdof_dot, cvel;
  std::vector<SpatialInertia> cinert, crb;
  std::vector<Real> qm, qm_chol, qfrc_bias, qfrc_passive, qfrc_actuator, qfrc_constraint, qacc, actuator_force;
  std::vector<char> actuator_clamped;
  std::vector<Contact> contacts;
  // scratch (data.rs:146-167)
  std::vector<Contact> prev_contacts;
  std::vector<std::vector<Real>> prev_lambda;  // [3] per contact
  size_t newton_iters = 0;
  // per-step counters used only by the op-count instrumentation
  size_t ls_steps = 0;
  size_t nrows = 0;
  std::vector<Spatial> cacc, cfrc;
};

// ---- model.rs ----
void compile(Model& m);  // model.rs:279-316
void rebuild_collision_pairs(Model& m);
std::pair<Vec3, Vec3> local_aabb(const Model& m, size_t g);
Real rbound(const Model& m, size_t g);
Model model_from_desc(const ph_model_desc& d);

// ---- data.rs ----
Data data_new(const Model& m);
void data_reset(Data& d, const Model& m);

// ---- forward.rs ----
void kinematics(const Model& m, Data& d);
void crba(const Model& m, Data& d);
void velocity(const Model& m, Data& d);
void rne(const Model& m, Data& d);
void passive(const Model& m, Data& d);
void actuation(const Model& m, Data& d);
void collide(const Model& m, Data& d);
void forward(const Model& m, Data& d);
void set_const(Model& m);
// Returns false when the mass matrix is not positive definite (the reference panics).
bool step(const Model& m, Data& d);
void mass_solve(const Model& m, const std::vector<Real>& l, Real* x);
void integrate_pos(const Model& m, Real* qpos, const Real* qvel, Real h);

// ---- collision.rs ----
void collide_pair(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out);
bool pair_supported(int a, int b);
bool needs_general(int a, int b);
// General convex pairs (collision.rs:124-147; parry's method restated in general.cpp).
void collide_general(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out);
// Keeps the deepest point and three spread around it (collision.rs:514-531).
void reduce_to_four(std::vector<std::pair<Real, Vec3>>& pts);
void contact_frame(const Contact& c, Vec3* t1, Vec3* t2);

// ---- newton.rs / solver.rs ----
Real impedance(const Real solimp[5], Real pos);
void newton_solve(const Model& m, Data& d, Real* a);
void pgs_solve(const Model& m, Data& d, Real* v, std::vector<Real>& v_pos);
void point_jac(const Model& m, const std::vector<Spatial>& cdof, size_t body, Vec3 p, Vec3 dir, Real sign,
               Real* row);

// ---- params.rs ----
int param_width(int p, const Model& m);
size_t param_count(int p, const Model& m);
std::vector<Real> get_param(const Model& m, int p, size_t id);
bool set_param(Model& m, int p, size_t id, const Real* v, size_t n, std::string* err);

}  // namespace orc
