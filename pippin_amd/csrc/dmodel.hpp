// Device-side model and state descriptors shared by the kernels and the host
// code that fills them. Everything is plain pointers into device memory plus
// sizes, passed to kernels by value.
#pragma once
#include <stdint.h>

#include "dmath.hpp"

namespace phk {

enum { J_FREE = 0, J_BALL = 1, J_SLIDE = 2, J_HINGE = 3 };
enum { G_PLANE = 0, G_SPHERE = 1, G_CAPSULE = 2, G_BOX = 3, G_CYLINDER = 4, G_MESH = 5 };

// Contact record in global memory: pos(3) normal(3) tangent(3) depth(1).
constexpr int CW = 10;
// Hits one geom pair can produce at most (plane-box, box-box, plane-cylinder).
constexpr int MAXHIT = 4;
// Status bits per env.
enum { ST_NOTPD = 1, ST_CONTACT_OVERFLOW = 2, ST_ROW_OVERFLOW = 4, ST_UNSUPPORTED = 8 };

// LDS layout: offsets (in doubles) of each per-env array inside one
// workgroup's dynamic shared memory, then an int region.
struct Layout {
  int qpos, qvel, ctrl, qacc, applied;
  int xpos, xquat, xmat, xipos, jq;
  int cdof, cinert, crb, cvel, cacc, cfrc;
  int qm, chol, H;
  int bias, passive, qfrc_act, actf, actclamp, smooth, qfrc_c, acc, vel;
  int gxpos, gxmat, aabb;
  int paircnt, pairoff, cact;  // int arrays stored in double slots
  int a, a0, grad, p, tmp, tmp2, mp;
  int J, aref, dd, jar, jp, rpos, rdiag, ract, rprev, rf;
  int scal;     // 16 doubles of per-env scalars
  int ndouble;  // total doubles
  // int region (offsets in ints, after the doubles)
  int i_iscal;
  int nint;
  int bytes;  // per env, 16-byte aligned
};

struct DModel {
  int nq, nv, nu, nbody, njnt, ngeom, ncam, neq, npair, nlevel, nblock;
  int integrator, solver_kind, iterations;
  Real timestep, gravity[3], tolerance, margin, solref[2], solimp[5], impratio, hertz, zeta, slop;
  int ncmax;  // contact capacity per env (the sum of per-pair maxima: never exceeded)
  int nrmax;  // row capacity per env (4 per contact for Newton, 3 for PGS, + limits + equalities)
  int nrl;    // rows kept in LDS; larger problems use the global row buffers

  const int *body_parent, *body_jntadr, *body_jntnum, *body_dofadr, *body_dofnum, *body_lastdof;
  const uint64_t* body_chain;  // dofs on the path world..body (bit k = dof k)
  const Real *body_pos, *body_quat;
  const int *level_adr, *level_body;  // bodies grouped by depth (parents first)
  const int *jnt_type, *jnt_qposadr, *jnt_dofadr, *jnt_limited;
  const Real *jnt_pos, *jnt_axis, *jnt_range, *jnt_stiffness, *jnt_springref;
  const int *dof_body, *dof_parent, *dof_jnt;
  const uint64_t* dof_anc;  // strict ancestors of each dof
  const Real* dof_armature;
  const int* block_adr;  // nblock + 1 dof boundaries
  const int *geom_type, *geom_body;
  const Real *geom_pos, *geom_quat;
  const int *pair_g1, *pair_g2;
  const Real* act_gear;
  const int *act_ctrllimited, *act_forcelimited;
  const Real *act_ctrlrange, *act_forcerange;
  const int *act_madr, *act_mnum, *mom_jnt;
  const Real* mom_coef;
  const int *dofact_adr, *dofact_num, *dofact_u;  // per dof: actuator contributions in (u, moment) order
  const Real* dofact_coef;
  const int *eq_j1, *eq_j2;
  const Real *eq_poly, *eq_solref, *eq_solimp;
  const Real* qpos0;

  // per-variant parameters (params.rs), variant of env e = env_var[e]
  const int* env_var;
  const Real *v_geom_size, *v_geom_aabb, *v_friction;
  const uint32_t *v_contype, *v_conaffinity;
  const Real *v_body_mass, *v_body_ipos, *v_body_inertia, *v_dof_damping, *v_act_gain, *v_act_bias;
  const Real *v_dof_invweight, *v_body_invweight, *v_qpos_reset;
  const int* v_damped;

  Layout L;
};

// Per-env state and scratch in global memory; env e owns slice e of each.
struct DState {
  int64_t n;
  Real *qpos, *qvel, *ctrl, *qacc, *applied, *time;
  Real *geom_xpos, *geom_xmat, *xpos, *xmat;  // poses after the last step
  int *ncon, *iters, *nrows, *status;
  Real* con;       // n x ncmax x CW
  int* cgeom;      // n x ncmax x 3 (g1, g2, feature)
  Real* pcon;      // PGS history: previous contacts
  int* pcgeom;
  Real* plambda;   // n x ncmax x 3
  int* npcon;
  Real* stage;     // n x npair x MAXHIT x CW
  Real* gjac;      // n x nrmax x nv
  Real* grow;      // n x nrmax x 10
};

}  // namespace phk
