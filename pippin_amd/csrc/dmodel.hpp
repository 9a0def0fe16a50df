// Device-side model and state descriptors shared by the kernels and the host
// code that fills them. Everything is plain pointers into device memory plus
// sizes, passed to kernels by value.
#pragma once
#include <stdint.h>
#include <string.h>

#include <cstring>

#include "dmath.hpp"
#include "mlp.hpp"

namespace phk {

// Pointers into device buffers. In device code every access goes through an
// explicit address space, never a generic (flat) pointer: flat accesses also
// count against the LDS wait counter, which would make every LDS wait also
// wait for HBM. Read-only model data (never written while a kernel runs) is
// read through the constant address space, so loads at wave-uniform indices
// become scalar loads; state buffers are read and written as global memory.
// On the host both are plain pointers; the layout is one pointer either way.
#if defined(__HIP_DEVICE_COMPILE__)
#define PH_AS_GLOBAL __attribute__((address_space(1)))
#define PH_AS_CONST __attribute__((address_space(4)))
#else
#define PH_AS_GLOBAL
#define PH_AS_CONST
#endif
template <class T>
struct gp {
  T* p;
  gp& operator=(T* q) {
    p = q;
    return *this;
  }
  PHD PH_AS_GLOBAL T& operator[](int64_t i) const { return ((PH_AS_GLOBAL T*)p)[i]; }
  PHD gp operator+(int64_t i) const {
    gp r;
    r.p = p + i;
    return r;
  }
};
template <class T>
struct gp<const T> {
  const T* p;
  gp& operator=(const T* q) {
    p = q;
    return *this;
  }
  PHD T operator[](int64_t i) const { return ((const PH_AS_CONST T*)p)[i]; }
  PHD gp operator+(int64_t i) const {
    gp r;
    r.p = p + i;
    return r;
  }
};

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
  int i_isl;  // Newton island bookkeeping (row block masks, gathered indices, islands); several dof blocks only
  int nint;
  int bytes;  // per env, 16-byte aligned
};

inline bool layout_equal(const Layout& a, const Layout& b) { return std::memcmp(&a, &b, sizeof(Layout)) == 0; }

struct DModel {
  int nq, nv, nu, nbody, njnt, ngeom, ncam, neq, npair, nlevel, nblock, nacc;
  int ngeneral;  // candidate pairs that take the general convex narrow phase (dgeneral.hpp)
  int integrator, solver_kind, iterations;
  Real timestep, gravity[3], tolerance, margin, solref[2], solimp[5], impratio, hertz, zeta, slop;
  int ncmax;  // contact capacity per env (the sum of per-pair maxima: never exceeded)
  int nrmax;  // row capacity per env (4 per contact for Newton, 3 for PGS, + limits + equalities)
  int nrl;    // rows kept in LDS; larger problems use the global row buffers
  int hess_mfma;  // 1: Newton's dense Hessian J'DJ on the FP64 matrix cores (one env per wavefront only)

  gp<const int> body_parent, body_jntadr, body_jntnum, body_dofadr, body_dofnum, body_lastdof;
  gp<const uint64_t> body_chain;  // dofs on the path world..body (bit k = dof k)
  gp<const Real> body_pos, body_quat;
  gp<const int> level_adr, level_body;  // bodies grouped by depth (parents first)
  gp<const int> acc_level_adr, acc_parent, acc_child_adr, acc_child;  // leaf-to-root sum schedule
  gp<const int> jnt_type, jnt_qposadr, jnt_dofadr, jnt_limited;
  gp<const Real> jnt_pos, jnt_axis, jnt_range, jnt_stiffness, jnt_springref;
  gp<const int> dof_body, dof_parent, dof_jnt;
  gp<const uint64_t> dof_anc;  // strict ancestors of each dof
  gp<const Real> dof_armature;
  gp<const int> block_adr;  // nblock + 1 dof boundaries
  gp<const int> dof_block;  // block of each dof
  int isl_fstride, isl_istride;  // per-env island scratch (doubles, ints); 0 with one block
  gp<const int> geom_type, geom_body;
  gp<const Real> geom_pos, geom_quat;
  gp<const int> pair_g1, pair_g2;
  gp<const Real> act_gear;
  gp<const int> act_ctrllimited, act_forcelimited;
  gp<const Real> act_ctrlrange, act_forcerange;
  gp<const int> act_madr, act_mnum, mom_jnt;
  gp<const Real> mom_coef;
  gp<const int> dofact_adr, dofact_num, dofact_u;  // per dof: actuator contributions in (u, moment) order
  gp<const Real> dofact_coef;
  gp<const int> eq_j1, eq_j2;
  gp<const Real> eq_poly, eq_solref, eq_solimp;
  gp<const Real> qpos0;
  // convex hulls of mesh geoms (geom_dataid: mesh index or -1)
  gp<const int> geom_dataid;
  gp<const int> mesh_vertadr, mesh_vertnum;  // per mesh: first hull vertex, count
  gp<const Real> mesh_vert;                  // hull vertices, geom frame, 3 per vertex
  gp<const Real> face_normal;                // unit normal per face, 3 per face
  gp<const int> face_vertadr, face_vertnum, face_vert;  // face vertex lists (ids within the face's mesh)
  gp<const int> vert_faceadr, vert_facenum, vert_face;  // per hull vertex: faces around it, increasing

  // per-variant parameters (params.rs), variant of env e = env_var[e]
  gp<const int> env_var;
  gp<const Real> v_geom_size, v_geom_aabb, v_friction;
  gp<const uint32_t> v_contype, v_conaffinity;
  gp<const Real> v_body_mass, v_body_ipos, v_body_inertia, v_dof_damping, v_act_gain, v_act_bias;
  gp<const Real> v_dof_invweight, v_body_invweight, v_qpos_reset;
  gp<const int> v_damped;

  Layout L;
  MlpW mlp;  // learned warm start (mode 0: off)
};

// Per-env state and scratch in global memory; env e owns slice e of each.
struct DState {
  int64_t n;
  gp<Real> qpos, qvel, ctrl, qacc, applied, time;
  gp<Real> geom_xpos, geom_xmat, xpos, xmat;  // poses after the last step
  gp<int> ncon, iters, nrows, status;
  gp<int> cost;  // Newton iterations summed over the env's last launch (bucketing key)
  gp<Real> con;       // n x ncmax x CW
  gp<int> cgeom;      // n x ncmax x 3 (g1, g2, feature)
  gp<Real> pcon;      // PGS history: previous contacts
  gp<int> pcgeom;
  gp<Real> plambda;   // n x ncmax x 3
  gp<int> npcon;
  gp<Real> stage;     // n x npair x MAXHIT x CW
  gp<Real> gjac;      // n x nrmax x nv
  gp<Real> grow;      // n x nrmax x 10
  gp<Real> gmj;       // n x nrmax x nv: PGS M^-1 J^T rows
  gp<Real> isl_f;     // n x isl_fstride: gathered island problems
  // learned warm start: per env, the first MLP_CL touching contacts of the last step
  gp<float> mlp_feat;  // n x MLP_CL x MLP_NF features
  gp<float> mlp_pred;  // n x MLP_CL x MLP_NO predicted log1p(lambda / D)
  gp<float> mlp_tgt;   // n x MLP_CL x MLP_NO solved log1p(lambda / D) (logging on)
  gp<int> mlp_cnt;     // n: contacts recorded
};

}  // namespace phk
