// The batched physics step, written once against an execution policy X:
//   X::par(n, f)  runs f(i) for i in [0, n) spread over the lanes of the
//                 env's wavefront, then makes the writes visible to all lanes;
//   X::lane()     this lane's index (0 on the serial policy).
// The device policy maps one environment to one 64-lane wavefront (one
// workgroup), so every loop whose trip count depends on the env's state
// (contacts, rows, Newton iterations, line search) is wave-uniform.
//
// Stages restate crates/pippin/src/forward.rs::step and its callees. Work is
// split across lanes only along axes where the reference's floating-point
// order is per element (bodies of one tree level, spatial-inertia components,
// dofs, pairs, matrix entries, Cholesky rows), so each lane performs exactly
// the reference's sequence of operations for the values it owns.
#pragma once
#include <type_traits>
#include "dgeneral.hpp"
#include "lanes.hpp"

namespace phk {

struct Env {
  Real* s;  // LDS doubles
  int* si;  // LDS ints
  int64_t e;
  int var;
  // row storage for the current solve (LDS when it fits, else global)
  Real *J, *aref, *dd, *jar, *jp, *rpos, *rdiag, *ract, *rprev, *rf;
};

#define LD(name) (E.s + M.L.name)

template <class X, class MT>
PHD void rows_bind(const MT& M, const DState& S, Env& E, int nr) {
  if (nr <= M.nrl) {
    E.J = LD(J);
    E.aref = LD(aref);
    E.dd = LD(dd);
    E.jar = LD(jar);
    E.jp = LD(jp);
    E.rpos = LD(rpos);
    E.rdiag = LD(rdiag);
    E.ract = LD(ract);
    E.rprev = LD(rprev);
    E.rf = LD(rf);
  } else {
    E.J = (S.gjac + E.e * (int64_t)M.nrmax * M.nv).p;
    Real* g = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
    E.aref = g;
    E.dd = g + M.nrmax;
    E.jar = g + 2 * M.nrmax;
    E.jp = g + 3 * M.nrmax;
    E.rpos = g + 4 * M.nrmax;
    E.rdiag = g + 5 * M.nrmax;
    E.ract = g + 6 * M.nrmax;
    E.rprev = g + 7 * M.nrmax;
    E.rf = g + 8 * M.nrmax;
  }
}

// ---------------- kinematics, forward.rs:14-90 ----------------
template <class X, class MT>
PHD void fk_body(const MT& M, Env& E, int b) {
  Real* xpos = LD(xpos);
  Real* xquat = LD(xquat);
  Real* xmat = LD(xmat);
  const Real* q = LD(qpos);
  Real* cdof = LD(cdof);
  int p = M.body_parent[b];
  V3 pos = ld3(xpos + 3 * p) + mul_vec(ldm(xmat + 9 * p), ld3(M.body_pos + 3 * b));
  Q4 quat = qmul(ldq(xquat + 4 * p), ldq(M.body_quat + 4 * b));
  int j0 = M.body_jntadr[b], j1 = j0 + M.body_jntnum[b];
  for (int j = j0; j < j1; j++) {
    int qa = M.jnt_qposadr[j], da = M.jnt_dofadr[j];
    int t = M.jnt_type[j];
    if (t == J_FREE) {
      pos = v3(q[qa], q[qa + 1], q[qa + 2]);
      quat = qnormalized(q4(q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]));
      M3 r = to_mat(quat);
      for (int k = 0; k < 3; k++) {
        V3 e = v3(k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0);
        stsp(cdof + 6 * (da + k), sp(v3z(), e));
        V3 a = col(r, k);
        stsp(cdof + 6 * (da + 3 + k), sp(a, cross(pos, a)));
      }
    } else if (t == J_BALL) {
      V3 jp = ld3(M.jnt_pos + 3 * j);
      V3 anchor = pos + qrotate(quat, jp);
      Q4 ql = qnormalized(q4(q[qa], q[qa + 1], q[qa + 2], q[qa + 3]));
      quat = qmul(quat, ql);
      pos = anchor - qrotate(quat, jp);
      M3 r = to_mat(quat);
      for (int k = 0; k < 3; k++) {
        V3 a = col(r, k);
        stsp(cdof + 6 * (da + k), sp(a, cross(anchor, a)));
      }
    } else if (t == J_HINGE) {
      V3 jp = ld3(M.jnt_pos + 3 * j), jax = ld3(M.jnt_axis + 3 * j);
      V3 anchor = pos + qrotate(quat, jp);
      V3 axis = qrotate(quat, jax);
      quat = qmul(quat, ldq(LD(jq) + 4 * j));  // from_axis_angle(axis, q - qpos0), precomputed
      pos = anchor - qrotate(quat, jp);
      stsp(cdof + 6 * da, sp(axis, cross(anchor, axis)));
    } else {
      V3 jp = ld3(M.jnt_pos + 3 * j), jax = ld3(M.jnt_axis + 3 * j);
      V3 anchor = pos + qrotate(quat, jp);
      (void)anchor;
      V3 axis = qrotate(quat, jax);
      pos = pos + axis * (q[qa] - M.qpos0[qa]);
      stsp(cdof + 6 * da, sp(v3z(), axis));
    }
  }
  quat = qnormalized(quat);
  M3 xm = to_mat(quat);
  st3(xpos + 3 * b, pos);
  stq(xquat + 4 * b, quat);
  stm(xmat + 9 * b, xm);
}

// Body centre of mass and composite-ready spatial inertia (the tail of each
// body's step in forward.rs:14-90). They depend only on the body's own pose,
// not on the tree order, so all bodies are done at once after the levels.
template <class X, class MT>
PHD void body_inertia(const MT& M, Env& E, int b) {
  V3 pos = ld3(LD(xpos) + 3 * b);
  M3 xm = ldm(LD(xmat) + 9 * b);
  V3 xipos = pos + mul_vec(xm, ld3(M.v_body_ipos + ((int64_t)E.var * M.nbody + b) * 3));
  st3(LD(xipos) + 3 * b, xipos);
  // SpatialInertia::from_com(mass, xipos, I.rotate(xmat)), math.rs:388-393
  Real mass = M.v_body_mass[(int64_t)E.var * M.nbody + b];
  M3 ic = mrotate(ldm(M.v_body_inertia + ((int64_t)E.var * M.nbody + b) * 9), xm);
  M3 cx = skew(xipos);
  M3 io = madd(ic, mscale(mul_mat(cx, transpose(cx)), mass));
  Real* ci = LD(cinert) + 13 * b;
  ci[0] = mass;
  st3(ci + 1, xipos * mass);
  stm(ci + 4, io);
}

template <class X, class MT>
PHD void kinematics(const MT& M, Env& E) {
  // hinge rotations do not depend on the parent's pose: evaluate their
  // sin/cos for all joints at once, off the tree's serial path
  X::par(M.njnt, [&](int j) {
    if (M.jnt_type[j] == J_HINGE) {
      int qa = M.jnt_qposadr[j];
      stq(LD(jq) + 4 * j, from_axis_angle(ld3(M.jnt_axis + 3 * j), LD(qpos)[qa] - M.qpos0[qa]));
    }
  });
  X::par(1, [&](int) {
    st3(LD(xpos), v3z());
    stq(LD(xquat), q4(1, 0, 0, 0));
    stm(LD(xmat), m3id());
    st3(LD(xipos), v3z());
    Real* ci = LD(cinert);
    for (int k = 0; k < 13; k++) ci[k] = 0.0;
  });
  for (int l = 0; l < M.nlevel; l++) {
    int b0 = M.level_adr[l], nb = M.level_adr[l + 1] - b0;
    X::par(nb, [&](int i) { fk_body<X>(M, E, M.level_body[b0 + i]); });
  }
  X::par(M.nbody - 1, [&](int i) { body_inertia<X>(M, E, i + 1); });
}

// Geom world poses, the last loop of forward.rs:14-90. Evaluated where they
// are consumed (collision, pose export) so that they need no LDS of their own
// across the dynamics stages.
template <class X, class MT>
PHD void geom_poses(const MT& M, Env& E) {
  X::par(M.ngeom, [&](int g) {
    int b = M.geom_body[g];
    M3 xm = ldm(LD(xmat) + 9 * b);
    st3(LD(gxpos) + 3 * g, ld3(LD(xpos) + 3 * b) + mul_vec(xm, ld3(M.geom_pos + 3 * g)));
    stm(LD(gxmat) + 9 * g, mul_mat(xm, to_mat(ldq(M.geom_quat + 4 * g))));
  });
}

// x[parent(b)] += x[b] for b from the last body down to 1 (W components per
// body), level by level. A level with at most 16 (parent, component) pairs
// gives each pair a lane; a larger one gives each parent a lane that sums all
// W components side by side (independent chains, one pass over the schedule).
// Each component adds the children in the same order either way.
template <class X, int W, class MT>
PHD void leaf_to_root(const MT& M, Real* x) {
  for (int L = 0; L < M.nacc; L++) {
    int g0 = M.acc_level_adr[L], ng = M.acc_level_adr[L + 1] - g0;
    if (ng * W <= GL) {
      X::par(ng * W, [&](int t) {
        int g = g0 + t / W, c = t % W;
        int p = M.acc_parent[g];
        Real v = x[W * p + c];
        for (int k = M.acc_child_adr[g]; k < M.acc_child_adr[g + 1]; k++) v = v + x[W * M.acc_child[k] + c];
        x[W * p + c] = v;
      });
    } else {
      X::par(ng, [&](int i) {
        const int g = g0 + i;
        const int p = M.acc_parent[g];
        Real v[W];
#pragma unroll
        for (int c = 0; c < W; c++) v[c] = x[W * p + c];
        for (int k = M.acc_child_adr[g]; k < M.acc_child_adr[g + 1]; k++) {
          const Real* xc = x + W * M.acc_child[k];
#pragma unroll
          for (int c = 0; c < W; c++) v[c] = v[c] + xc[c];
        }
#pragma unroll
        for (int c = 0; c < W; c++) x[W * p + c] = v[c];
      });
    }
  }
}

// ---------------- CRBA, forward.rs:93-112 ----------------
template <class X, class MT>
PHD void crba(const MT& M, Env& E) {
  const int nb = M.nbody, nv = M.nv;
  Real* crb = LD(crb);
  const Real* ci = LD(cinert);
  Real* qm = LD(qm);
  // composite inertia leaf to root: each of the 13 components is its own
  // sequence of additions in the reference's (decreasing body) order
  X::par(13 * nb, [&](int t) { crb[t] = ci[t]; });
  leaf_to_root<X, 13>(M, crb);
  X::par(nv * nv, [&](int k) { qm[k] = 0.0; });
  X::par(nv, [&](int i) {
    const Real* cd = LD(cdof);
    SP f = mul_motion(crb + 13 * M.dof_body[i], ldsp(cd + 6 * i));
    // i, then the dof_parent chain; that chain is the set of strict
    // ancestors in decreasing index order (parents precede children), read
    // from the ancestor mask instead of one dependent load per step
    uint64_t anc = M.dof_anc[i];
    int j = i;
    while (true) {
      Real v = sdot(ldsp(cd + 6 * j), f);
      qm[i * nv + j] = v;
      qm[j * nv + i] = v;
      if (!anc) break;
      j = 63 - __builtin_clzll(anc);
      anc &= ~((uint64_t)1 << j);
    }
    qm[i * nv + i] += M.dof_armature[i];
  });
}

// ---------------- velocity + RNE, forward.rs:115-171 ----------------
// One pass over the tree levels computes, per body, cvel and cdof_dot
// (forward.rs:115-144) and immediately the RNE forward terms cacc and cfrc
// (forward.rs:147-165): cdof_dot of a dof is consumed by its own body's
// acceleration sum in dof order, exactly where the reference's second loop
// would read it, so it never needs storage. Then the backward force sum and
// qfrc_bias (forward.rs:166-171).
template <class X, class MT>
PHD void velocity_rne(const MT& M, Env& E) {
  Real* cvel = LD(cvel);
  Real* cacc = LD(cacc);
  Real* cfrc = LD(cfrc);
  const Real* cd = LD(cdof);
  const Real* ci = LD(cinert);
  const Real* qvel = LD(qvel);
  X::par(1, [&](int) {
    stsp(cvel, sp(v3z(), v3z()));
    stsp(cacc, sp(v3z(), -v3(M.gravity[0], M.gravity[1], M.gravity[2])));
    stsp(cfrc, sp(v3z(), v3z()));
  });
  for (int l = 0; l < M.nlevel; l++) {
    int b0 = M.level_adr[l], nb = M.level_adr[l + 1] - b0;
    X::par(nb, [&](int i) {
      int b = M.level_body[b0 + i];
      int par = M.body_parent[b];
      SP v = ldsp(cvel + 6 * par);
      SP a = ldsp(cacc + 6 * par);
      int j0 = M.body_jntadr[b], j1 = j0 + M.body_jntnum[b];
      for (int j = j0; j < j1; j++) {
        int da = M.jnt_dofadr[j], t = M.jnt_type[j];
        if (t == J_FREE || t == J_BALL) {
          int lin = t == J_FREE ? 3 : 0, rot = t == J_FREE ? da + 3 : da;
          for (int k = 0; k < lin; k++) {
            a = spadd(a, sscale(sp(v3z(), v3z()), qvel[da + k]));
            v = spadd(v, sscale(ldsp(cd + 6 * (da + k)), qvel[da + k]));
          }
          for (int k = 0; k < 3; k++) a = spadd(a, sscale(cross_motion(v, ldsp(cd + 6 * (rot + k))), qvel[rot + k]));
          for (int k = 0; k < 3; k++) v = spadd(v, sscale(ldsp(cd + 6 * (rot + k)), qvel[rot + k]));
        } else {
          a = spadd(a, sscale(cross_motion(v, ldsp(cd + 6 * da)), qvel[da]));
          v = spadd(v, sscale(ldsp(cd + 6 * da), qvel[da]));
        }
      }
      stsp(cvel + 6 * b, v);
      stsp(cacc + 6 * b, a);
      SP h = mul_motion(ci + 13 * b, v);
      stsp(cfrc + 6 * b, spadd(mul_motion(ci + 13 * b, a), cross_force(v, h)));
    });
  }
  leaf_to_root<X, 6>(M, cfrc);
  X::par(M.nv, [&](int k) {
    LD(bias)[k] = sdot(ldsp(LD(cdof) + 6 * k), ldsp(cfrc + 6 * M.dof_body[k]));
  });
}

// ---------------- passive + actuation, forward.rs:173-211 ----------------
template <class X, class MT>
PHD void passive_actuation(const MT& M, Env& E) {
  const Real* qpos = LD(qpos);
  const Real* qvel = LD(qvel);
  X::par(M.nv, [&](int k) {
    Real f = -M.v_dof_damping[(int64_t)E.var * M.nv + k] * qvel[k];
    int j = M.dof_jnt[k];
    Real kp = M.jnt_stiffness[j];
    int t = M.jnt_type[j];
    if (kp != 0.0 && (t == J_HINGE || t == J_SLIDE)) f -= kp * (qpos[M.jnt_qposadr[j]] - M.jnt_springref[j]);
    LD(passive)[k] = f;
  });
  X::par(M.nu, [&](int u) {
    Real gear = M.act_gear[u];
    int a0 = M.act_madr[u], a1 = a0 + M.act_mnum[u];
    Real sl = rsum0(), sv = rsum0();
    for (int i = a0; i < a1; i++) sl += M.mom_coef[i] * qpos[M.jnt_qposadr[M.mom_jnt[i]]];
    for (int i = a0; i < a1; i++) sv += M.mom_coef[i] * qvel[M.jnt_dofadr[M.mom_jnt[i]]];
    Real length = gear * sl, vel = gear * sv;
    Real ctrl = LD(ctrl)[u];
    if (M.act_ctrllimited[u]) ctrl = rclamp(ctrl, M.act_ctrlrange[2 * u], M.act_ctrlrange[2 * u + 1]);
    auto b = M.v_act_bias + ((int64_t)E.var * M.nu + u) * 3;
    Real f = M.v_act_gain[(int64_t)E.var * M.nu + u] * ctrl + b[0] + b[1] * length + b[2] * vel;
    Real clamped = 0.0;
    if (M.act_forcelimited[u]) {
      Real lo = M.act_forcerange[2 * u], hi = M.act_forcerange[2 * u + 1];
      clamped = (f < lo || f > hi) ? 1.0 : 0.0;
      f = rclamp(f, lo, hi);
    }
    LD(actf)[u] = f;
    LD(actclamp)[u] = clamped;
  });
  X::par(M.nv, [&](int k) {
    Real s = 0.0;
    int a0 = M.dofact_adr[k], a1 = a0 + M.dofact_num[k];
    for (int i = a0; i < a1; i++) {
      int u = M.dofact_u[i];
      s += M.act_gear[u] * M.dofact_coef[i] * LD(actf)[u];
    }
    LD(qfrc_act)[k] = s;
  });
}

// The hull of geom g (a mesh geom), or an empty reference.
template <class MT>
PHD HullRef hull_ref(const MT& M, int g) {
  HullRef h{};
  const int k = M.geom_dataid[g];
  if (k < 0) return h;
  const int va = M.mesh_vertadr[k];
  h.vert = M.mesh_vert + 3 * (int64_t)va;
  h.vfadr = M.vert_faceadr + va;
  h.vfnum = M.vert_facenum + va;
  h.vface = M.vert_face;
  h.fnormal = M.face_normal;
  h.fvadr = M.face_vertadr;
  h.fvnum = M.face_vertnum;
  h.fvert = M.face_vert;
  h.nvert = M.mesh_vertnum[k];
  return h;
}

// ---------------- collision, forward.rs:214-260 ----------------
template <class X, class MT>
PHD void collide(const MT& M, const DState& S, Env& E) {
  const int ng = M.ngeom;
  geom_poses<X>(M, E);
  Real* aabb = LD(aabb);
  const Real margin = M.margin;
  X::par(ng, [&](int g) {
    auto la = M.v_geom_aabb + ((int64_t)E.var * ng + g) * 6;
    V3 c = ld3(la), h = ld3(la + 3);
    const Real* r = LD(gxmat) + 9 * g;
    if (!rfinite(h.x)) {
      st3(aabb + 6 * g, v3z());
      st3(aabb + 6 * g + 3, h);
      return;
    }
    M3 rm = ldm(r);
    st3(aabb + 6 * g, ld3(LD(gxpos) + 3 * g) + mul_vec(rm, c));
    st3(aabb + 6 * g + 3,
        v3(fabs(r[0]) * h.x + fabs(r[1]) * h.y + fabs(r[2]) * h.z, fabs(r[3]) * h.x + fabs(r[4]) * h.y + fabs(r[5]) * h.z,
           fabs(r[6]) * h.x + fabs(r[7]) * h.y + fabs(r[8]) * h.z));
  });
  int* cnt = reinterpret_cast<int*>(LD(paircnt));
  int* isc = E.si + M.L.i_iscal;
  auto stage = S.stage + E.e * (int64_t)M.npair * MAXHIT * CW;
  // broad phase: cnt[p] = 1 for a pair that passes the masks and the AABB test
  X::par(M.npair, [&](int p) {
    int g1 = M.pair_g1[p], g2 = M.pair_g2[p];
    cnt[p] = 0;
    auto ct = M.v_contype + (int64_t)E.var * ng;
    auto ca = M.v_conaffinity + (int64_t)E.var * ng;
    if ((ct[g1] & ca[g2]) == 0 && (ct[g2] & ca[g1]) == 0) return;  // per-env masks, model.rs:439-443
    const Real* A = aabb + 6 * g1;
    const Real* B = aabb + 6 * g2;
    for (int k = 0; k < 3; k++)
      if (fabs(A[k] - B[k]) > A[3 + k] + B[3 + k] + margin) return;
    cnt[p] = 1;
  });
  // the passing pairs, in pair order, listed in the offset array (free until
  // the compaction below), so that the narrow phase takes one lane per passing
  // pair instead of one round per 16 candidate pairs
  int* list = reinterpret_cast<int*>(LD(pairoff));
  {
    using LV = typename X::LV;
    const int np = M.npair, run = (np + GL - 1) / GL;
    const LV part = X::gen([&](int l) {
      int t = 0;
      for (int p = l * run; p < (l + 1) * run && p < np; p++) t += cnt[p];
      return (Real)t;
    });
    LV pb[GL];
    bcast_all<X>(part, pb);
    X::each([&](int l) {
      int o = 0;
#pragma unroll
      for (int k = 0; k < GL; k++)
        if (k < l) o += (int)X::at(pb[k], l);
      for (int p = l * run; p < (l + 1) * run && p < np; p++)
        if (cnt[p]) list[o++] = p;
      if (l == 0) {
        int tot = 0;
#pragma unroll
        for (int k = 0; k < GL; k++) tot += (int)X::at(pb[k], l);
        isc[12] = tot;
      }
    });
    X::sync();
  }
  const int npass = isc[12];
  X::par(npass, [&](int i) {
    const int p = list[i];
    int g1 = M.pair_g1[p], g2 = M.pair_g2[p];
    GPose a, b;
    a.typ = M.geom_type[g1];
    a.pos = ld3(LD(gxpos) + 3 * g1);
    a.mat = ldm(LD(gxmat) + 9 * g1);
    a.size = ld3(M.v_geom_size + ((int64_t)E.var * ng + g1) * 3);
    b.typ = M.geom_type[g2];
    b.pos = ld3(LD(gxpos) + 3 * g2);
    b.mat = ldm(LD(gxmat) + 9 * g2);
    b.size = ld3(M.v_geom_size + ((int64_t)E.var * ng + g2) * 3);
    Sink sink;
    sink.out = stage + (int64_t)p * MAXHIT * CW;
    sink.n = 0;
    if (M.ngeneral > 0 && (a.typ == G_MESH || b.typ == G_MESH)) {
      a.hull = hull_ref(M, g1);
      b.hull = hull_ref(M, g2);
    }
    if (M.ngeneral > 0 && needs_general(a.typ, b.typ)) collide_general(a, b, margin, sink);
    else collide_pair(a, b, margin, sink);
    cnt[p] = sink.n;
  });
  // pair-ordered compaction (contacts are listed in pair order, then hit
  // order): each lane of the 16-lane row sums a contiguous run of pairs, the
  // run totals are scanned across lanes, then each lane writes its offsets
  int* off = reinterpret_cast<int*>(LD(pairoff));
  {
    using LV = typename X::LV;
    const int np = M.npair, run = (np + GL - 1) / GL;
    const LV part = X::gen([&](int l) {
      int t = 0;
      for (int p = l * run; p < (l + 1) * run && p < np; p++) t += cnt[p];
      return (Real)t;
    });
    LV pb[GL];
    bcast_all<X>(part, pb);
    X::each([&](int l) {
      int o = 0;
#pragma unroll
      for (int k = 0; k < GL; k++)
        if (k < l) o += (int)X::at(pb[k], l);
      for (int p = l * run; p < (l + 1) * run && p < np; p++) {
        off[p] = o;
        o += cnt[p];
      }
      if (l == 0) {
        int tot = 0;
#pragma unroll
        for (int k = 0; k < GL; k++) tot += (int)X::at(pb[k], l);
        if (tot > M.ncmax) {  // only reachable with a user-lowered contact capacity
          isc[7] |= ST_CONTACT_OVERFLOW;
          tot = M.ncmax;
        }
        isc[0] = tot;  // ncon
      }
    });
    X::sync();
  }
  auto con = S.con + E.e * (int64_t)M.ncmax * CW;
  auto cg = S.cgeom + E.e * (int64_t)M.ncmax * 3;
  X::par(M.npair, [&](int p) {
    int n = cnt[p];
    for (int h = 0; h < n && off[p] + h < M.ncmax; h++) {
      auto src = stage + ((int64_t)p * MAXHIT + h) * CW;
      auto dst = con + (int64_t)(off[p] + h) * CW;
      for (int k = 0; k < CW; k++) dst[k] = src[k];
      auto gd = cg + (off[p] + h) * 3;
      gd[0] = M.pair_g1[p];
      gd[1] = M.pair_g2[p];
      gd[2] = h;
    }
  });
}

// Contact::frame, collision.rs:47-55
template <class P>
PHD void contact_frame(P c, V3* t1, V3* t2) {
  V3 n = ld3(c + 3), tg = ld3(c + 6);
  V3 t = tg - n * dot(n, tg);
  if (norm2(t) > 1e-12) {
    *t1 = normalized(t);
    *t2 = cross(n, *t1);
  } else {
    tangents(n, t1, t2);
  }
}

// newton.rs:52-66
PHD Real impedance(const Real* solimp, Real pos) {
  Real dmin = solimp[0], dmax = solimp[1], width = solimp[2], mid = solimp[3], power = solimp[4];
  Real x = rmin(fabs(pos) / rmax(width, 1e-12), 1.0);
  if (x >= 1.0 || dmin == dmax) return dmax;
  Real y;
  if (x <= mid) {
    Real num = power == 2.0 ? x * x : (power == 1.0 ? x : pow(x, power));
    Real pm = power - 1.0;
    Real den = pm == 2.0 ? mid * mid : (pm == 1.0 ? mid : pow(mid, pm));
    y = num / den;
  } else {
    Real b = 1.0 - x;
    Real num = power == 2.0 ? b * b : (power == 1.0 ? b : pow(b, power));
    Real bm = 1.0 - mid, pm = power - 1.0;
    Real den = pm == 2.0 ? bm * bm : (pm == 1.0 ? bm : pow(bm, pm));
    y = 1.0 - num / den;
  }
  return rclamp(dmin + y * (dmax - dmin), 1e-4, 0.9999);
}

// ---------------- block Cholesky, math.rs:425-463 ----------------
// Factor the diagonal block [s0, e) of the nv x nv row-major matrix A in place.
// Column j: the diagonal (one lane), then the entries below it (one lane per
// row, each the reference's 4-accumulator dot). Returns false if not PD.
template <class X, class MT>
PHD bool chol_block(const MT& M, Env& E, Real* A, int nv, int s0, int e) {
  if (e - s0 <= GL) {
    bool ok = chol_lanes<X>(A, nv, s0, e - s0);
    X::sync();
    return ok;
  }
  int* isc = E.si + M.L.i_iscal;
  for (int j = s0; j < e; j++) {
    X::par(1, [&](int) {
      Real* rj = A + j * nv + s0;
      Real d = A[j * nv + j] - sdot_n(rj, rj, j - s0);
      if (d <= 0.0) {
        isc[1] = 1;
        return;
      }
      A[j * nv + j] = sqrt(d);
    });
    if (isc[1]) return false;
    X::par(e - j - 1, [&](int t) {
      int i = j + 1 + t;
      Real inv = 1.0 / A[j * nv + j];
      Real* ri = A + i * nv + s0;
      int k = j - s0;
      ri[k] = (ri[k] - sdot_n(ri, A + j * nv + s0, k)) * inv;
    });
  }
  return true;
}

// Forward then backward substitution on block [s0, e); x is an LDS vector.
template <class X, class MT>
PHD void chol_solve_block(const MT& M, Env& E, const Real* L, int nv, int s0, int e, Real* x) {
  (void)M;
  (void)E;
  if (e - s0 <= GL) {
    solve_lanes<X>(L, nv, s0, e - s0, x);
    X::sync();
    return;
  }
  X::par(1, [&](int) {
    for (int i = s0; i < e; i++) x[i] = (x[i] - sdot_n(L + i * nv + s0, x + s0, i - s0)) / L[i * nv + i];
    for (int i = e - 1; i >= s0; i--) {
      x[i] /= L[i * nv + i];
      Real xi = x[i];
      const Real* li = L + i * nv + s0;
      for (int k = 0; k < i - s0; k++) x[s0 + k] += -xi * li[k];
    }
  });
}

template <class X, class MT>
PHD bool factor_mass(const MT& M, Env& E, Real h) {
  const int nv = M.nv;
  Real* chol = LD(chol);
  const Real* qm = LD(qm);
  X::par(nv * nv, [&](int k) { chol[k] = qm[k]; });
  // add_implicit_terms, forward.rs:407-431
  X::par(nv, [&](int k) { chol[k * nv + k] += h * M.v_dof_damping[(int64_t)E.var * nv + k]; });
  if (M.integrator == 1) {
    X::par(1, [&](int) {
      for (int u = 0; u < M.nu; u++) {
        Real kv = -M.v_act_bias[((int64_t)E.var * M.nu + u) * 3 + 2] * M.act_gear[u] * M.act_gear[u];
        if (kv == 0.0 || LD(actclamp)[u] != 0.0) continue;
        int a0 = M.act_madr[u], a1 = a0 + M.act_mnum[u];
        for (int ia = a0; ia < a1; ia++)
          for (int ib = a0; ib < a1; ib++) {
            int a = M.jnt_dofadr[M.mom_jnt[ia]], b = M.jnt_dofadr[M.mom_jnt[ib]];
            bool anc = ((M.dof_anc[b] >> a) & 1) || ((M.dof_anc[a] >> b) & 1);
            if (a == b || anc) chol[a * nv + b] += h * kv * M.mom_coef[ia] * M.mom_coef[ib];
          }
      }
    });
  }
  for (int blk = 0; blk < M.nblock; blk++)
    if (!chol_block<X>(M, E, chol, nv, M.block_adr[blk], M.block_adr[blk + 1])) return false;
  return true;
}

template <class X, class MT>
PHD void mass_solve(const MT& M, Env& E, Real* x) {
  for (int blk = 0; blk < M.nblock; blk++)
    chol_solve_block<X>(M, E, LD(chol), M.nv, M.block_adr[blk], M.block_adr[blk + 1], x);
}

// factor_mass then mass_solve for the common case (Euler, one dof block of
// at most 16 dofs) in registers: M + h diag(damping) one row per lane, the
// register Cholesky, then the two triangular solves. Other models take the
// general path. Returns false when the matrix is not positive definite.
template <class X, class MT>
PHD bool factor_solve_mass(const MT& M, Env& E, Real h, Real* x) {
  if (M.integrator == 0 && M.nblock == 1 && M.nv <= GL) {
    using LV = typename X::LV;
    const int nv = M.nv;
    const Real* qm = LD(qm);
    auto damp = M.v_dof_damping + (int64_t)E.var * nv;
    LV r[GL];
#pragma unroll
    for (int k = 0; k < GL; k++)
      r[k] = X::gen([&](int l) {
        Real v = qm[clampi(l, nv) * nv + clampi(k, nv)];
        return (l < nv && k < nv) ? v : 0.0;
      });
    // add_implicit_terms (forward.rs:407-431): the diagonal gains h * damping
    const LV hd = X::gen([&](int l) { return h * damp[clampi(l, nv)]; });
#pragma unroll
    for (int k = 0; k < GL; k++) r[k] = X::sel([&](int l) { return l == k; }, r[k] + hd, r[k]);
    if (!chol_rows<X>(r, nv)) return false;
    store_lower<X>(r, LD(chol), nv, 0, nv);
    X::sync();
    LV lr[GL], lc[GL], lii;
    load_factor<X>(LD(chol), nv, 0, nv, lr, lc, &lii);
    LV xv = X::gen([&](int l) {
      Real v = x[clampi(l, nv)];
      return l < nv ? v : 0.0;
    });
    xv = solve_rows<X>(r, lc, lii, xv, nv);
    X::each([&](int l) {
      if (l < nv) x[l] = X::at(xv, l);
    });
    X::sync();
    return true;
  }
  if (!factor_mass<X>(M, E, h)) return false;
  mass_solve<X>(M, E, x);
  return true;
}

// ---------------- constraint rows, newton.rs:113-218 ----------------
// Returns nr; rows are bound to LDS or global storage by their count.
template <class X, class MT>
PHD int build_rows(const MT& M, const DState& S, Env& E) {
  const int nv = M.nv;
  int* isc = E.si + M.L.i_iscal;
  int* cact = reinterpret_cast<int*>(LD(cact));
  auto con = S.con + E.e * (int64_t)M.ncmax * CW;
  auto cg = S.cgeom + E.e * (int64_t)M.ncmax * 3;
  const Real* qpos = LD(qpos);
  // Violated limits per joint (bit 0 lower, bit 1 upper), in joint order.
  // With at most 16 joints each lane owns a joint and the row positions come
  // from a scan across the lanes; otherwise one lane walks the joints.
  const bool lane_limits = M.njnt <= GL;
  using LV = typename X::LV;
  auto seps_of = [&](int j, Real* s0, Real* s1) {
    int t = M.jnt_type[j];
    if (!M.jnt_limited[j] || !(t == J_HINGE || t == J_SLIDE)) return 0;
    Real q = qpos[M.jnt_qposadr[j]];
    *s0 = q - M.jnt_range[2 * j];
    *s1 = M.jnt_range[2 * j + 1] - q;
    return (*s0 < 0.0 ? 1 : 0) + (*s1 < 0.0 ? 2 : 0);
  };
  LV lim_before = X::splat(0.0), lim_flags = X::splat(0.0);
  if (lane_limits) {
    lim_flags = X::gen([&](int l) {
      Real s0 = 0.0, s1 = 0.0;
      return l < M.njnt ? (Real)seps_of(l, &s0, &s1) : 0.0;
    });
    LV fb[GL];
    bcast_all<X>(lim_flags, fb);
    lim_before = X::gen([&](int l) {
      int n = 0;
#pragma unroll
      for (int k = 0; k < GL; k++) {
        int f = (int)X::at(fb[k], l);
        if (k < l) n += (f & 1) + (f >> 1);
      }
      return (Real)n;
    });
    X::each([&](int l) {
      if (l == 0) {
        int n = 0;
#pragma unroll
        for (int k = 0; k < GL; k++) {
          int f = (int)X::at(fb[k], l);
          n += (f & 1) + (f >> 1);
        }
        isc[4] = n;
      }
    });
    X::sync();
  }
  // list of touching contacts, then the row count
  X::par(1, [&](int) {
    int ncon = isc[0], na = 0;
    for (int c = 0; c < ncon; c++)
      if (con[c * CW + 9] > 0.0) cact[na++] = c;
    int nl = 0;
    if (lane_limits) {
      nl = isc[4];
    } else {
      for (int j = 0; j < M.njnt; j++) {
        Real s0, s1;
        int f = seps_of(j, &s0, &s1);
        nl += (f & 1) + (f >> 1);
      }
    }
    isc[2] = na;
    isc[3] = 4 * na + nl + M.neq;
    isc[4] = nl;
  });
  const int na = isc[2], nr = isc[3];
  if (nr > M.nrmax) {
    X::par(1, [&](int) { isc[5] = 1; });
    return -1;
  }
  rows_bind<X>(M, S, E, nr);
  if (nr == 0) return 0;
  Real* J = E.J;
  // contact rows: lane per (contact, dof) builds J_n, J_t1, J_t2 at that dof
  X::par(na * nv, [&](int t) {
    int k = t / nv, dof = t % nv;
    int c = cact[k];
    auto cc = con + c * CW;
    int g1 = cg[3 * c], g2 = cg[3 * c + 1];
    int b1 = M.geom_body[g1], b2 = M.geom_body[g2];
    Real mu = rmax(M.v_friction[(int64_t)E.var * M.ngeom + g1], M.v_friction[(int64_t)E.var * M.ngeom + g2]);
    V3 t1, t2;
    contact_frame(cc, &t1, &t2);
    V3 p = ld3(cc);
    V3 dirs[3] = {ld3(cc + 3), t1, t2};
    Real jv[3];
    SP cd = ldsp(LD(cdof) + 6 * dof);
    bool on2 = (M.body_chain[b2] >> dof) & 1, on1 = (M.body_chain[b1] >> dof) & 1;
    for (int r = 0; r < 3; r++) {
      Real v = 0.0;
      if (on2) v += 1.0 * dot(dirs[r], point_velocity(cd, p));
      if (on1) v += -1.0 * dot(dirs[r], point_velocity(cd, p));
      jv[r] = v;
    }
    for (int ti = 0; ti < 2; ti++)
      for (int si = 0; si < 2; si++) {
        Real s = si == 0 ? 1.0 : -1.0;
        J[(4 * k + 2 * ti + si) * nv + dof] = jv[0] + s * mu * jv[1 + ti];
      }
  });
  X::par(na, [&](int k) {
    int c = cact[k];
    int g1 = cg[3 * c], g2 = cg[3 * c + 1];
    int b1 = M.geom_body[g1], b2 = M.geom_body[g2];
    Real mu = rmax(M.v_friction[(int64_t)E.var * M.ngeom + g1], M.v_friction[(int64_t)E.var * M.ngeom + g2]);
    auto iw = M.v_body_invweight + (int64_t)E.var * M.nbody * 2;
    Real tran = iw[2 * b1] + iw[2 * b2];
    Real dpyr = 2.0 * mu * mu * (1.0 + mu * mu) * tran / M.impratio;
    for (int r = 0; r < 4; r++) {
      E.rpos[4 * k + r] = -con[c * CW + 9];
      E.rdiag[4 * k + r] = dpyr;
    }
  });
  // joint limit rows, in joint order (lower then upper), then equalities
  auto iw = M.v_dof_invweight + (int64_t)E.var * nv;
  auto limit_rows = [&](int j, int f, int r) {
    Real seps[2];
    seps_of(j, &seps[0], &seps[1]);
    int da = M.jnt_dofadr[j];
    for (int s = 0; s < 2; s++) {
      if ((f >> s) & 1) {
        for (int k = 0; k < nv; k++) J[r * nv + k] = 0.0;
        J[r * nv + da] = s == 0 ? 1.0 : -1.0;
        E.rpos[r] = seps[s];
        E.rdiag[r] = iw[da];
        r++;
      }
    }
    return r;
  };
  if (lane_limits) {
    X::each([&](int l) {
      int f = (int)X::at(lim_flags, l);
      if (l < M.njnt && f) limit_rows(l, f, 4 * na + (int)X::at(lim_before, l));
    });
    X::sync();
  }
  X::par(1, [&](int) {
    int r = 4 * na;
    if (lane_limits) {
      r += isc[4];
    } else {
      for (int j = 0; j < M.njnt; j++) {
        Real s0, s1;
        int f = seps_of(j, &s0, &s1);
        if (f) r = limit_rows(j, f, r);
      }
    }
    for (int e = 0; e < M.neq; e++) {
      int j1 = M.eq_j1[e], j2 = M.eq_j2[e];
      auto c = M.eq_poly + 5 * e;
      Real q1 = qpos[M.jnt_qposadr[j1]] - M.qpos0[M.jnt_qposadr[j1]];
      Real poly = c[0], dpoly = 0.0;
      if (j2 >= 0) {
        Real x = qpos[M.jnt_qposadr[j2]] - M.qpos0[M.jnt_qposadr[j2]];
        poly = c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4])));
        dpoly = c[1] + x * (2.0 * c[2] + x * (3.0 * c[3] + x * 4.0 * c[4]));
      }
      for (int k = 0; k < nv; k++) J[r * nv + k] = 0.0;
      J[r * nv + M.jnt_dofadr[j1]] += 1.0;
      Real dg = iw[M.jnt_dofadr[j1]];
      if (j2 >= 0) {
        J[r * nv + M.jnt_dofadr[j2]] -= dpoly;
        dg += iw[M.jnt_dofadr[j2]];
      }
      E.rpos[r] = q1 - poly;
      E.rdiag[r] = dg;
      r++;
    }
  });
  // reference acceleration and regularisation per row
  const int n_ineq = nr - M.neq;
  X::par(nr, [&](int r) {
    bool bil = r >= n_ineq;
    Real solref[2], solimp[5];
    for (int i = 0; i < 2; i++) solref[i] = bil ? M.eq_solref[2 * (r - n_ineq) + i] : M.solref[i];
    for (int i = 0; i < 5; i++) solimp[i] = bil ? M.eq_solimp[5 * (r - n_ineq) + i] : M.solimp[i];
    Real tc = rmax(solref[0], 2.0 * M.timestep), dr = solref[1];
    Real dmax = solimp[1];
    Real b = 2.0 / (dmax * tc);
    Real k = 1.0 / (dmax * dmax * tc * tc * dr * dr);
    Real vel = rsum0();
    const Real* jr = J + r * nv;
    const Real* qvel = LD(qvel);
    for (int i = 0; i < nv; i++) vel += jr[i] * qvel[i];
    Real imp = impedance(solimp, E.rpos[r]);
    Real reg = rmax((1.0 - imp) / imp * E.rdiag[r], 1e-15);
    E.aref[r] = -b * vel - k * imp * E.rpos[r];
    E.dd[r] = 1.0 / reg;
  });
  return nr;
}

// ---------------- Newton, newton.rs:224-371 ----------------
// One Newton problem: the whole model (one island) or one island gathered
// into its own arrays (newton.rs:436-471).
struct NP {
  int nv, nr, n_ineq;                  // dofs, rows, rows below n_ineq are inequalities
  const Real *qm, *J, *aref, *dd;      // qm: nv x nv, J: nr x nv (row stride nv)
  const Real* warm;                    // previous acceleration (warm start candidate)
  Real* a;                             // in: unconstrained acceleration; out: solution
  Real* qf;                            // out: constraint force, nv
  Real* fac;                           // nv x nv factor storage
  const Real* warm2;                   // learned warm-start candidate, or null
};

template <class X>
PHD void mul_m(const NP& P, const Real* x, Real* y) {
  const int nv = P.nv;
  X::par(nv, [&](int i) { y[i] = sdot_n(P.qm + i * nv, x, nv) + 0.0 * x[i]; });
}

// newton.rs:224-371 for nv <= 16 and nr <= 16 rows held in LDS, with the
// vectors in registers: a dof vector holds dof k in lane k, a row vector holds
// row r in lane r, exchanged by broadcasts inside the 16-lane row. M and J stay
// in LDS and are read with independent loads; the Hessian is assembled one row
// per lane directly into the register Cholesky. Every sum keeps the
// reference's order, so the result equals newton_dense's bit for bit.
template <class X, class MT>
PHD int newton_lanes(const MT& M, Env& E, const NP& P) {
  using LV = typename X::LV;
  const int nv = P.nv, nr = P.nr, n_ineq = P.n_ineq;
  // every array of P is in LDS on this path
  const Real* qm = P.qm;
  const Real* J = P.J;
  const Real* rdd = P.dd;
  Real* Lf = P.fac;
  const LV aref = X::gen([&](int l) {
    Real v = P.aref[clampi(l, nr)];
    return l < nr ? v : 0.0;
  });
  const LV dd = X::gen([&](int l) {
    Real v = rdd[clampi(l, nr)];
    return l < nr ? v : 0.0;
  });
  // active(v): bilateral rows always, inequality rows when v < 0
  auto active_mask = [&](const LV& v) {
    return X::ballot([&](int l) { return l < nr && (l >= n_ineq || X::at(v, l) < 0.0); });
  };
  // M x (damp_h is zero inside Newton: damping is applied after the solve)
  auto mul_m = [&](const LV& x) {
    LV xb[GL], mr[GL];
    bcast_all<X>(x, xb);
#pragma unroll
    for (int k = 0; k < GL; k++)
      mr[k] = X::gen([&](int l) {
        Real v = qm[clampi(l, nv) * nv + clampi(k, nv)];
        return (l < nv && k < nv) ? v : 0.0;
      });
    return lsdot_rt<X>(mr, xb, nv) + 0.0 * x;
  };
  // J x, one row per lane
  auto jdot = [&](const LV& x) {
    LV xb[GL], jr[GL];
    bcast_all<X>(x, xb);
#pragma unroll
    for (int k = 0; k < GL; k++)
      jr[k] = X::gen([&](int l) {
        Real v = J[clampi(l, nr) * nv + clampi(k, nv)];
        return (l < nr && k < nv) ? v : 0.0;
      });
    return lsdot_rt<X>(jr, xb, nv);
  };
  // column r of J^T: lane k gets J[r][k]
  auto jrow = [&](int r) {
    return X::gen([&](int l) {
      Real v = J[r * nv + clampi(l, nv)];
      return l < nv ? v : 0.0;
    });
  };
  // sum over active rows of f_r J_r (a dof vector), rows in order, from `init`
  // (act only has bits of rows < nr; the updates are selects, not branches)
  auto add_rows = [&](LV g, const LV& f, uint32_t act) {
    LV fb[GL];
    bcast_all<X>(f, fb);
#pragma unroll
    for (int r = 0; r < GL; r++) {
      const LV jr = X::gen([&](int l) {
        Real v = J[clampi(r, nr) * nv + clampi(l, nv)];
        return l < nv ? v : 0.0;
      });
      g = X::sel([&](int) { return ((act >> r) & 1u) != 0; }, g + fb[r] * jr, g);
    }
    return g;
  };
  auto cost = [&](const LV& x, const LV& a0v) {
    LV t = x - a0v;
    LV t2 = mul_m(t);
    Real c = 0.5 * X::uni(lsum_lanes<X>(t * t2, nv, -0.0));
    LV jar = jdot(x) - aref;
    uint32_t on = active_mask(jar);
    LV term = ((0.5 * dd) * jar) * jar;
    LV tb[GL];
    bcast_all<X>(term, tb);
#pragma unroll
    for (int r = 0; r < GL; r++) {
      Real cr = c + X::uni(tb[r]);
      c = ((on >> r) & 1u) ? cr : c;
    }
    return c;
  };

  const LV a0 = X::gen([&](int l) {
    Real v = P.a[clampi(l, nv)];
    return l < nv ? v : 0.0;
  });
  LV a = a0;
  {
    const LV warm = X::gen([&](int l) {
      Real v = P.warm[clampi(l, nv)];
      return l < nv ? v : 0.0;
    });
    Real c0 = cost(a0, a0), c1 = cost(warm, a0);
    if (c1 < c0) a = warm;
    if (P.warm2) {
      // the learned candidate replaces the reference's choice only when cheaper
      const LV w2 = X::gen([&](int l) {
        Real v = P.warm2[clampi(l, nv)];
        return l < nv ? v : 0.0;
      });
      Real c2 = cost(w2, a0);
      if (c2 < (c1 < c0 ? c1 : c0)) a = w2;
    }
  }
  const LV ma0 = mul_m(a0);
  const Real scale = 1.0 + sqrt_cr(X::uni(lsum_lanes<X>(ma0 * ma0, nv, -0.0)));
  const Real tol = M.tolerance;
  const int cap = M.iterations > 1 ? M.iterations : 1;
  uint64_t prev = ~(uint64_t)0;
  int iters = 0;
  // residuals and active set at the current a, kept for the final force while a is unchanged
  LV jar_a = X::splat(0.0);
  uint32_t act_a = 0;
  bool fresh = false;
  while (iters < cap) {
    iters++;
    const LV tmp = a - a0;
    const LV jar = jdot(a) - aref;
    const uint32_t act = active_mask(jar);
    jar_a = jar;
    act_a = act;
    fresh = true;
    const LV grad = add_rows(mul_m(tmp), dd * jar, act);
    const Real gnorm = sqrt_cr(X::uni(lsum_lanes<X>(grad * grad, nv, -0.0)));
    if (gnorm < tol * scale || ((uint64_t)act == prev && iters > 1)) break;
    prev = act;
    // H = M + sum_active D_r J_r^T J_r, row l in lane l
    LV h[GL];
#pragma unroll
    for (int j = 0; j < GL; j++) {
      h[j] = X::gen([&](int l) {
        Real v = qm[clampi(l, nv) * nv + clampi(j, nv)];
        return (l < nv && j < nv) ? v : 0.0;
      });
      h[j] = X::sel([&](int l) { return l == j; }, h[j] + 0.0, h[j]);
    }
    for (int r = 0; r < nr; r++) {
      const bool on = ((act >> r) & 1u) != 0;
      const LV ji = jrow(r);
      const LV di = rdd[r] * ji;
#pragma unroll
      for (int j = 0; j < GL; j++) {
        Real jrj = J[r * nv + clampi(j, nv)];
        h[j] = X::sel([&](int l) { return on && X::at(ji, l) != 0.0; }, h[j] + di * jrj, h[j]);
      }
    }
    if (!chol_rows<X>(h, nv)) break;
    store_lower<X>(h, Lf, nv, 0, nv);
    X::sync();
    LV lr[GL], lc[GL], lii;
    load_factor<X>(Lf, nv, 0, nv, lr, lc, &lii);
    const LV p = solve_rows<X>(h, lc, lii, -grad, nv);
    // exact line search along p (newton.rs:310-351)
    const LV mp = mul_m(p);
    const Real mpp = X::uni(lsum_lanes<X>(p * mp, nv, -0.0));
    const Real mpg = X::uni(lsum_lanes<X>(tmp * mp, nv, -0.0));
    const LV jp = jdot(p);
    // Row r's terms, one row per lane: D_r jp_r v_r and D_r jp_r^2 when the
    // row is active at step al, v_r = jar_r + al jp_r; they are then added in
    // row order. An inactive row (or a lane past nr) adds +0.0, which leaves
    // the sums equal except for the sign of an exact zero, and no later use
    // (|d1|, d1 < 0, al - d1 / max(d2, 1e-30)) depends on that sign.
    const LV dq = dd * jp;
    const LV dq2 = dq * jp;
    auto deriv = [&](Real al, Real* d1o, Real* d2o) {
      const LV v = jar + al * jp;
      auto on = [&](int l) { return l < nr && (l >= n_ineq || X::at(v, l) < 0.0); };
      const LV t1 = X::sel(on, dq * v, X::splat(0.0));
      const LV t2 = X::sel(on, dq2, X::splat(0.0));
      LV b1[GL], b2[GL];
      bcast_all<X>(t1, b1);
      bcast_all<X>(t2, b2);
      Real d1 = al * mpp + mpg, d2 = mpp;
#pragma unroll
      for (int r = 0; r < GL; r++) {
        d1 = d1 + X::uni(b1[r]);
        d2 = d2 + X::uni(b2[r]);
      }
      *d1o = d1;
      *d2o = d2;
    };
    Real lo = 0.0, hi = INFINITY, al = 1.0, d1, d2;
    deriv(0.0, &d1, &d2);
    const Real ltol = 1e-12 * rmax(fabs(d1), 1e-30);
    for (int it = 0; it < 50; it++) {
      deriv(al, &d1, &d2);
      if (fabs(d1) <= ltol) break;
      if (d1 < 0.0) lo = al;
      else hi = al;
      Real next = al - d1 / rmax(d2, 1e-30);
      if (next <= lo || next >= hi) next = rfinite(hi) ? 0.5 * (lo + hi) : 2.0 * rmax(al, lo);
      if (fabs(next - al) <= 1e-14 * rmax(fabs(al), 1.0)) {
        al = next;
        break;
      }
      al = next;
    }
    a = a + al * p;
    fresh = false;
  }
  // add_constraint_force (newton.rs:361-371), from zero, at the final a
  if (!fresh) {
    jar_a = jdot(a) - aref;
    act_a = active_mask(jar_a);
  }
  const LV f = add_rows(X::splat(0.0), (-dd) * jar_a, act_a);
  X::each([&](int l) {
    if (l < nv) {
      P.a[l] = X::at(a, l);
      P.qf[l] = X::at(f, l);
    }
  });
  X::sync();
  return iters;
}

// Lanes an env owns in executor X (0 for executors without lane groups).
template <class X, class = void>
struct wave_lanes {
  static constexpr int value = 0;
};
template <class X>
struct wave_lanes<X, std::void_t<decltype(X::G)>> {
  static constexpr int value = X::G;
};

#if defined(__HIP_DEVICE_COMPILE__)
typedef double f64x4 __attribute__((ext_vector_type(4)));
// H = M + sum over active rows of (d_r J_r)' J_r on the FP64 matrix cores, for
// one env owning the whole wavefront: 16 x 16 tiles of H, each accumulated by
// v_mfma_f64_16x16x4f64 over the rows four at a time (inactive rows and the
// padding beyond nv give zeros). Lane l feeds A[i = l % 16][k = l / 16] and
// B[k = l / 16][j = l % 16] and holds H's column j = l % 16, rows l / 16,
// l / 16 + 4, l / 16 + 8 and l / 16 + 12 of the tile. The fused multiply-adds and the summation order
// differ from the reference's (d_r J_ri) J_rj added row by row, so results are
// not bit-identical with it; used only when the sim asks for it.
__device__ __noinline__ void hessian_mfma(const Real* qm, const Real* J, const Real* dd, const Real* act, int nr, int nv,
                                          Real* H) {
  const int l = (int)threadIdx.x, c16 = l & 15, q = l >> 4;
  const int nt = (nv + 15) / 16;  // tiles per side (nv <= 64)
  for (int ti = 0; ti < nt; ti++) {
    f64x4 acc[4];
    for (int tj = 0; tj < 4; tj++) acc[tj] = f64x4{0.0, 0.0, 0.0, 0.0};
    const int i = 16 * ti + c16;
    for (int s = 0; s < nr; s += 4) {
      const int r = s + q;
      const bool on = r < nr && act[r] != 0.0;
      const Real a = (on && i < nv) ? dd[r] * J[(int64_t)r * nv + i] : 0.0;
      for (int tj = 0; tj < 4; tj++) {
        if (tj >= nt) break;
        const int j = 16 * tj + c16;
        const Real b = (on && j < nv) ? J[(int64_t)r * nv + j] : 0.0;
        acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[tj], 0, 0, 0);
      }
    }
    // accumulator layout: lane l holds column l % 16, rows l / 16 + 4 t
    for (int tj = 0; tj < nt; tj++) {
      const int j = 16 * tj + c16;
      for (int t = 0; t < 4; t++) {
        const int ii = 16 * ti + q + 4 * t;
        if (ii < nv && j < nv) H[ii * nv + j] = qm[ii * nv + j] + acc[tj][t];
      }
    }
  }
}
#endif

template <class X, class MT>
PHD int newton_dense(const MT& M, Env& E, const NP& P) {
  const int nv = P.nv, nr = P.nr;
  Real* a = P.a;
  Real* a0 = LD(a0);
  Real* grad = LD(grad);
  Real* p = LD(p);
  Real* tmp = LD(tmp);
  Real* tmp2 = LD(tmp2);
  Real* mp = LD(mp);
  Real* H = P.fac;
  Real* sc = LD(scal);
  int* isc = E.si + M.L.i_iscal;
  const Real* qm = P.qm;
  const Real* J = P.J;
  Real* jar = E.jar;
  Real* jp = E.jp;
  Real* act = E.ract;
  Real* prev = E.rprev;
  X::par(nv, [&](int k) { a0[k] = a[k]; });
  const int nb = (nr + 63) / 64;
  // cost(x) = 1/2 (x-a0)'M(x-a0) + sum_active 1/2 D r^2 with r = J x - aref
  auto cost_into = [&](const Real* x, int slot) {
    X::par(nv, [&](int k) { tmp[k] = x[k] - a0[k]; });
    mul_m<X>(P, tmp, tmp2);
    X::par(nr, [&](int r) { jar[r] = sdot_n(J + r * nv, x, nv) - P.aref[r]; });
    X::par(1, [&](int) {
      Real s = rsum0();
      for (int k = 0; k < nv; k++) s += tmp[k] * tmp2[k];
      Real c = 0.5 * s;
      for (int r = 0; r < nr; r++) {
        bool on = r >= P.n_ineq || jar[r] < 0.0;
        if (on) c += 0.5 * P.dd[r] * jar[r] * jar[r];
      }
      sc[slot] = c;
    });
  };
  // warm start from the previous acceleration when cheaper (newton.rs:238-245)
  const Real* warm = P.warm;
  cost_into(a0, 0);
  cost_into(warm, 1);
  if (sc[1] < sc[0]) X::par(nv, [&](int k) { a[k] = warm[k]; });
  if (P.warm2) {
    // the learned candidate replaces the reference's choice only when cheaper
    cost_into(P.warm2, 4);
    if (sc[4] < (sc[1] < sc[0] ? sc[1] : sc[0])) X::par(nv, [&](int k) { a[k] = P.warm2[k]; });
  }
  mul_m<X>(P, a0, tmp);
  X::par(1, [&](int) {
    Real s = rsum0();
    for (int k = 0; k < nv; k++) s += tmp[k] * tmp[k];
    sc[2] = 1.0 + sqrt(s);  // scale
  });
  const Real scale = sc[2];
  const Real tol = M.tolerance;
  const int cap = M.iterations > 1 ? M.iterations : 1;
  const int n_ineq = P.n_ineq;
  X::par(nr, [&](int r) { prev[r] = -1.0; });  // u128::MAX sentinel
  int iters = 0;
  while (iters < cap) {
    iters++;
    X::par(nv, [&](int k) { tmp[k] = a[k] - a0[k]; });
    mul_m<X>(P, tmp, grad);
    X::par(nr, [&](int r) {
      Real v = sdot_n(J + r * nv, a, nv) - P.aref[r];
      jar[r] = v;
      bool on = r >= n_ineq || v < 0.0;
      act[r] = on ? 1.0 : 0.0;
      E.rf[r] = P.dd[r] * v;
    });
    X::par(nv, [&](int k) {
      Real g = grad[k];
      for (int r = 0; r < nr; r++)
        if (act[r] != 0.0) g += E.rf[r] * J[r * nv + k];
      grad[k] = g;
    });
    X::par(1, [&](int) {
      Real s = rsum0();
      for (int k = 0; k < nv; k++) s += grad[k] * grad[k];
      Real gnorm = sqrt(s);
      bool same = nr <= 128 && iters > 1;
      if (same)
        for (int r = 0; r < nr && r < 128; r++)
          if (act[r] != prev[r]) {
            same = false;
            break;
          }
      isc[6] = (gnorm < tol * scale || same) ? 1 : 0;
    });
    if (isc[6]) break;
    X::par(nr, [&](int r) { prev[r] = act[r]; });
    bool on_mfma = false;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (wave_lanes<X>::value == 64) {
      if (M.hess_mfma && nv <= 64) {  // four 16-column tiles per row of tiles
        hessian_mfma(qm, J, P.dd, act, nr, nv, H);
        __syncthreads();
        on_mfma = true;
      }
    }
#endif
    // H = M + J_A' D J_A, entry-parallel, rows in order
    if (!on_mfma) X::par(nv * nv, [&](int t) {
      int i = t / nv, j = t % nv;
      Real h = qm[t];
      if (i == j) h += 0.0;
      for (int r = 0; r < nr; r++) {
        if (act[r] == 0.0) continue;
        Real ji = J[r * nv + i];
        if (ji == 0.0) continue;
        h += (P.dd[r] * ji) * J[r * nv + j];
      }
      H[t] = h;
    });
    X::par(1, [&](int) { isc[1] = 0; });
    if (!chol_block<X>(M, E, H, nv, 0, nv)) break;
    X::par(nv, [&](int k) { p[k] = -grad[k]; });
    chol_solve_block<X>(M, E, H, nv, 0, nv, p);
    // exact line search along p (newton.rs:310-351)
    mul_m<X>(P, p, mp);
    X::par(nr, [&](int r) { jp[r] = sdot_n(J + r * nv, p, nv); });
    X::par(1, [&](int) {
      Real mpp = rsum0(), mpg = rsum0();
      for (int k = 0; k < nv; k++) mpp += p[k] * mp[k];
      for (int k = 0; k < nv; k++) mpg += tmp[k] * mp[k];
      auto deriv = [&](Real al, Real* d1o, Real* d2o) {
        Real d1 = al * mpp + mpg, d2 = mpp;
        for (int r = 0; r < nr; r++) {
          Real v = jar[r] + al * jp[r];
          if (r >= n_ineq || v < 0.0) {
            Real dq = P.dd[r] * jp[r];
            d1 += dq * v;
            d2 += dq * jp[r];
          }
        }
        *d1o = d1;
        *d2o = d2;
      };
      Real lo = 0.0, hi = INFINITY, al = 1.0, d1, d2;
      deriv(0.0, &d1, &d2);
      Real ltol = 1e-12 * rmax(fabs(d1), 1e-30);
      for (int it = 0; it < 50; it++) {
        deriv(al, &d1, &d2);
        if (fabs(d1) <= ltol) break;
        if (d1 < 0.0) lo = al;
        else hi = al;
        Real next = al - d1 / rmax(d2, 1e-30);
        if (next <= lo || next >= hi) next = rfinite(hi) ? 0.5 * (lo + hi) : 2.0 * rmax(al, lo);
        if (fabs(next - al) <= 1e-14 * rmax(fabs(al), 1.0)) {
          al = next;
          break;
        }
        al = next;
      }
      sc[3] = al;
    });
    const Real al = sc[3];
    X::par(nv, [&](int k) { a[k] += al * p[k]; });
  }
  (void)nb;
  // add_constraint_force (newton.rs:361-371)
  X::par(nr, [&](int r) {
    Real v = sdot_n(J + r * nv, a, nv) - P.aref[r];
    act[r] = (r >= n_ineq || v < 0.0) ? 1.0 : 0.0;
    E.rf[r] = -P.dd[r] * v;
  });
  X::par(nv, [&](int k) {
    Real f = 0.0;
    for (int r = 0; r < nr; r++)
      if (act[r] != 0.0) f += E.rf[r] * J[r * nv + k];
    P.qf[k] = f;
  });
  return iters;
}

// newton.rs:378-473. Rows join the dof blocks they touch into islands
// (union-find; an island is named by its smallest block, which is where the
// reference's parent[max] = min links leave the root). One island spanning
// every block is solved in place; otherwise each island, in increasing order,
// is gathered (its dofs block by block, its rows in order), solved, and
// scattered back. Iterations: the maximum over islands.
template <class X, class MT>
PHD int newton_solve(const MT& M, const DState& S, Env& E, int nr, const Real* warm2) {
  const int nv = M.nv;
  const int n_ineq = nr - M.neq;
  auto whole = [&]() {
    if (nv <= GL && nr <= GL && nr <= M.nrl) {
      NP P{nv, nr, n_ineq, LD(qm), LD(J), LD(aref), LD(dd), LD(qacc), LD(a), LD(qfrc_c), LD(chol), warm2};
      return newton_lanes<X>(M, E, P);
    }
    NP P{nv, nr, n_ineq, LD(qm), E.J, E.aref, E.dd, LD(qacc), LD(a), LD(qfrc_c), LD(chol), warm2};
    return newton_dense<X>(M, E, P);
  };
  if (M.nblock <= 1) return whole();
  // per-env island scratch: floats in HBM (gathered into LDS below when an
  // island fits), the bookkeeping ints in the env's LDS int region
  Real* fs = (S.isl_f + E.e * (int64_t)M.isl_fstride).p;
  int* is = E.si + M.L.i_isl;
  int* rmask = is;                 // nrmax: blocks each row touches
  int* sidx = rmask + M.nrmax;     // nv: gathered dofs
  int* sridx = sidx + nv;          // nrmax: gathered rows
  int* root = sridx + M.nrmax;     // nblock: island of each block
  int* isl = root + M.nblock;      // nblock + 1: island list, then its length
  Real* sqm = fs;
  Real* sJ = sqm + nv * nv;
  Real* saref = sJ + (int64_t)M.nrmax * nv;
  Real* sdd = saref + M.nrmax;
  Real* sa = sdd + M.nrmax;
  Real* swarm = sa + nv;
  Real* sqf = swarm + nv;
  Real* swarm2 = sqf + nv;
  int* isc = E.si + M.L.i_iscal;
  const Real* J = E.J;
  X::par(nr, [&](int r) {
    uint32_t m = 0;
    for (int k = 0; k < nv; k++)
      if (J[r * nv + k] != 0.0) m |= 1u << M.dof_block[k];
    rmask[r] = (int)m;
  });
  X::par(1, [&](int) {
    const int nb = M.nblock;
    for (int b = 0; b < nb; b++) root[b] = b;
    auto find = [&](int x) {
      while (root[x] != x) x = root[x];
      return x;
    };
    for (int r = 0; r < nr; r++) {
      uint32_t m = (uint32_t)rmask[r];
      if (!m) continue;
      int first = __builtin_ctz(m);
      for (int b = first + 1; b < nb; b++)
        if ((m >> b) & 1u) {
          int x = find(first), y = find(b);
          if (x != y) root[x > y ? x : y] = x < y ? x : y;
        }
    }
    for (int b = 0; b < nb; b++) root[b] = find(b);
    // islands holding a row, in increasing order
    int ni = 0;
    for (int b = 0; b < nb; b++) {
      bool used = false;
      for (int r = 0; r < nr && !used; r++) used = rmask[r] && root[__builtin_ctz((uint32_t)rmask[r])] == b;
      if (used) isl[ni++] = b;
    }
    bool all = ni == 1;
    for (int b = 0; b < nb && all; b++) all = root[b] == isl[0];
    isl[nb] = all ? -1 : ni;
  });
  if (isl[M.nblock] < 0) return whole();
  const int ni = isl[M.nblock];
  int iters = 0;
  for (int ii = 0; ii < ni; ii++) {
    X::par(1, [&](int) {
      const int id = isl[ii];
      int k = 0, q = 0, qi = 0;
      for (int b = 0; b < M.nblock; b++)
        if (root[b] == id)
          for (int d = M.block_adr[b]; d < M.block_adr[b + 1]; d++) sidx[k++] = d;
      for (int r = 0; r < nr; r++)
        if (rmask[r] && root[__builtin_ctz((uint32_t)rmask[r])] == id) {
          sridx[q++] = r;
          if (r < n_ineq) qi++;
        }
      isc[10] = k;
      isc[11] = q;
      isc[12] = qi;
    });
    const int k = isc[10], q = isc[11], qi = isc[12];
    // An island the register solver takes (k, q <= 16) is gathered into the
    // LDS of the mass-factor slot, after the k x k Hessian factor that the
    // solver keeps there, when it fits; otherwise into the HBM island scratch.
    const bool lds_isl = k <= GL && q <= GL && 2 * k * k + q * k + 2 * q + 4 * k <= nv * nv;
    if (lds_isl) {
      sqm = LD(chol) + k * k;
      sJ = sqm + k * k;
      saref = sJ + q * k;
      sdd = saref + q;
      sa = sdd + q;
      swarm = sa + k;
      sqf = swarm + k;
      swarm2 = sqf + k;
    } else {
      sqm = fs;
      sJ = sqm + nv * nv;
      saref = sJ + (int64_t)M.nrmax * nv;
      sdd = saref + M.nrmax;
      sa = sdd + M.nrmax;
      swarm = sa + nv;
      sqf = swarm + nv;
      swarm2 = sqf + nv;
    }
    X::par(k * k, [&](int t) { sqm[t] = LD(qm)[sidx[t / k] * nv + sidx[t % k]]; });
    X::par(q * k, [&](int t) { sJ[t] = J[sridx[t / k] * nv + sidx[t % k]]; });
    X::par(q, [&](int t) {
      saref[t] = E.aref[sridx[t]];
      sdd[t] = E.dd[sridx[t]];
    });
    X::par(k, [&](int t) {
      sa[t] = LD(a)[sidx[t]];
      swarm[t] = LD(qacc)[sidx[t]];
      if (warm2) swarm2[t] = warm2[sidx[t]];
    });
    NP P{k, q, qi, sqm, sJ, saref, sdd, swarm, sa, sqf, LD(chol), warm2 ? swarm2 : nullptr};
    int it = (k <= GL && q <= GL) ? newton_lanes<X>(M, E, P) : newton_dense<X>(M, E, P);
    iters = it > iters ? it : iters;
    X::par(k, [&](int t) {
      LD(a)[sidx[t]] = sa[t];
      LD(qfrc_c)[sidx[t]] = sqf[t];
    });
  }
  return iters;
}

// ---------------- PGS, solver.rs:44-289 ----------------
// math.rs:449-463 on one lane: the substitutions in the reference's order.
PHD void solve_serial(const Real* L, int nv, int s0, int e, Real* x) {
  for (int i = s0; i < e; i++) x[i] = (x[i] - sdot_n(L + i * nv + s0, x + s0, i - s0)) / L[i * nv + i];
  for (int i = e - 1; i >= s0; i--) {
    x[i] /= L[i * nv + i];
    Real xi = x[i];
    const Real* li = L + i * nv + s0;
    for (int k = 0; k < i - s0; k++) x[s0 + k] += -xi * li[k];
  }
}

// Row record of the PGS problem (HBM, 10 doubles per row).
enum { PR_SEP = 0, PR_MU, PR_INVA, PR_LAM, PR_KIND, PR_NORMAL, PR_CONTACT };
enum { PK_NORMAL = 0, PK_FRICTION = 1, PK_BILATERAL = 2 };

// Soft-constraint coefficients and stopping rule of the sweeps (solver.rs:44-50, :229-240).
struct PgsConsts {
  Real h, slop, bias_rate, mass_scale, impulse_scale, max_bias, tol;
  int iterations;
};

// One row's new force (solver.rs:175-207): jv = J_r v, lam = the row's force,
// lam_n = the force of its normal row (friction rows).
PHD Real pgs_row_force(const PgsConsts& C, int kind, Real inv_a, Real sep, Real mu, Real lam, Real lam_n, Real jv,
                       bool use_bias) {
  if (kind == PK_NORMAL) {
    Real sp = sep + C.slop;
    Real bias, ms, is;
    if (sp > 0.0) {
      bias = sp / C.h;
      ms = 1.0;
      is = 0.0;
    } else if (use_bias) {
      bias = rmax(C.bias_rate * sp, -C.max_bias);
      ms = C.mass_scale;
      is = C.impulse_scale;
    } else {
      bias = 0.0;
      ms = 1.0;
      is = 0.0;
    }
    Real delta = -inv_a * ms * (jv + bias) - is * lam;
    return rmax(lam + delta, 0.0);
  }
  if (kind == PK_BILATERAL) {
    Real bias, ms, is;
    if (use_bias) {
      bias = rclamp(C.bias_rate * sep, -C.max_bias, C.max_bias);
      ms = C.mass_scale;
      is = C.impulse_scale;
    } else {
      bias = 0.0;
      ms = 1.0;
      is = 0.0;
    }
    return lam - inv_a * ms * (jv + bias) - is * lam;
  }
  Real bound = mu * lam_n;
  return rclamp(lam - inv_a * jv, -bound, bound);
}

#if defined(__HIP_DEVICE_COMPILE__)
// x as held by lane l (l wave-uniform).
__device__ __forceinline__ Real lane_read(Real x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// All Gauss-Seidel sweeps of one env that owns the whole wavefront (nv <= 64,
// nr <= 128), with the env's state in registers: lane k holds v_k and the
// J_rk and (M^-1 J^T)_rk of the next AHEAD rows, loaded ahead into buffers
// that the rows take in turn (lanes past nv hold values nothing reads); lane l
// holds rows l and l + 64: their forces, records and the terms of their
// updates that stay fixed over the sweeps (bias, mass and impulse scale of
// each phase, with the reference's operations). Every lane computes each
// row's new force, so no barrier or LDS round trip sits between rows. J_r v is
// summed in dof order over the dofs where J_rk is nonzero: adding a zero
// product to a nonzero partial sum leaves it unchanged, so this equals the
// reference's in-order sum over all dofs unless the result is zero or not
// finite, or v holds a non-finite entry (a zero J_rk times an infinite v_k is
// NaN); in those cases the full sum runs. The sweep's largest velocity change
// max_r |dlambda_r| / inv_a_r is formed after the sweep, from each lane's
// rows: each row changes at most once per sweep and the maximum does not
// depend on the order. Each lane passes its entry
// of v and receives its entries of v and vpos; the records' forces are
// written before the return.
//
// The loads ahead are issued in one fixed order (J then M^-1 J^T, row after
// row) on every path, each followed by a compiler barrier, so that the
// compiler neither sinks them to their use nor waits for more than the row it
// needs.
#define PH_LOAD_FENCE() asm volatile("" ::: "memory")
constexpr int AHEAD = 4;  // rows whose J and M^-1 J^T are in flight
struct LaneVel {
  Real v, vpos;  // this lane's entry of the solved velocity and of the one positions integrate with
  int sweeps;    // sweeps run, both phases
};
constexpr int PGS_WAVE_ROWS = 128;  // rows the register sweeps hold: two per lane
__device__ __forceinline__ LaneVel pgs_sweeps_body(const PgsConsts C, const Real* Jp, const Real* mjp, Real* rwp,
                                                   int nr_in, int nv_in, Real v_in) {
  // sizes and counts are the same in every lane: keep them scalar so that the
  // loops are uniform
  const int nr = __builtin_amdgcn_readfirstlane(nr_in), nv = __builtin_amdgcn_readfirstlane(nv_in);
  const int iterations = __builtin_amdgcn_readfirstlane(C.iterations);
  typedef const PH_AS_GLOBAL Real* GC;
  const GC J = (GC)Jp, mj = (GC)mjp;
  PH_AS_GLOBAL Real* rw = (PH_AS_GLOBAL Real*)rwp;
  const int lane = (int)threadIdx.x;
  const bool dof_lane = lane < nv;
  const int kl = dof_lane ? lane : nv - 1;
  Real vk = dof_lane ? v_in : 0.0;
  // row lane + 64 s in slot s: force, record, and the bias and scales of the
  // phase with bias (the phase without has the bias bias0, scale 1, impulse 0)
  Real inva[2] = {0.0, 0.0}, mu[2] = {0.0, 0.0}, lam[2] = {0.0, 0.0};
  Real bias1[2] = {0.0, 0.0}, ms1[2] = {1.0, 1.0}, is1[2] = {0.0, 0.0}, bias0[2] = {0.0, 0.0};
  int kind[2] = {0, 0}, nrm[2] = {0, 0};
#pragma unroll
  for (int sl = 0; sl < 2; sl++) {
    const int row_id = lane + 64 * sl;
    if (row_id >= nr) continue;
    const PH_AS_GLOBAL Real* q = rw + row_id * 10;
    inva[sl] = q[PR_INVA];
    const Real sep = q[PR_SEP];
    mu[sl] = q[PR_MU];
    lam[sl] = q[PR_LAM];
    kind[sl] = (int)q[PR_KIND];
    nrm[sl] = (int)q[PR_NORMAL];
    if (kind[sl] == PK_NORMAL) {
      const Real sp = sep + C.slop;
      if (sp > 0.0) {
        bias1[sl] = bias0[sl] = sp / C.h;
      } else {
        bias1[sl] = rmax(C.bias_rate * sp, -C.max_bias);
        ms1[sl] = C.mass_scale;
        is1[sl] = C.impulse_scale;
      }
    } else if (kind[sl] == PK_BILATERAL) {
      bias1[sl] = rclamp(C.bias_rate * sep, -C.max_bias, C.max_bias);
      ms1[sl] = C.mass_scale;
      is1[sl] = C.impulse_scale;
    }
  }
  // the value row r holds (r wave-uniform)
  auto at = [&](const Real (&x)[2], int r) { return r < 64 ? lane_read(x[0], r) : lane_read(x[1], r - 64); };
  auto at_i = [&](const int (&x)[2], int r) {
    return r < 64 ? __builtin_amdgcn_readlane(x[0], r) : __builtin_amdgcn_readlane(x[1], r - 64);
  };
  auto v_finite = [&]() { return __ballot(dof_lane && !__builtin_isfinite(vk)) == 0; };
  bool vfin = v_finite();
  auto sweep = [&](bool use_bias) {
    Real dq[2] = {0.0, 0.0};  // |dlambda| of this lane's rows, if they changed in this sweep
    bool upd[2] = {false, false};
    // row r with its J and M^-1 J^T entries in jb, mb; reloads them with row
    // r + AHEAD. Every lane loads (lanes past nv repeat the last dof, rows past
    // the end repeat the last row), so no branch sits around a load.
    auto row = [&](int r, Real& jb, Real& mb) {
      const int64_t off = (int64_t)(r + AHEAD < nr ? r + AHEAD : nr - 1) * nv + kl;
      const Real inv_a = at(inva, r);
      const Real p = dof_lane ? jb * vk : 0.0;  // lanes past nv hold 0
      const uint64_t nz = __ballot(dof_lane && jb != 0.0);
      // both loads sit on the path of every row, skipped or not, so that the
      // loads in flight are the same whichever way a row goes
      jb = J[off];
      PH_LOAD_FENCE();
      if (inv_a != 0.0) {
        Real jv = rsum0();
        for (uint64_t m = nz; m; m &= m - 1) jv += lane_read(p, __builtin_ctzll(m));
        if (!vfin || jv == 0.0 || !__builtin_isfinite(jv)) {
          jv = rsum0();
          for (int k = 0; k < nv; k++) jv += lane_read(p, k);
        }
        const Real lm = at(lam, r);
        const int kd = at_i(kind, r);
        Real nw;
        if (kd == PK_FRICTION) {
          const Real bound = at(mu, r) * at(lam, at_i(nrm, r));
          nw = rclamp(lm - inv_a * jv, -bound, bound);
        } else {
          const Real bias = use_bias ? at(bias1, r) : at(bias0, r);
          const Real ms = use_bias ? at(ms1, r) : 1.0;
          const Real is = use_bias ? at(is1, r) : 0.0;
          if (kd == PK_NORMAL) {
            const Real delta = -inv_a * ms * (jv + bias) - is * lm;
            nw = rmax(lm + delta, 0.0);
          } else {
            nw = lm - inv_a * ms * (jv + bias) - is * lm;
          }
        }
        const Real dl = nw - lm;
        if (dl != 0.0) {
          const int sl = r < 64 ? 0 : 1;
          if (lane == r - 64 * sl) {
            if (sl == 0) {
              lam[0] = nw;
              dq[0] = fabs(dl);
              upd[0] = true;
            } else {
              lam[1] = nw;
              dq[1] = fabs(dl);
              upd[1] = true;
            }
          }
          vk += mb * dl;
          vfin = v_finite();
        }
      }
      mb = mj[off];
      PH_LOAD_FENCE();
    };
    Real jbuf[AHEAD], mbuf[AHEAD];
#pragma unroll
    for (int b = 0; b < AHEAD; b++) {
      const int64_t off = (int64_t)(b < nr ? b : nr - 1) * nv + kl;
      jbuf[b] = J[off];
      PH_LOAD_FENCE();
      mbuf[b] = mj[off];
      PH_LOAD_FENCE();
    }
    int r = 0;
    for (; r + AHEAD <= nr; r += AHEAD) {  // every buffer on every pass: the loads in flight are alike at each pass
#pragma unroll
      for (int b = 0; b < AHEAD; b++) row(r + b, jbuf[b], mbuf[b]);
    }
#pragma unroll
    for (int b = 0; b < AHEAD - 1; b++)
      if (r + b < nr) row(r + b, jbuf[b], mbuf[b]);
    Real q = upd[0] ? dq[0] / inva[0] : 0.0;
    if (upd[1]) q = rmax(q, dq[1] / inva[1]);
    for (int o = 32; o >= 1; o >>= 1) q = rmax(q, __shfl_xor(q, o));
    return q;
  };
  LaneVel out;
  out.sweeps = 0;
  for (int it = 0; it < iterations; it++) {
    out.sweeps++;
    if (sweep(true) < C.tol) break;
  }
  out.vpos = vk;
  const int relax = iterations / 3 > 2 ? iterations / 3 : 2;
  for (int it = 0; it < relax; it++) {
    out.sweeps++;
    if (sweep(false) < C.tol) break;
  }
  out.v = vk;
#pragma unroll
  for (int sl = 0; sl < 2; sl++)
    if (lane + 64 * sl < nr) rw[(lane + 64 * sl) * 10 + PR_LAM] = lam[sl];
  __syncthreads();
  return out;
}
// Out of line inside the step kernel (whose registers it would otherwise add
// to); the sweeps kernel inlines the body, so its register budget applies.
__device__ __noinline__ LaneVel pgs_sweeps_wave(const PgsConsts C, const Real* Jp, const Real* mjp, Real* rwp, int nr_in,
                                                int nv_in, Real v_in) {
  return pgs_sweeps_body(C, Jp, mjp, rwp, nr_in, nv_in, v_in);
}
#undef PH_LOAD_FENCE
#endif

// solver.rs:63-251, in three parts: pgs_prepare (rows, M^-1 J^T, warm
// start), the Gauss-Seidel sweeps, pgs_finish (constraint force, history for
// the next warm start). v: velocity after the smooth update (in, out); vpos:
// the biased velocity the positions integrate with. Rows, M^-1 J^T and the row
// records live in HBM; the rows, the M^-1 J^T columns and the constraint force
// are formed lane-parallel. A sweep is sequential over rows: when one env owns
// the wavefront it runs on registers (pgs_sweeps_wave), otherwise one lane per
// env forms each row's force between lane-parallel products and velocity
// updates.
//
// v += M^-1 J^T lambda over the rows' warm-start forces (solver.rs:155-161).
template <class X, class MT>
PHD void pgs_apply_warm(const MT& M, const DState& S, Env& E, int nr, Real* v) {
  const int nv = M.nv;
  const Real* mj = (S.gmj + E.e * (int64_t)M.nrmax * nv).p;
  const Real* rw = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
  X::par(nv, [&](int k) {
    for (int r = 0; r < nr; r++) {
      Real l = rw[r * 10 + PR_LAM];
      if (l != 0.0) v[k] += mj[r * nv + k] * l;
    }
  });
}

// pgs_prepare returns the row count, 0 when there is nothing to solve (then v,
// vpos and the history are final), or -1 when the rows overflowed. Without
// apply_warm it stops with the warm-start forces in the row records, before
// they move v (pgs_apply_warm).
template <class X, class MT>
PHD int pgs_prepare(const MT& M, const DState& S, Env& E, Real* v, Real* vpos, bool apply_warm = true) {
  const int nv = M.nv;
  const Real h = M.timestep;
  int* isc = E.si + M.L.i_iscal;
  auto con = S.con + E.e * (int64_t)M.ncmax * CW;
  auto cg = S.cgeom + E.e * (int64_t)M.ncmax * 3;
  Real* J = (S.gjac + E.e * (int64_t)M.nrmax * nv).p;
  Real* mj = (S.gmj + E.e * (int64_t)M.nrmax * nv).p;
  Real* rw = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
  const Real* qpos = LD(qpos);
  const int ncon = isc[0];
  // contact rows: normal, then the two friction directions (solver.rs:67-91)
  X::par(ncon * nv, [&](int t) {
    int c = t / nv, dof = t % nv;
    auto cc = con + c * CW;
    int b1 = M.geom_body[cg[3 * c]], b2 = M.geom_body[cg[3 * c + 1]];
    V3 t1, t2;
    contact_frame(cc, &t1, &t2);
    V3 p = ld3(cc);
    V3 dirs[3] = {ld3(cc + 3), t1, t2};
    SP cd = ldsp(LD(cdof) + 6 * dof);
    bool on2 = (M.body_chain[b2] >> dof) & 1, on1 = (M.body_chain[b1] >> dof) & 1;
    for (int r = 0; r < 3; r++) {
      Real jv = 0.0;
      if (on2) jv += 1.0 * dot(dirs[r], point_velocity(cd, p));
      if (on1) jv += -1.0 * dot(dirs[r], point_velocity(cd, p));
      J[(3 * c + r) * nv + dof] = jv;
    }
  });
  X::par(ncon, [&](int c) {
    int g1 = cg[3 * c], g2 = cg[3 * c + 1];
    Real mu = rmax(M.v_friction[(int64_t)E.var * M.ngeom + g1], M.v_friction[(int64_t)E.var * M.ngeom + g2]);
    for (int r = 0; r < 3; r++) {
      Real* q = rw + (3 * c + r) * 10;
      q[PR_SEP] = -con[c * CW + 9];
      q[PR_MU] = mu;
      q[PR_INVA] = 0.0;
      q[PR_LAM] = 0.0;
      q[PR_KIND] = r == 0 ? PK_NORMAL : PK_FRICTION;
      q[PR_NORMAL] = 3 * c;
      q[PR_CONTACT] = c;
    }
  });
  // limits within reach of this step, then equalities (solver.rs:92-127)
  X::par(1, [&](int) {
    int r = 3 * ncon;
    auto row = [&](Real sep, int kind) {
      Real* q = rw + r * 10;
      q[PR_SEP] = sep;
      q[PR_MU] = 0.0;
      q[PR_INVA] = 0.0;
      q[PR_LAM] = 0.0;
      q[PR_KIND] = kind;
      q[PR_NORMAL] = 0;
      q[PR_CONTACT] = -1;
      r++;
    };
    for (int j = 0; j < M.njnt; j++) {
      int t = M.jnt_type[j];
      if (!M.jnt_limited[j] || !(t == J_HINGE || t == J_SLIDE)) continue;
      Real q = qpos[M.jnt_qposadr[j]];
      int da = M.jnt_dofadr[j];
      Real reach = M.margin + 2.0 * h * fabs(v[da]);
      Real seps[2] = {q - M.jnt_range[2 * j], M.jnt_range[2 * j + 1] - q}, signs[2] = {1.0, -1.0};
      for (int sg = 0; sg < 2; sg++)
        if (seps[sg] < reach) {
          if (r >= M.nrmax) {
            isc[7] |= ST_ROW_OVERFLOW;
            return;
          }
          for (int k = 0; k < nv; k++) J[r * nv + k] = 0.0;
          J[r * nv + da] = signs[sg];
          row(seps[sg], PK_NORMAL);
        }
    }
    for (int e = 0; e < M.neq; e++) {
      if (r >= M.nrmax) {
        isc[7] |= ST_ROW_OVERFLOW;
        return;
      }
      int j1 = M.eq_j1[e], j2 = M.eq_j2[e];
      auto c = M.eq_poly + 5 * e;
      Real q1 = qpos[M.jnt_qposadr[j1]] - M.qpos0[M.jnt_qposadr[j1]];
      for (int k = 0; k < nv; k++) J[r * nv + k] = 0.0;
      J[r * nv + M.jnt_dofadr[j1]] += 1.0;
      Real poly = c[0];
      if (j2 >= 0) {
        Real x = qpos[M.jnt_qposadr[j2]] - M.qpos0[M.jnt_qposadr[j2]];
        poly = c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4])));
        J[r * nv + M.jnt_dofadr[j2]] -= c[1] + x * (2.0 * c[2] + x * (3.0 * c[3] + x * 4.0 * c[4]));
      }
      row(q1 - poly, PK_BILATERAL);
    }
    isc[9] = r;
  });
  const int nr = isc[9];
  if (isc[7] & ST_ROW_OVERFLOW) return -1;
  if (nr == 0) {
    X::par(nv, [&](int k) {
      vpos[k] = v[k];
      LD(qfrc_c)[k] = 0.0;
    });
    X::par(1, [&](int) { S.npcon[E.e] = 0; });
    return 0;
  }
  // M^-1 J_r^T and the row's inverse effective mass, one row per lane
  X::par(nr, [&](int r) {
    Real* x = mj + r * nv;
    for (int k = 0; k < nv; k++) x[k] = J[r * nv + k];
    for (int blk = 0; blk < M.nblock; blk++) solve_serial(LD(chol), nv, M.block_adr[blk], M.block_adr[blk + 1], x);
    Real a = rsum0();
    for (int k = 0; k < nv; k++) a += J[r * nv + k] * x[k];
    rw[r * 10 + PR_INVA] = a > 1e-12 ? 1.0 / a : 0.0;
  });
  // warm start from the matching previous contact (solver.rs:253-289)
  auto pcon = S.pcon + E.e * (int64_t)M.ncmax * CW;
  auto pcg = S.pcgeom + E.e * (int64_t)M.ncmax * 3;
  auto plam = S.plambda + E.e * (int64_t)M.ncmax * 3;
  const int npc = S.npcon[E.e];
  X::par(ncon, [&](int c) {
    if (npc == 0) return;
    auto cc = con + c * CW;
    Real bd = 0.01 * 0.01;
    int bi = -1;
    V3 cp = ld3(cc);
    for (int pi = 0; pi < npc; pi++) {
      if (pcg[3 * pi] == cg[3 * c] && pcg[3 * pi + 1] == cg[3 * c + 1]) {
        Real d2 = norm2(ld3(pcon + pi * CW) - cp);
        if (d2 < bd) {
          bd = d2;
          bi = pi;
        }
      }
    }
    if (bi >= 0) {
      V3 pt1, pt2, t1, t2;
      contact_frame(pcon + bi * CW, &pt1, &pt2);
      contact_frame(cc, &t1, &t2);
      V3 f = pt1 * plam[3 * bi + 1] + pt2 * plam[3 * bi + 2];
      rw[(3 * c) * 10 + PR_LAM] = plam[3 * bi];
      rw[(3 * c + 1) * 10 + PR_LAM] = dot(f, t1);
      rw[(3 * c + 2) * 10 + PR_LAM] = dot(f, t2);
    }
  });
  if (apply_warm) pgs_apply_warm<X>(M, S, E, nr, v);
  return nr;
}

// soft-constraint coefficients (solver.rs:44-50) and the sweeps' stopping rule
template <class MT>
PHD PgsConsts pgs_consts(const MT& M) {
  const Real h = M.timestep;
  const Real omega = 2.0 * 3.141592653589793 * M.hertz;
  const Real a1 = 2.0 * M.zeta + h * omega;
  const Real a2 = h * omega * a1;
  const Real a3 = 1.0 / (1.0 + a2);
  return PgsConsts{h, M.slop, omega / a1, a2 * a3, a3, 1.0, 1e-5, M.iterations};
}

// The Gauss-Seidel sweeps over nr > 0 rows (solver.rs:149-245); returns the
// number of sweeps run.
template <class X, class MT>
PHD int pgs_sweeps(const MT& M, const DState& S, Env& E, int nr, Real* v, Real* vpos) {
  const int nv = M.nv;
  Real* J = (S.gjac + E.e * (int64_t)M.nrmax * nv).p;
  Real* mj = (S.gmj + E.e * (int64_t)M.nrmax * nv).p;
  Real* rw = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
  const PgsConsts C = pgs_consts(M);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (wave_lanes<X>::value == 64) {
    if (nr <= PGS_WAVE_ROWS) {
      const int lane = (int)threadIdx.x;
      const LaneVel r = pgs_sweeps_wave(C, J, mj, rw, nr, nv, lane < nv ? v[lane] : 0.0);
      if (lane < nv) {
        v[lane] = r.v;
        vpos[lane] = r.vpos;
      }
      __syncthreads();
      return r.sweeps;
    }
  }
#endif
  // Gauss-Seidel sweeps (solver.rs:149-245): rows in order; for each row the
  // env's lanes form the products J_rk v_k and, after one lane has summed
  // them in dof order and updated the row's force, the update v += M^-1 J_r^T
  // dlambda, so that the rows' HBM data is read by all lanes at once
  Real* prod = LD(tmp);
  Real* sc = LD(scal);
  auto iterate = [&](bool use_bias) {
    X::par(1, [&](int) { sc[8] = 0.0; });  // largest velocity change of the sweep
    for (int r = 0; r < nr; r++) {
      Real* q = rw + r * 10;
      const Real inv_a = q[PR_INVA];
      if (inv_a == 0.0) continue;
      const Real* jr = J + r * nv;
      X::par(nv, [&](int k) { prod[k] = jr[k] * v[k]; });
      X::par(1, [&](int) {
        Real jv = rsum0();
        for (int k = 0; k < nv; k++) jv += prod[k];
        const Real lam = q[PR_LAM];
        const int kind = (int)q[PR_KIND];
        const Real lam_n = kind == PK_FRICTION ? rw[(int)q[PR_NORMAL] * 10 + PR_LAM] : 0.0;
        const Real nw = pgs_row_force(C, kind, inv_a, q[PR_SEP], q[PR_MU], lam, lam_n, jv, use_bias);
        const Real dl = nw - lam;
        sc[9] = dl;
        if (dl != 0.0) {
          sc[8] = rmax(sc[8], fabs(dl) / inv_a);
          q[PR_LAM] = nw;
        }
      });
      const Real dl = sc[9];
      if (dl != 0.0) {
        const Real* mr = mj + r * nv;
        X::par(nv, [&](int k) { v[k] += mr[k] * dl; });
      }
    }
    return sc[8];
  };
  int sweeps = 0;
  for (int it = 0; it < M.iterations; it++) {
    sweeps++;
    if (iterate(true) < C.tol) break;
  }
  X::par(nv, [&](int k) { vpos[k] = v[k]; });
  const int relax = M.iterations / 3 > 2 ? M.iterations / 3 : 2;
  for (int it = 0; it < relax; it++) {
    sweeps++;
    if (iterate(false) < C.tol) break;
  }
  return sweeps;
}

// The constraint force and the history for the next warm start, after the
// sweeps over nr > 0 rows.
template <class X, class MT>
PHD void pgs_finish(const MT& M, const DState& S, Env& E, int nr) {
  const int nv = M.nv;
  const Real h = M.timestep;
  int* isc = E.si + M.L.i_iscal;
  auto con = S.con + E.e * (int64_t)M.ncmax * CW;
  auto cg = S.cgeom + E.e * (int64_t)M.ncmax * 3;
  Real* J = (S.gjac + E.e * (int64_t)M.nrmax * nv).p;
  Real* rw = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
  auto pcon = S.pcon + E.e * (int64_t)M.ncmax * CW;
  auto pcg = S.pcgeom + E.e * (int64_t)M.ncmax * 3;
  auto plam = S.plambda + E.e * (int64_t)M.ncmax * 3;
  const int ncon = isc[0];
  // qfrc_constraint = sum_r J_r lambda_r / h, rows in order
  X::par(nv, [&](int k) {
    Real f = 0.0;
    for (int r = 0; r < nr; r++) f += J[r * nv + k] * rw[r * 10 + PR_LAM] / h;
    LD(qfrc_c)[k] = f;
  });
  // history for the next warm start: this step's contacts and their forces
  X::par(ncon * CW, [&](int t) { pcon[t] = con[t]; });
  X::par(ncon, [&](int c) {
    for (int i = 0; i < 3; i++) {
      pcg[3 * c + i] = cg[3 * c + i];
      plam[3 * c + i] = rw[(3 * c + i) * 10 + PR_LAM];
    }
  });
  X::par(1, [&](int) { S.npcon[E.e] = ncon; });
}

template <class X, class MT>
PHD void pgs_solve(const MT& M, const DState& S, Env& E, Real* v, Real* vpos) {
  const int nr = pgs_prepare<X>(M, S, E, v, vpos);
  if (nr <= 0) return;
  pgs_sweeps<X>(M, S, E, nr, v, vpos);
  pgs_finish<X>(M, S, E, nr);
}

// ---------------- integrate_pos, forward.rs:459-480 ----------------
template <class X, class MT>
PHD void integrate_pos(const MT& M, Env& E, const Real* v, Real h) {
  Real* qpos = LD(qpos);
  X::par(M.njnt, [&](int j) {
    int qa = M.jnt_qposadr[j], da = M.jnt_dofadr[j], t = M.jnt_type[j];
    if (t == J_FREE) {
      for (int k = 0; k < 3; k++) qpos[qa + k] += h * v[da + k];
      Q4 q = ldq(qpos + qa + 3);
      Q4 r = integrate_local(q, v3(v[da + 3], v[da + 4], v[da + 5]), h);
      stq(qpos + qa + 3, r);
    } else if (t == J_BALL) {
      Q4 r = integrate_local(ldq(qpos + qa), v3(v[da], v[da + 1], v[da + 2]), h);
      stq(qpos + qa, r);
    } else {
      qpos[qa] += h * v[da];
    }
  });
}

// ---------------- one step, forward.rs:336-389 (Newton) ----------------
enum { STEP_FAILED = -1, STEP_DONE = -2 };

// The end of the PGS step (forward.rs:336-389, PGS branch): qacc from the
// velocity change, positions integrated with vpos (LD(acc)), qvel = LD(vel).
template <class X, class MT>
PHD void pgs_step_end(const MT& M, Env& E) {
  const Real h = M.timestep;
  Real* vel = LD(vel);
  X::par(M.nv, [&](int k) { LD(qacc)[k] = (vel[k] - LD(qvel)[k]) / h; });
  integrate_pos<X>(M, E, LD(acc), h);
  X::par(M.nv, [&](int k) { LD(qvel)[k] = vel[k]; });
}

// The step up to the constraint rows: forward dynamics, collision, the
// unconstrained acceleration a0 (in LD(a)) and the rows. Returns the row count,
// STEP_FAILED when the env cannot be stepped (status set) or STEP_DONE when the
// PGS branch has completed the step. With pgs_stop the PGS branch stops
// early and returns the row count: PGS_BEFORE_SWEEPS with the warm start
// applied (velocity in LD(vel); with no rows the step's velocities are final
// and LD(acc) holds vpos), PGS_BEFORE_WARM with the warm-start forces still in
// the row records.
enum { PGS_FULL = 0, PGS_BEFORE_SWEEPS = 1, PGS_BEFORE_WARM = 2 };
template <class X, class MT>
PHD int step_front(const MT& M, const DState& S, Env& E, int pgs_stop = PGS_FULL) {
  const int nv = M.nv;
  const Real h = M.timestep;
  int* isc = E.si + M.L.i_iscal;
  kinematics<X>(M, E);
  crba<X>(M, E);
  velocity_rne<X>(M, E);
  passive_actuation<X>(M, E);
  collide<X>(M, S, E);
  if (isc[7] & ST_CONTACT_OVERFLOW) return STEP_FAILED;
  Real* smooth = LD(smooth);
  Real* a = LD(a);
  X::par(nv, [&](int k) {
    Real s = LD(passive)[k] + LD(qfrc_act)[k] + LD(applied)[k] - LD(bias)[k];
    smooth[k] = s;
    a[k] = s;
  });
  X::par(1, [&](int) {
    isc[1] = 0;
    isc[5] = 0;
  });
  if (M.solver_kind == 1) {
    // PGS branch of forward.rs:336-389: M + h D, the smooth velocity, then
    // PGS on velocities; positions integrate with the biased velocity
    if (!factor_solve_mass<X>(M, E, h, a)) {
      X::par(1, [&](int) { isc[7] |= ST_NOTPD; });
      return STEP_FAILED;
    }
    Real* vel = LD(vel);
    Real* vpos = LD(acc);
    X::par(nv, [&](int k) { vel[k] = LD(qvel)[k] + h * a[k]; });
    if (pgs_stop != PGS_FULL) {
      const int nr = pgs_prepare<X>(M, S, E, vel, vpos, pgs_stop == PGS_BEFORE_SWEEPS);
      return nr < 0 ? STEP_FAILED : nr;
    }
    pgs_solve<X>(M, S, E, vel, vpos);
    if (isc[7] & ST_ROW_OVERFLOW) return STEP_FAILED;
    pgs_step_end<X>(M, E);
    return STEP_DONE;
  }
  if (!factor_solve_mass<X>(M, E, 0.0, a)) {
    X::par(1, [&](int) { isc[7] |= ST_NOTPD; });
    return STEP_FAILED;
  }
  X::par(nv, [&](int k) { LD(qfrc_c)[k] = 0.0; });
  int nr = build_rows<X>(M, S, E);
  if (nr < 0) {
    X::par(1, [&](int) { isc[7] |= ST_ROW_OVERFLOW; });
    return STEP_FAILED;
  }
  X::par(1, [&](int) { isc[9] = nr; });
  return nr;
}

// The rest of the Newton step: solve, damping, integration. warm2 is an extra
// warm-start candidate (null: the reference's rule alone).
template <class X, class MT>
PHD bool step_back(const MT& M, const DState& S, Env& E, int nr, const Real* warm2) {
  const int nv = M.nv;
  const Real h = M.timestep;
  int* isc = E.si + M.L.i_iscal;
  Real* smooth = LD(smooth);
  Real* acc = LD(acc);
  Real* a = LD(a);
  if (nr > 0) {
    int it = newton_solve<X>(M, S, E, nr, warm2);
    X::par(1, [&](int) { isc[8] = it; });
  }
  const bool damped = M.v_damped[E.var] != 0;
  if (damped) {
    X::par(1, [&](int) { isc[1] = 0; });
    X::par(nv, [&](int k) { acc[k] = smooth[k] + LD(qfrc_c)[k]; });
    if (!factor_solve_mass<X>(M, E, h, acc)) {
      X::par(1, [&](int) { isc[7] |= ST_NOTPD; });
      return false;
    }
  } else {
    X::par(nv, [&](int k) { acc[k] = a[k]; });
  }
  Real* vel = LD(vel);
  X::par(nv, [&](int k) {
    Real v = LD(qvel)[k] + h * acc[k];
    vel[k] = v;
    LD(qacc)[k] = (v - LD(qvel)[k]) / h;
  });
  integrate_pos<X>(M, E, vel, h);
  X::par(nv, [&](int k) { LD(qvel)[k] = vel[k]; });
  return true;
}

// Returns false when the env cannot be stepped (status set).
template <class X, class MT>
PHD bool step_env(const MT& M, const DState& S, Env& E) {
  const int nr = step_front<X>(M, S, E);
  if (nr < 0) return nr == STEP_DONE;
  return step_back<X>(M, S, E, nr, nullptr);
}

// ---------------- set_const, forward.rs:275-333 ----------------
// Inverse inertia at the reset configuration (the caller loads qpos_reset
// into LDS qpos): dof_invweight (nv) and body_invweight (nbody x 2). Runs the
// solves on one lane; it is called once per parameter variant, not per step.
template <class X, class MT>
PHD void set_const_env(const MT& M, Env& E, Real* dof_iw, Real* body_iw) {
  const int nv = M.nv, nb = M.nbody;
  kinematics<X>(M, E);
  crba<X>(M, E);
  Real* L = LD(chol);
  const Real* qm = LD(qm);
  X::par(nv, [&](int k) { dof_iw[k] = 0.0; });
  X::par(nb, [&](int b) {
    body_iw[2 * b] = 0.0;
    body_iw[2 * b + 1] = 0.0;
  });
  if (nv == 0) return;
  X::par(nv * nv, [&](int k) { L[k] = qm[k]; });
  int* isc = E.si + M.L.i_iscal;
  X::par(1, [&](int) { isc[1] = 0; });
  if (!chol_block<X>(M, E, L, nv, 0, nv)) return;
  Real* row = LD(tmp);
  Real* x = LD(tmp2);
  X::par(1, [&](int) {
    auto quad = [&]() {
      for (int i = 0; i < nv; i++) x[i] = row[i];
      for (int i = 0; i < nv; i++) x[i] = (x[i] - sdot_n(L + i * nv, x, i)) / L[i * nv + i];
      for (int i = nv - 1; i >= 0; i--) {
        x[i] /= L[i * nv + i];
        Real xi = x[i];
        for (int k = 0; k < i; k++) x[k] += -xi * L[i * nv + k];
      }
      Real s = rsum0();
      for (int i = 0; i < nv; i++) s += row[i] * x[i];
      return s;
    };
    for (int k = 0; k < nv; k++) {
      for (int i = 0; i < nv; i++) row[i] = 0.0;
      row[k] = 1.0;
      dof_iw[k] = quad();
    }
    for (int j = 0; j < M.njnt; j++) {
      int da = M.jnt_dofadr[j], t = M.jnt_type[j];
      int ng = t == J_FREE ? 2 : (t == J_BALL ? 1 : 0);
      for (int gi = 0; gi < ng; gi++) {
        int g = gi * 3;
        Real s = rsum0();
        for (int i = 0; i < 3; i++) s += dof_iw[da + g + i];
        Real avg = s / 3.0;
        for (int i = 0; i < 3; i++) dof_iw[da + g + i] = avg;
      }
    }
    const Real* cd = LD(cdof);
    for (int b = 1; b < nb; b++) {
      Real w0 = 0.0, w1 = 0.0;
      V3 xi = ld3(LD(xipos) + 3 * b);
      for (int axis = 0; axis < 3; axis++) {
        V3 dir = v3(axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0, axis == 2 ? 1.0 : 0.0);
        for (int kind = 0; kind < 2; kind++) {
          for (int i = 0; i < nv; i++) row[i] = 0.0;
          int k = M.body_lastdof[b];
          while (k >= 0) {
            SP c = ldsp(cd + 6 * k);
            row[k] = kind == 0 ? dot(dir, point_velocity(c, xi)) : dot(dir, c.a);
            k = M.dof_parent[k];
          }
          Real qd = quad() / 3.0;
          if (kind == 0) w0 += qd;
          else w1 += qd;
        }
      }
      body_iw[2 * b] = w0;
      body_iw[2 * b + 1] = w1;
    }
  });
}

// ---------------- per-env driver (shared by the kernels and the serial policy) ----------------
enum { IS_NCON = 0, IS_STATUS = 7, IS_ITERS = 8, IS_NROWS = 9 };

template <class X, class MT>
PHD void load_state(const MT& M, const DState& S, Env& E) {
  Real* s = E.s;
  const int64_t e = E.e;
  X::par(M.nq, [&](int k) { s[M.L.qpos + k] = S.qpos[e * M.nq + k]; });
  X::par(M.nv, [&](int k) {
    s[M.L.qvel + k] = S.qvel[e * M.nv + k];
    s[M.L.qacc + k] = S.qacc[e * M.nv + k];
    s[M.L.applied + k] = S.applied[e * M.nv + k];
  });
  X::par(M.nu, [&](int k) { s[M.L.ctrl + k] = S.ctrl[e * M.nu + k]; });
  int* isc = E.si + M.L.i_iscal;
  X::par(1, [&](int) {
    for (int k = 0; k < 16; k++) isc[k] = 0;
    isc[IS_NCON] = S.ncon[e];
    isc[IS_STATUS] = S.status[e];
    isc[IS_ITERS] = S.iters[e];
    isc[IS_NROWS] = S.nrows[e];
  });
}

template <class X, class MT>
PHD void store_poses(const MT& M, const DState& S, Env& E) {
  geom_poses<X>(M, E);
  const Real* s = E.s;
  const int64_t e = E.e;
  X::par(3 * M.ngeom, [&](int k) { S.geom_xpos[e * 3 * M.ngeom + k] = s[M.L.gxpos + k]; });
  X::par(9 * M.ngeom, [&](int k) { S.geom_xmat[e * 9 * M.ngeom + k] = s[M.L.gxmat + k]; });
  X::par(3 * M.nbody, [&](int k) { S.xpos[e * 3 * M.nbody + k] = s[M.L.xpos + k]; });
  X::par(9 * M.nbody, [&](int k) { S.xmat[e * 9 * M.nbody + k] = s[M.L.xmat + k]; });
}

// Load one env, run nstep steps and store it back.
template <class X, class MT>
PHD void run_env(const MT& M, const DState& S, Env& E, int64_t nstep, const Real* ctrl_seq, Real* qpos_out,
                 int64_t rel) {
  Real* lds = E.s;
  load_state<X>(M, S, E);
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  int cost = 0;
  for (int64_t t = 0; t < nstep; t++) {
    if (isc[IS_STATUS] != 0) break;  // a failed env stays put (the reference panics)
    if (ctrl_seq) X::par(M.nu, [&](int k) { lds[M.L.ctrl + k] = ctrl_seq[(rel * nstep + t) * M.nu + k]; });
    if (!step_env<X>(M, S, E)) break;
    cost += isc[IS_ITERS];
    if (qpos_out) X::par(M.nq, [&](int k) { qpos_out[(rel * nstep + t) * M.nq + k] = lds[M.L.qpos + k]; });
  }
  // poses are not refreshed here: they are derived from qpos on demand (k_kin)
  X::par(M.nq, [&](int k) { S.qpos[e * M.nq + k] = lds[M.L.qpos + k]; });
  X::par(M.nv, [&](int k) {
    S.qvel[e * M.nv + k] = lds[M.L.qvel + k];
    S.qacc[e * M.nv + k] = lds[M.L.qacc + k];
  });
  X::par(M.nu, [&](int k) { S.ctrl[e * M.nu + k] = lds[M.L.ctrl + k]; });
  X::par(1, [&](int) {
    S.ncon[e] = isc[IS_NCON];
    S.status[e] = isc[IS_STATUS];
    S.iters[e] = isc[IS_ITERS];
    S.nrows[e] = isc[IS_NROWS];
    S.cost[e] = cost;
    S.time[e] += (Real)nstep * M.timestep;
  });
}

// The PGS step split around its Gauss-Seidel sweeps, one physics step (step t
// of nstep) per launch of each part: run_pgs_front runs the step up to the
// sweeps and leaves in HBM the velocity (vel_g, the env's nv entries), vpos
// when there are no rows, and the contact and row counts and status;
// k_pgs_sweeps runs the sweeps from there; run_pgs_back forms the constraint
// force and the history, integrates and stores the state, as run_env does for
// the step. Everything a part reads from another is in HBM or recomputed, so
// the three give run_env's results bit for bit.
template <class X, class MT>
PHD void run_pgs_front(const MT& M, const DState& S, Env& E, int64_t nstep, int64_t t, const Real* ctrl_seq,
                       int64_t rel, Real* vel_g, Real* vpos_g) {
  Real* lds = E.s;
  load_state<X>(M, S, E);
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  if (isc[IS_STATUS] != 0) return;  // a failed env stays put (the reference panics)
  if (ctrl_seq) {
    X::par(M.nu, [&](int k) { lds[M.L.ctrl + k] = ctrl_seq[(rel * nstep + t) * M.nu + k]; });
    X::par(M.nu, [&](int k) { S.ctrl[e * M.nu + k] = lds[M.L.ctrl + k]; });
  }
  const int nr = step_front<X>(M, S, E, PGS_BEFORE_SWEEPS);
  X::par(M.nv, [&](int k) {
    vel_g[k] = LD(vel)[k];
    vpos_g[k] = LD(acc)[k];
  });
  X::par(1, [&](int) {
    S.ncon[e] = isc[IS_NCON];
    S.status[e] = isc[IS_STATUS];
    S.nrows[e] = nr >= 0 ? nr : isc[IS_NROWS];
  });
}

template <class X, class MT>
PHD void run_pgs_back(const MT& M, const DState& S, Env& E, int64_t nstep, int64_t t, Real* qpos_out, int64_t rel,
                      const Real* vel_g, const Real* vpos_g) {
  Real* lds = E.s;
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  const Real h = M.timestep;
  const int status = S.status[e];
  if (status == 0) {
    Real* vel = LD(vel);
    Real* vpos = LD(acc);
    X::par(M.nq, [&](int k) { lds[M.L.qpos + k] = S.qpos[e * M.nq + k]; });
    X::par(M.nv, [&](int k) {
      lds[M.L.qvel + k] = S.qvel[e * M.nv + k];
      vel[k] = vel_g[k];
      vpos[k] = vpos_g[k];
    });
    X::par(1, [&](int) { isc[IS_NCON] = S.ncon[e]; });
    const int nr = S.nrows[e];
    if (nr > 0) pgs_finish<X>(M, S, E, nr);
    X::par(M.nv, [&](int k) { LD(qacc)[k] = (vel[k] - LD(qvel)[k]) / h; });
    integrate_pos<X>(M, E, vpos, h);
    X::par(M.nv, [&](int k) { LD(qvel)[k] = vel[k]; });
    if (qpos_out) X::par(M.nq, [&](int k) { qpos_out[(rel * nstep + t) * M.nq + k] = lds[M.L.qpos + k]; });
    X::par(M.nq, [&](int k) { S.qpos[e * M.nq + k] = lds[M.L.qpos + k]; });
    X::par(M.nv, [&](int k) {
      S.qvel[e * M.nv + k] = LD(qvel)[k];
      S.qacc[e * M.nv + k] = LD(qacc)[k];
    });
  }
  // run_env's per-launch updates (the PGS step leaves the iteration count alone, so its cost is 0)
  if (t == nstep - 1)
    X::par(1, [&](int) {
      S.cost[e] = 0;
      S.time[e] += (Real)nstep * M.timestep;
    });
}

// The sweeps of one env for the split PGS step when it has more rows than the
// register sweeps hold, on one lane, in the reference's order (solver.rs:149-245).
PHD void pgs_sweeps_serial(const PgsConsts& C, const Real* J, const Real* mj, Real* rw, int nr, int nv, Real* v,
                           Real* vpos) {
  auto sweep = [&](bool use_bias) {
    Real maxdv = 0.0;
    for (int r = 0; r < nr; r++) {
      Real* q = rw + r * 10;
      const Real inv_a = q[PR_INVA];
      if (inv_a == 0.0) continue;
      Real jv = rsum0();
      for (int k = 0; k < nv; k++) jv += J[r * nv + k] * v[k];
      const Real lam = q[PR_LAM];
      const int kind = (int)q[PR_KIND];
      const Real lam_n = kind == PK_FRICTION ? rw[(int)q[PR_NORMAL] * 10 + PR_LAM] : 0.0;
      const Real nw = pgs_row_force(C, kind, inv_a, q[PR_SEP], q[PR_MU], lam, lam_n, jv, use_bias);
      const Real dl = nw - lam;
      if (dl != 0.0) {
        maxdv = rmax(maxdv, fabs(dl) / inv_a);
        for (int k = 0; k < nv; k++) v[k] += mj[r * nv + k] * dl;
        q[PR_LAM] = nw;
      }
    }
    return maxdv;
  };
  for (int it = 0; it < C.iterations; it++)
    if (sweep(true) < C.tol) break;
  for (int k = 0; k < nv; k++) vpos[k] = v[k];
  const int relax = C.iterations / 3 > 2 ? C.iterations / 3 : 2;
  for (int it = 0; it < relax; it++)
    if (sweep(false) < C.tol) break;
}

// ---------------- learned warm start (no reference counterpart) ----------------
// Runs between the rows and Newton, on every env of a wavefront at once (the
// network runs on the matrix cores over the wavefront's contacts): features of
// the env's first MLP_CL touching contacts, the network (X::mlp_tiles), then
// the candidate a0 + M^-1 J^T lambda = M^-1 (smooth + J^T lambda) in LD(vel),
// which step_back reads before it reuses that array. code is the row count of
// an env stepping now, STEP_FAILED / STEP_DONE for an env that is not, or
// MLP_ABSENT for a lane group without an env (tail of the last workgroup).
enum { MLP_ABSENT = -3 };

template <class X, class MT>
PHD const Real* mlp_stage(const MT& M, const DState& S, Env& E, int code) {
  const int nv = M.nv;
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  int cnt = 0;
  if (code > 0) {
    const int na = isc[2];
    cnt = na < MLP_CL ? na : MLP_CL;
    const int nrows = code;
    const Real* J = E.J;
    const Real* aref = E.aref;
    const Real* dd = E.dd;
    const Real* a0 = LD(a);
    const Real* warm = LD(qacc);
    const Real* qv = LD(qvel);
    const int* cact = reinterpret_cast<const int*>(LD(cact));
    auto con = S.con + e * (int64_t)M.ncmax * CW;
    auto cg = S.cgeom + e * (int64_t)M.ncmax * 3;
    X::par(cnt, [&](int k) {
      auto f = S.mlp_feat + (e * MLP_CL + k) * MLP_NF;
      for (int j = 0; j < 4; j++) {
        const int r = 4 * k + j;
        const Real* jr = J + r * nv;
        f[j] = (float)mlp_slog(sdot_n(jr, a0, nv) - aref[r]);      // residual at a0
        f[4 + j] = (float)mlp_slog(sdot_n(jr, warm, nv) - aref[r]);  // at the previous acceleration
        f[8 + j] = (float)mlp_slog(aref[r]);
        f[12 + j] = (float)mlp_slog(sdot_n(jr, qv, nv));             // row velocity
      }
      const int c = cact[k];
      auto cc = con + c * CW;
      const int g1 = cg[3 * c], g2 = cg[3 * c + 1];
      const Real mu = rmax(M.v_friction[(int64_t)E.var * M.ngeom + g1], M.v_friction[(int64_t)E.var * M.ngeom + g2]);
      f[16] = (float)(0.1 * log(dd[4 * k]));
      f[17] = (float)mlp_slog(1000.0 * cc[9]);  // depth in mm
      f[18] = (float)mu;
      f[19] = (float)cc[3];
      f[20] = (float)cc[4];
      f[21] = (float)cc[5];
      f[22] = (float)(na / 16.0);
      f[23] = (float)((nrows - 4 * na) / 16.0);  // limit and equality rows
    });
  }
  X::par(1, [&](int) {
    if (code != MLP_ABSENT) S.mlp_cnt[e] = cnt;
    isc[13] = cnt;
    isc[14] = 0;
  });
  X::mlp_tiles(M, S, e, isc[13]);
  if (code <= 0 || cnt == 0 || !(M.mlp.mode & MLP_WARM)) return nullptr;
  const Real* J = E.J;
  const Real* dd = E.dd;
  Real* w = LD(vel);
  auto pred = S.mlp_pred + e * MLP_CL * MLP_NO;
  X::par(nv, [&](int k) {
    Real f = LD(smooth)[k];
    for (int c = 0; c < cnt; c++)
      for (int j = 0; j < MLP_NO; j++) {
        const int r = 4 * c + j;
        const Real z = (Real)pred[c * MLP_NO + j];
        const Real lam = dd[r] * expm1(z > 0.0 ? z : 0.0);
        f += lam * J[r * nv + k];
      }
    w[k] = f;
  });
  mass_solve<X>(M, E, w);
  X::par(1, [&](int) {
    bool ok = true;
    for (int k = 0; k < nv; k++) ok = ok && rfinite(w[k]);
    isc[14] = ok ? 1 : 0;
  });
  return isc[14] ? w : nullptr;
}

// Training targets after the solve: log1p(lambda_r / D_r) = log1p(max(0,
// aref_r - J_r a)) for the rows of the recorded contacts.
template <class X, class MT>
PHD void mlp_targets(const MT& M, const DState& S, Env& E) {
  const int nv = M.nv;
  const int* isc = E.si + M.L.i_iscal;
  const int cnt = isc[13];
  const Real* J = E.J;
  const Real* aref = E.aref;
  const Real* a = LD(a);
  auto tg = S.mlp_tgt + E.e * MLP_CL * MLP_NO;
  X::par(cnt * MLP_NO, [&](int r) {
    const Real v = aref[r] - sdot_n(J + r * nv, a, nv);
    tg[r] = (float)log1p(v > 0.0 ? v : 0.0);
  });
}

// The learned warm start for PGS (no reference counterpart; the reference
// warm-starts each contact from the nearest previous contact of the same geom
// pair, solver.rs:253-289, and leaves the others at zero). Between the rows and
// the warm start's velocity update, on every env of a wavefront at once:
// features of the env's first MLP_CL contacts, the network, and with
// MLP_WARM, for each of those contacts that the matching left at zero (normal
// and both friction forces), the predicted forces. Outputs 0..2 are
// sign(x) log1p(|x|) of 100 lambda_r / inv_a_r, the velocity change (cm/s) each
// row's force makes along the row, for the normal and the two friction rows;
// the normal force is clamped at 0 and the friction forces within the rows'
// bounds (solver.rs:190-196). code as for mlp_stage (here the row count of a
// PGS env).
template <class X, class MT>
PHD void mlp_stage_pgs(const MT& M, const DState& S, Env& E, int code) {
  const int nv = M.nv;
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  Real* rw = (S.grow + e * (int64_t)M.nrmax * 10).p;
  int cnt = 0;
  if (code > 0) {
    const int ncon = isc[0];
    cnt = ncon < MLP_CL ? ncon : MLP_CL;
    const Real* J = (S.gjac + e * (int64_t)M.nrmax * nv).p;
    const Real* v = LD(vel);
    auto con = S.con + e * (int64_t)M.ncmax * CW;
    X::par(cnt, [&](int c) {
      auto f = S.mlp_feat + (e * MLP_CL + c) * MLP_NF;
      bool matched = false;
      for (int j = 0; j < 3; j++) {
        const int r = 3 * c + j;
        const Real* q = rw + r * 10;
        f[j] = (float)mlp_slog(sdot_n(J + r * nv, v, nv));  // row velocity before the warm start
        f[3 + j] = (float)(0.1 * log(q[PR_INVA] > 0.0 ? q[PR_INVA] : 1e-30));
        f[6 + j] = (float)mlp_slog(q[PR_LAM]);              // the reference's warm start
        matched = matched || q[PR_LAM] != 0.0;
      }
      auto cc = con + c * CW;
      f[9] = matched ? 1.0f : 0.0f;
      f[10] = (float)mlp_slog(1000.0 * cc[9]);  // depth in mm
      f[11] = (float)rw[3 * c * 10 + PR_MU];
      f[12] = (float)cc[3];
      f[13] = (float)cc[4];
      f[14] = (float)cc[5];
      f[15] = (float)(ncon / 16.0);
      f[16] = (float)((code - 3 * ncon) / 16.0);  // limit and equality rows
      for (int k = 17; k < MLP_NF; k++) f[k] = 0.0f;
    });
  }
  X::par(1, [&](int) {
    if (code != MLP_ABSENT) S.mlp_cnt[e] = cnt;
    isc[13] = cnt;
  });
  X::mlp_tiles(M, S, e, isc[13]);
  if (code <= 0 || cnt == 0 || !(M.mlp.mode & MLP_WARM)) return;
  auto pred = S.mlp_pred + e * MLP_CL * MLP_NO;
  X::par(cnt, [&](int c) {
    Real* qn = rw + 3 * c * 10;
    if (qn[PR_LAM] != 0.0 || qn[10 + PR_LAM] != 0.0 || qn[20 + PR_LAM] != 0.0) return;  // matched
    auto inv = [&](int j) {  // lambda from 100 lambda / inv_a
      const Real z = (Real)pred[c * MLP_NO + j];
      const Real ia = qn[10 * j + PR_INVA];
      return ia > 0.0 ? (z < 0.0 ? -expm1(-z) : expm1(z)) * 0.01 * ia : 0.0;
    };
    Real ln = inv(0);
    if (!(ln > 0.0)) ln = 0.0;  // also drops a NaN
    const Real bound = qn[PR_MU] * ln;
    qn[PR_LAM] = ln;
    for (int j = 1; j < 3; j++) {
      const Real t = inv(j);
      qn[10 * j + PR_LAM] = rfinite(t) ? rclamp(t, -bound, bound) : 0.0;
    }
  });
}

// Training targets after the PGS solve: sign(x) log1p(|x|) of 100 lambda_r /
// inv_a_r for the normal and friction rows of the recorded contacts (output 3
// unused; 0 for a row without effective mass).
template <class X, class MT>
PHD void mlp_targets_pgs(const MT& M, const DState& S, Env& E) {
  const int* isc = E.si + M.L.i_iscal;
  const int cnt = isc[13];
  const Real* rw = (S.grow + E.e * (int64_t)M.nrmax * 10).p;
  auto tg = S.mlp_tgt + E.e * MLP_CL * MLP_NO;
  X::par(cnt * MLP_NO, [&](int k) {
    const int c = k / MLP_NO, j = k % MLP_NO;
    const Real* q = rw + (3 * c + (j < 3 ? j : 0)) * 10;
    tg[k] = j < 3 && q[PR_INVA] > 0.0 ? (float)mlp_slog(100.0 * q[PR_LAM] / q[PR_INVA]) : 0.0f;
  });
}

// run_env with the learned warm start. Every lane group of the wavefront
// reaches mlp_stage once per step, whether its env steps, has failed or does
// not exist, because the network runs on the whole wavefront.
template <class X, class MT>
PHD void run_env_mlp(const MT& M, const DState& S, Env& E, int64_t nstep, const Real* ctrl_seq, Real* qpos_out,
                     int64_t rel, bool live) {
  Real* lds = E.s;
  int* isc = E.si + M.L.i_iscal;
  const int64_t e = E.e;
  if (live) load_state<X>(M, S, E);
  for (int64_t t = 0; t < nstep; t++) {
    const bool go = live && isc[IS_STATUS] == 0;
    int code = live ? (int)STEP_FAILED : (int)MLP_ABSENT;
    if (go) {
      if (ctrl_seq) X::par(M.nu, [&](int k) { lds[M.L.ctrl + k] = ctrl_seq[(rel * nstep + t) * M.nu + k]; });
      code = step_front<X>(M, S, E, PGS_BEFORE_WARM);
    }
    if (M.solver_kind == 1) {
      // PGS: the network sets the warm start of unmatched contacts; the step's
      // sweep count goes to the iteration counter
      mlp_stage_pgs<X>(M, S, E, code);
      if (code >= 0) {
        Real* vel = LD(vel);
        Real* vpos = LD(acc);
        int sweeps = 0;
        if (code > 0) {
          pgs_apply_warm<X>(M, S, E, code, vel);
          sweeps = pgs_sweeps<X>(M, S, E, code, vel, vpos);
          pgs_finish<X>(M, S, E, code);
          if (M.mlp.mode & MLP_LOG) mlp_targets_pgs<X>(M, S, E);
        }
        pgs_step_end<X>(M, E);
        X::par(1, [&](int) { isc[IS_ITERS] = sweeps; });
        if (qpos_out) X::par(M.nq, [&](int k) { qpos_out[(rel * nstep + t) * M.nq + k] = lds[M.L.qpos + k]; });
      }
      continue;
    }
    const Real* w2 = mlp_stage<X>(M, S, E, code);
    bool ok = go && code == STEP_DONE;
    if (code >= 0) {
      ok = step_back<X>(M, S, E, code, w2);
      if (ok && (M.mlp.mode & MLP_LOG)) mlp_targets<X>(M, S, E);
      if (!ok) X::par(1, [&](int) { S.mlp_cnt[e] = 0; });
    }
    if (ok && qpos_out) X::par(M.nq, [&](int k) { qpos_out[(rel * nstep + t) * M.nq + k] = lds[M.L.qpos + k]; });
  }
  if (!live) return;
  X::par(M.nq, [&](int k) { S.qpos[e * M.nq + k] = lds[M.L.qpos + k]; });
  X::par(M.nv, [&](int k) {
    S.qvel[e * M.nv + k] = lds[M.L.qvel + k];
    S.qacc[e * M.nv + k] = lds[M.L.qacc + k];
  });
  X::par(M.nu, [&](int k) { S.ctrl[e * M.nu + k] = lds[M.L.ctrl + k]; });
  X::par(1, [&](int) {
    S.ncon[e] = isc[IS_NCON];
    S.status[e] = isc[IS_STATUS];
    S.iters[e] = isc[IS_ITERS];
    S.nrows[e] = isc[IS_NROWS];
    S.time[e] += (Real)nstep * M.timestep;
  });
}

#undef LD
}  // namespace phk
