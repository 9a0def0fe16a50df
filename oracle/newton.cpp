// ORACLE — test infrastructure only (see math.hpp).
// Constraint solvers: the convex primal Newton solver of crates/pippin/src/newton.rs
// and the soft PGS of crates/pippin/src/solver.rs.
#include <algorithm>

#include "engine.hpp"

namespace orc {

typedef unsigned __int128 u128;

// newton.rs:52-66
Real impedance(const Real solimp[5], Real pos) {
  Real dmin = solimp[0], dmax = solimp[1], width = solimp[2], mid = solimp[3], power = solimp[4];
  Real x = rmin(r_fabs(pos) / rmax(width, 1e-12), 1.0);
  if (x >= 1.0 || dmin == dmax) return dmax;
  auto pw = [](Real b, Real e) { return e == 2.0 ? b * b : (e == 1.0 ? b : r_pow(b, e)); };
  Real y = x <= mid ? pw(x, power) / pw(mid, power - 1.0) : 1.0 - pw(1.0 - x, power) / pw(1.0 - mid, power - 1.0);
  return rclamp(dmin + y * (dmax - dmin), 1e-4, 0.9999);
}

// solver.rs:53-59
void point_jac(const Model& m, const std::vector<Spatial>& cdof, size_t body, Vec3 p, Vec3 dir, Real sign, Real* row) {
  size_t k = m.body_lastdof[body];
  while (k != NONE) {
    row[k] += sign * dot(dir, point_velocity(cdof[k], p));
    k = m.dof_parent[k];
  }
}

// newton.rs:68-110
struct Workspace {
  size_t nv, nr;
  const Real* jac;
  const NRow* rows;
  const Real* qm;
  const Real* damp_h;
  void mul_m(const Real* x, Real* y) const {
    for (size_t i = 0; i < nv; i++) y[i] = sdot_n(qm + i * nv, x, nv) + damp_h[i] * x[i];
  }
  const Real* jrow(size_t r) const { return jac + r * nv; }
  Real dot_j(size_t r, const Real* x) const { return sdot_n(jrow(r), x, nv); }
  Real cost(const Real* a, const Real* a0, Real* jar, Real* tmp, Real* tmp2) const {
    for (size_t k = 0; k < nv; k++) tmp[k] = a[k] - a0[k];
    mul_m(tmp, tmp2);
    Real s = rsum_start();
    for (size_t k = 0; k < nv; k++) s += tmp[k] * tmp2[k];
    Real c = 0.5 * s;
    for (size_t r = 0; r < nr; r++) {
      jar[r] = dot_j(r, a) - rows[r].aref;
      if (rows[r].active(jar[r])) c += 0.5 * rows[r].d * jar[r] * jar[r];
    }
    return c;
  }
};

// newton.rs:113-218. Returns the row count; fills jac (nr x nv) and rows.
static size_t build_rows(const Model& m, Data& d, std::vector<Real>& jac, std::vector<NRow>& rows) {
  size_t nv = m.nv;
  Real h = m.timestep;
  const SolverOptions& opt = m.solver;
  jac.clear();
  rows.clear();
  std::vector<Real> pos, diag;
  std::vector<Real> jn(nv), jt1(nv), jt2(nv);
  for (const Contact& c : d.contacts) {
    if (c.depth <= 0.0) continue;
    size_t g1 = c.geom[0], g2 = c.geom[1];
    size_t b1 = m.geom_body[g1], b2 = m.geom_body[g2];
    Real mu = rmax(m.geom_friction[g1], m.geom_friction[g2]);
    Vec3 t1, t2;
    contact_frame(c, &t1, &t2);
    std::vector<Real>* rr[3] = {&jn, &jt1, &jt2};
    Vec3 dirs[3] = {c.normal, t1, t2};
    for (int i = 0; i < 3; i++) {
      std::fill(rr[i]->begin(), rr[i]->end(), 0.0);
      point_jac(m, d.cdof, b2, c.pos, dirs[i], 1.0, rr[i]->data());
      point_jac(m, d.cdof, b1, c.pos, dirs[i], -1.0, rr[i]->data());
    }
    Real tran = m.body_invweight[b1].first + m.body_invweight[b2].first;
    Real dpyr = 2.0 * mu * mu * (1.0 + mu * mu) * tran / opt.impratio;
    const std::vector<Real>* ts[2] = {&jt1, &jt2};
    for (int ti = 0; ti < 2; ti++) {
      const Real ss[2] = {1.0, -1.0};
      for (Real s : ss) {
        for (size_t k = 0; k < nv; k++) jac.push_back(jn[k] + s * mu * (*ts[ti])[k]);
        pos.push_back(-c.depth);
        diag.push_back(dpyr);
      }
    }
  }
  for (size_t j = 0; j < m.njnt(); j++) {
    if (!m.jnt_limited[j] || !(m.jnt_type[j] == Hinge || m.jnt_type[j] == Slide)) continue;
    Real q = d.qpos[m.jnt_qposadr[j]];
    size_t da = m.jnt_dofadr[j];
    Real lo = m.jnt_range_lo[j], hi = m.jnt_range_hi[j];
    Real seps[2] = {q - lo, hi - q}, signs[2] = {1.0, -1.0};
    for (int s = 0; s < 2; s++) {
      if (seps[s] < 0.0) {
        size_t start = jac.size();
        jac.resize(start + nv, 0.0);
        jac[start + da] = signs[s];
        pos.push_back(seps[s]);
        diag.push_back(m.dof_invweight[da]);
      }
    }
  }
  size_t n_ineq = pos.size();
  for (size_t e = 0; e < m.eq_joint1.size(); e++) {
    size_t j1 = m.eq_joint1[e], j2 = m.eq_joint2[e];
    const std::vector<Real>& c = m.eq_polycoef[e];
    Real q1 = d.qpos[m.jnt_qposadr[j1]] - m.qpos0[m.jnt_qposadr[j1]];
    Real poly = c[0], dpoly = 0.0;
    if (j2 != NONE) {
      Real x = d.qpos[m.jnt_qposadr[j2]] - m.qpos0[m.jnt_qposadr[j2]];
      poly = c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4])));
      dpoly = c[1] + x * (2.0 * c[2] + x * (3.0 * c[3] + x * 4.0 * c[4]));
    }
    size_t start = jac.size();
    jac.resize(start + nv, 0.0);
    jac[start + m.jnt_dofadr[j1]] += 1.0;
    Real dg = m.dof_invweight[m.jnt_dofadr[j1]];
    if (j2 != NONE) {
      jac[start + m.jnt_dofadr[j2]] -= dpoly;
      dg += m.dof_invweight[m.jnt_dofadr[j2]];
    }
    pos.push_back(q1 - poly);
    diag.push_back(dg);
  }
  size_t nr = pos.size();
  if (nr == 0) return 0;
  for (size_t r = 0; r < nr; r++) {
    bool bilateral = r >= n_ineq;
    const Real* solref = bilateral ? m.eq_solref[r - n_ineq].data() : opt.solref;
    const Real* solimp = bilateral ? m.eq_solimp[r - n_ineq].data() : opt.solimp;
    Real tc = rmax(solref[0], 2.0 * h), dr = solref[1];
    Real dmax = solimp[1];
    Real b = 2.0 / (dmax * tc);
    Real k = 1.0 / (dmax * dmax * tc * tc * dr * dr);
    Real vel = rsum_start();
    for (size_t i = 0; i < nv; i++) vel += jac[r * nv + i] * d.qvel[i];
    Real imp = impedance(solimp, pos[r]);
    Real reg = rmax((1.0 - imp) / imp * diag[r], 1e-15);
    NRow row;
    row.aref = -b * vel - k * imp * pos[r];
    row.d = 1.0 / reg;
    row.bilateral = bilateral;
    rows.push_back(row);
  }
  return nr;
}

// newton.rs:224-358
static size_t newton_dense(const Workspace& w, Real* a, const Real* warm, size_t max_iters, Real tolerance,
                           size_t* ls_steps) {
  size_t nv = w.nv, nr = w.nr;
  std::vector<Real> buf(6 * nv + 2 * nr + nv * nv, 0.0);
  Real* a0 = buf.data();
  Real* grad = a0 + nv;
  Real* p = grad + nv;
  Real* tmp = p + nv;
  Real* tmp2 = tmp + nv;
  Real* mp = tmp2 + nv;
  Real* jar = mp + nv;
  Real* jp = jar + nr;
  Real* hess = jp + nr;
  for (size_t k = 0; k < nv; k++) a0[k] = a[k];
  if (warm) {
    Real c0 = w.cost(a0, a0, jar, tmp, tmp2);
    Real c1 = w.cost(warm, a0, jar, tmp, tmp2);
    if (c1 < c0)
      for (size_t k = 0; k < nv; k++) a[k] = warm[k];
  }
  w.mul_m(a0, tmp);
  Real s0 = rsum_start();
  for (size_t k = 0; k < nv; k++) s0 += tmp[k] * tmp[k];
  Real scale = 1.0 + r_sqrt(s0);

  size_t iters = 0;
  u128 prev_active = ~(u128)0;
  size_t cap = max_iters > 1 ? max_iters : 1;
  while (iters < cap) {
    iters++;
    u128 active = 0;
    for (size_t k = 0; k < nv; k++) tmp[k] = a[k] - a0[k];
    w.mul_m(tmp, grad);
    for (size_t r = 0; r < nr; r++) {
      jar[r] = w.dot_j(r, a) - w.rows[r].aref;
      if (w.rows[r].active(jar[r])) {
        if (r < 128) active |= (u128)1 << r;
        Real f = w.rows[r].d * jar[r];
        const Real* j = w.jrow(r);
        for (size_t k = 0; k < nv; k++) grad[k] += f * j[k];
      }
    }
    Real gs = rsum_start();
    for (size_t k = 0; k < nv; k++) gs += grad[k] * grad[k];
    Real gnorm = r_sqrt(gs);
    if (gnorm < tolerance * scale || (nr <= 128 && active == prev_active && iters > 1)) break;
    prev_active = active;

    for (size_t i = 0; i < nv; i++) {
      for (size_t j = 0; j < nv; j++) hess[i * nv + j] = w.qm[i * nv + j];
      hess[i * nv + i] += w.damp_h[i];
    }
    for (size_t r = 0; r < nr; r++) {
      if (w.rows[r].active(jar[r])) {
        Real dr = w.rows[r].d;
        const Real* jr = w.jrow(r);
        for (size_t i = 0; i < nv; i++) {
          if (jr[i] == 0.0) continue;
          Real di = dr * jr[i];
          for (size_t j = 0; j < nv; j++) hess[i * nv + j] += di * jr[j];
        }
      }
    }
    if (!cholesky(hess, nv)) break;
    for (size_t k = 0; k < nv; k++) p[k] = -grad[k];
    cholesky_solve(hess, nv, p);

    w.mul_m(p, mp);
    Real mpp = rsum_start(), mpg = rsum_start();
    for (size_t k = 0; k < nv; k++) mpp += p[k] * mp[k];
    for (size_t k = 0; k < nv; k++) mpg += tmp[k] * mp[k];
    for (size_t r = 0; r < nr; r++) jp[r] = w.dot_j(r, p);
    auto deriv = [&](Real al, Real* d1o, Real* d2o) {
      Real d1 = al * mpp + mpg, d2 = mpp;
      for (size_t r = 0; r < nr; r++) {
        Real v = jar[r] + al * jp[r];
        if (w.rows[r].active(v)) {
          Real dq = w.rows[r].d * jp[r];
          d1 += dq * v;
          d2 += dq * jp[r];
        }
      }
      *d1o = d1;
      *d2o = d2;
    };
    Real lo = 0.0, hi = INFINITY, al = 1.0;
    Real d1_0, d2_0;
    deriv(0.0, &d1_0, &d2_0);
    Real tol = 1e-12 * rmax(r_fabs(d1_0), 1e-30);
    for (int it = 0; it < 50; it++) {
      Real d1, d2;
      deriv(al, &d1, &d2);
      if (ls_steps) (*ls_steps)++;
      if (r_fabs(d1) <= tol) break;
      if (d1 < 0.0) lo = al;
      else hi = al;
      Real next = al - d1 / rmax(d2, 1e-30);
      if (next <= lo || next >= hi) next = r_isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * rmax(al, lo);
      if (r_fabs(next - al) <= 1e-14 * rmax(r_fabs(al), 1.0)) {
        al = next;
        break;
      }
      al = next;
    }
    for (size_t k = 0; k < nv; k++) a[k] += al * p[k];
  }
  return iters;
}

// newton.rs:361-371
static void add_constraint_force(const Workspace& w, const Real* a, const size_t* idx, Real* qfrc) {
  for (size_t r = 0; r < w.nr; r++) {
    Real v = w.dot_j(r, a) - w.rows[r].aref;
    if (w.rows[r].active(v)) {
      Real f = -w.rows[r].d * v;
      const Real* j = w.jrow(r);
      for (size_t k = 0; k < w.nv; k++) qfrc[idx ? idx[k] : k] += f * j[k];
    }
  }
}

// newton.rs:378-473
void newton_solve(const Model& m, Data& d, Real* a) {
  size_t nv = m.nv;
  std::vector<Real> jac;
  std::vector<NRow> rows;
  size_t nr = build_rows(m, d, jac, rows);
  d.nrows = nr;
  d.ls_steps = 0;
  std::fill(d.qfrc_constraint.begin(), d.qfrc_constraint.end(), 0.0);
  if (nr == 0) return;
  std::vector<Real> damp_h(nv, 0.0);
  size_t max_iters = m.solver.iterations;
  Real tol = m.solver.tolerance;

  size_t nb = m.dof_blocks.size();
  std::vector<size_t> block_of(nv, 0);
  for (size_t b = 0; b < nb; b++)
    for (size_t k = m.dof_blocks[b].start; k < m.dof_blocks[b].end; k++) block_of[k] = b;
  std::vector<size_t> parent(nb);
  for (size_t b = 0; b < nb; b++) parent[b] = b;
  auto find = [&](size_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  std::vector<size_t> row_block(nr, NONE);
  for (size_t r = 0; r < nr; r++) {
    for (size_t k = 0; k < nv; k++) {
      if (jac[r * nv + k] != 0.0) {
        size_t b = block_of[k];
        if (row_block[r] == NONE) row_block[r] = b;
        else {
          size_t x = find(row_block[r]), y = find(b);
          if (x != y) parent[x > y ? x : y] = x < y ? x : y;
        }
      }
    }
  }
  std::vector<size_t> island_of_block(nb);
  for (size_t b = 0; b < nb; b++) island_of_block[b] = find(b);
  std::vector<size_t> islands;
  for (size_t r = 0; r < nr; r++)
    if (row_block[r] != NONE) islands.push_back(island_of_block[row_block[r]]);
  std::sort(islands.begin(), islands.end());
  islands.erase(std::unique(islands.begin(), islands.end()), islands.end());

  size_t iters = 0;
  bool all_in_one = islands.size() == 1;
  for (size_t b = 0; b < nb && all_in_one; b++)
    if (island_of_block[b] != islands[0]) all_in_one = false;
  if (all_in_one) {
    Workspace w{nv, nr, jac.data(), rows.data(), d.qm.data(), damp_h.data()};
    iters = newton_dense(w, a, d.qacc.data(), max_iters, tol, &d.ls_steps);
    add_constraint_force(w, a, nullptr, d.qfrc_constraint.data());
  } else {
    for (size_t isl : islands) {
      std::vector<size_t> idx, ridx;
      for (size_t b = 0; b < nb; b++)
        if (island_of_block[b] == isl)
          for (size_t k = m.dof_blocks[b].start; k < m.dof_blocks[b].end; k++) idx.push_back(k);
      for (size_t r = 0; r < nr; r++)
        if (row_block[r] != NONE && island_of_block[row_block[r]] == isl) ridx.push_back(r);
      size_t k = idx.size();
      std::vector<Real> qm, sjac, zeros(k, 0.0), sub, warm;
      for (size_t i : idx)
        for (size_t j : idx) qm.push_back(d.qm[i * nv + j]);
      for (size_t r : ridx)
        for (size_t c : idx) sjac.push_back(jac[r * nv + c]);
      std::vector<NRow> srows;
      for (size_t r : ridx) srows.push_back(rows[r]);
      for (size_t i : idx) sub.push_back(a[i]);
      for (size_t i : idx) warm.push_back(d.qacc[i]);
      Workspace w{k, ridx.size(), sjac.data(), srows.data(), qm.data(), zeros.data()};
      size_t it = newton_dense(w, sub.data(), warm.data(), max_iters, tol, &d.ls_steps);
      if (it > iters) iters = it;
      for (size_t t = 0; t < k; t++) a[idx[t]] = sub[t];
      add_constraint_force(w, sub.data(), idx.data(), d.qfrc_constraint.data());
    }
  }
  d.newton_iters = iters;
}

// ---------------- PGS, solver.rs ----------------
static const Real MAX_BIAS_VELOCITY = 1.0;
static const Real PGS_TOLERANCE = 1e-5;

// solver.rs:254-289
static void warm_start(const Data& d, const std::vector<Row>& rows, std::vector<Real>& lambda) {
  const Real MATCH_DIST2 = 0.01 * 0.01;
  const auto& prev = d.prev_contacts;
  if (prev.empty()) return;
  size_t r = 0;
  while (r < rows.size()) {
    size_t ci = rows[r].contact;
    if (ci == NONE) {
      r++;
      continue;
    }
    const Contact& c = d.contacts[ci];
    Real bd = MATCH_DIST2;
    size_t bi = NONE;
    for (size_t pi = 0; pi < prev.size(); pi++) {
      const Contact& p = prev[pi];
      if (p.geom[0] == c.geom[0] && p.geom[1] == c.geom[1]) {
        Real d2 = norm2(p.pos - c.pos);
        if (d2 < bd) {
          bd = d2;
          bi = pi;
        }
      }
    }
    if (bi != NONE) {
      const Contact& p = prev[bi];
      const std::vector<Real>& l = d.prev_lambda[bi];
      Vec3 pt1, pt2, t1, t2;
      contact_frame(p, &pt1, &pt2);
      contact_frame(c, &t1, &t2);
      Vec3 f = pt1 * l[1] + pt2 * l[2];
      lambda[r] = l[0];
      lambda[r + 1] = dot(f, t1);
      lambda[r + 2] = dot(f, t2);
    }
    r += 3;
  }
}

// solver.rs:63-251
void pgs_solve(const Model& m, Data& d, Real* v, std::vector<Real>& v_pos) {
  size_t nv = m.nv;
  Real h = m.timestep;
  const SolverOptions& opt = m.solver;
  std::vector<Row> rows;
  std::vector<Real> jac;
  auto push_row = [&]() {
    jac.resize(jac.size() + nv, 0.0);
    return jac.size() / nv - 1;
  };
  for (size_t ci = 0; ci < d.contacts.size(); ci++) {
    const Contact& c = d.contacts[ci];
    size_t g1 = c.geom[0], g2 = c.geom[1];
    size_t b1 = m.geom_body[g1], b2 = m.geom_body[g2];
    Real mu = rmax(m.geom_friction[g1], m.geom_friction[g2]);
    Vec3 t1, t2;
    contact_frame(c, &t1, &t2);
    size_t base = rows.size();
    Vec3 dirs[3] = {c.normal, t1, t2};
    for (int r = 0; r < 3; r++) {
      size_t i = push_row();
      Real* row = jac.data() + i * nv;
      point_jac(m, d.cdof, b2, c.pos, dirs[r], 1.0, row);
      point_jac(m, d.cdof, b1, c.pos, dirs[r], -1.0, row);
      Row rw;
      rw.kind = r == 0 ? RNormal : RFriction;
      rw.sep = -c.depth;
      rw.normal = base;
      rw.mu = mu;
      rw.inv_a = 0.0;
      rw.contact = ci;
      rows.push_back(rw);
    }
  }
  for (size_t j = 0; j < m.njnt(); j++) {
    if (!m.jnt_limited[j] || !(m.jnt_type[j] == Hinge || m.jnt_type[j] == Slide)) continue;
    Real q = d.qpos[m.jnt_qposadr[j]];
    size_t da = m.jnt_dofadr[j];
    Real lo = m.jnt_range_lo[j], hi = m.jnt_range_hi[j];
    Real reach = opt.contact_margin + 2.0 * h * r_fabs(v[da]);
    Real seps[2] = {q - lo, hi - q}, signs[2] = {1.0, -1.0};
    for (int s = 0; s < 2; s++) {
      if (seps[s] < reach) {
        size_t i = push_row();
        jac[i * nv + da] = signs[s];
        Row rw;
        rw.kind = RNormal;
        rw.sep = seps[s];
        rw.normal = 0;
        rw.mu = 0;
        rw.inv_a = 0.0;
        rw.contact = NONE;
        rows.push_back(rw);
      }
    }
  }
  for (size_t e = 0; e < m.eq_joint1.size(); e++) {
    size_t j1 = m.eq_joint1[e], j2 = m.eq_joint2[e];
    const std::vector<Real>& c = m.eq_polycoef[e];
    Real q1 = d.qpos[m.jnt_qposadr[j1]] - m.qpos0[m.jnt_qposadr[j1]];
    size_t i = push_row();
    jac[i * nv + m.jnt_dofadr[j1]] += 1.0;
    Real poly = c[0];
    if (j2 != NONE) {
      Real x = d.qpos[m.jnt_qposadr[j2]] - m.qpos0[m.jnt_qposadr[j2]];
      poly = c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4])));
      jac[i * nv + m.jnt_dofadr[j2]] -= c[1] + x * (2.0 * c[2] + x * (3.0 * c[3] + x * 4.0 * c[4]));
    }
    Row rw;
    rw.kind = RBilateral;
    rw.sep = q1 - poly;
    rw.normal = 0;
    rw.mu = 0;
    rw.inv_a = 0.0;
    rw.contact = NONE;
    rows.push_back(rw);
  }
  size_t nr = rows.size();
  d.nrows = nr;
  v_pos.assign(v, v + nv);
  if (nr == 0) {
    std::fill(d.qfrc_constraint.begin(), d.qfrc_constraint.end(), 0.0);
    d.prev_contacts.clear();
    d.prev_lambda.clear();
    return;
  }
  std::vector<Real> mj = jac;
  for (size_t r = 0; r < nr; r++) {
    Real* x = mj.data() + r * nv;
    mass_solve(m, d.qm_chol, x);
    Real a = rsum_start();
    for (size_t k = 0; k < nv; k++) a += jac[r * nv + k] * x[k];
    rows[r].inv_a = a > 1e-12 ? 1.0 / a : 0.0;
  }
  std::vector<Real> lambda(nr, 0.0);
  warm_start(d, rows, lambda);
  for (size_t r = 0; r < nr; r++)
    if (lambda[r] != 0.0)
      for (size_t k = 0; k < nv; k++) v[k] += mj[r * nv + k] * lambda[r];

  // solver.rs:44-50
  Real omega = 2.0 * M_PI * opt.contact_hertz;
  Real a1 = 2.0 * opt.contact_damping_ratio + h * omega;
  Real a2 = h * omega * a1;
  Real a3 = 1.0 / (1.0 + a2);
  Real bias_rate = omega / a1, mass_scale = a2 * a3, impulse_scale = a3;

  auto iterate = [&](bool use_bias) {
    Real max_dv = 0.0;
    for (size_t r = 0; r < nr; r++) {
      const Row row = rows[r];
      if (row.inv_a == 0.0) continue;
      const Real* jr = jac.data() + r * nv;
      Real jv = rsum_start();
      for (size_t k = 0; k < nv; k++) jv += jr[k] * v[k];
      Real nw;
      if (row.kind == RNormal) {
        Real s = row.sep + opt.contact_slop;
        Real bias, ms, is;
        if (s > 0.0) {
          bias = s / h;
          ms = 1.0;
          is = 0.0;
        } else if (use_bias) {
          bias = rmax(bias_rate * s, -MAX_BIAS_VELOCITY);
          ms = mass_scale;
          is = impulse_scale;
        } else {
          bias = 0.0;
          ms = 1.0;
          is = 0.0;
        }
        Real delta = -row.inv_a * ms * (jv + bias) - is * lambda[r];
        nw = rmax(lambda[r] + delta, 0.0);
      } else if (row.kind == RBilateral) {
        Real bias, ms, is;
        if (use_bias) {
          bias = rclamp(bias_rate * row.sep, -MAX_BIAS_VELOCITY, MAX_BIAS_VELOCITY);
          ms = mass_scale;
          is = impulse_scale;
        } else {
          bias = 0.0;
          ms = 1.0;
          is = 0.0;
        }
        nw = lambda[r] - row.inv_a * ms * (jv + bias) - is * lambda[r];
      } else {
        Real bound = row.mu * lambda[row.normal];
        nw = rclamp(lambda[r] - row.inv_a * jv, -bound, bound);
      }
      Real dl = nw - lambda[r];
      if (dl != 0.0) {
        max_dv = rmax(max_dv, r_fabs(dl) / row.inv_a);
        const Real* mr = mj.data() + r * nv;
        for (size_t k = 0; k < nv; k++) v[k] += mr[k] * dl;
        lambda[r] = nw;
      }
    }
    return max_dv;
  };
  for (size_t it = 0; it < opt.iterations; it++)
    if (iterate(true) < PGS_TOLERANCE) break;
  v_pos.assign(v, v + nv);
  size_t relax = opt.iterations / 3 > 2 ? opt.iterations / 3 : 2;
  for (size_t it = 0; it < relax; it++)
    if (iterate(false) < PGS_TOLERANCE) break;

  std::fill(d.qfrc_constraint.begin(), d.qfrc_constraint.end(), 0.0);
  for (size_t r = 0; r < nr; r++)
    for (size_t k = 0; k < nv; k++) d.qfrc_constraint[k] += jac[r * nv + k] * lambda[r] / h;
  d.prev_contacts = d.contacts;
  d.prev_lambda.assign(d.contacts.size(), std::vector<Real>(3, 0.0));
  for (size_t r = 0; r < nr; r++) {
    size_t c = rows[r].contact;
    if (c != NONE) {
      size_t slot;
      if (rows[r].kind == RNormal) slot = 0;
      else if (rows[r].kind == RFriction) slot = r - rows[r].normal;
      else continue;
      d.prev_lambda[c][slot] = lambda[r];
    }
  }
}

}  // namespace orc
