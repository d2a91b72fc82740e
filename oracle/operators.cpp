// CPU oracle (test infrastructure only; see oracle.h): the solid (neo-Hookean
// tangent + Newmark mass) and heat (M + dt c K) matrix-free operators, restated
// from the reference's arithmetic.  Sum factorisation here uses shape values
// AND shape derivatives at the Gauss points (not the collocation derivative the
// device kernels use), so the two code paths share no contraction code.
#include <omp.h>

#include "fe.hpp"
#include "oracle.h"

namespace orc {

struct Basis {
  int N = 0, Q = 0;
  vector<double> S, Sd;  // Q x N (row q, column node a)
  vector<double> qx, qw;
  void init(const Mesh &m, int nq) {
    N = m.N; Q = nq;
    qx.resize(Q); qw.resize(Q);
    gauss_rule(Q, qx.data(), qw.data());
    S.resize(Q * N); Sd.resize(Q * N);
    for (int q = 0; q < Q; ++q) lagrange_eval(N, m.gll.data(), qx[q], &S[q * N], &Sd[q * N]);
  }
};

// Contract matrix M (R x C, row-major) along direction dir of the tensor
// `in` (dims[0] fastest).  transpose=false: out[r] = sum_c M[r][c] in[c];
// transpose=true: out[c] = sum_r M[r][c] in[r].
static void apply1d(const double *M, int R, int C, int dir, const int *dims, const double *in,
                    double *out, bool transpose) {
  int od[3] = {dims[0], dims[1], dims[2]};
  const int nin = transpose ? R : C, nout = transpose ? C : R;
  od[dir] = nout;
  const int stride_in = dir == 0 ? 1 : (dir == 1 ? dims[0] : dims[0] * dims[1]);
  const int stride_out = dir == 0 ? 1 : (dir == 1 ? od[0] : od[0] * od[1]);
  for (int k = 0; k < od[2]; ++k)
    for (int j = 0; j < od[1]; ++j)
      for (int i = 0; i < od[0]; ++i) {
        int idx[3] = {i, j, k};
        int o = idx[dir];
        idx[dir] = 0;
        int base_in = idx[0] + dims[0] * (idx[1] + dims[1] * idx[2]);
        double s = 0;
        for (int t = 0; t < nin; ++t) {
          double coef = transpose ? M[t * C + o] : M[o * C + t];
          s += coef * in[base_in + t * stride_in];
        }
        out[i + od[0] * (j + od[1] * k)] = s;
        (void)stride_out;
      }
}

// values and reference gradients of one scalar field at the Gauss points
static void evaluate(const Basis &B, const double *u, double *val, double *g0, double *g1,
                     double *g2) {
  const int N = B.N, Q = B.Q;
  vector<double> a(Q * N * N), ad(Q * N * N), b(Q * Q * N), bd(Q * Q * N), c(Q * Q * N);
  int d0[3] = {N, N, N};
  apply1d(B.S.data(), Q, N, 0, d0, u, a.data(), false);
  apply1d(B.Sd.data(), Q, N, 0, d0, u, ad.data(), false);
  int d1[3] = {Q, N, N};
  apply1d(B.S.data(), Q, N, 1, d1, a.data(), b.data(), false);    // S S
  apply1d(B.Sd.data(), Q, N, 1, d1, a.data(), bd.data(), false);  // S Sd
  apply1d(B.S.data(), Q, N, 1, d1, ad.data(), c.data(), false);   // Sd S
  int d2[3] = {Q, Q, N};
  if (val) apply1d(B.S.data(), Q, N, 2, d2, b.data(), val, false);
  if (g2) apply1d(B.Sd.data(), Q, N, 2, d2, b.data(), g2, false);
  if (g1) apply1d(B.S.data(), Q, N, 2, d2, bd.data(), g1, false);
  if (g0) apply1d(B.S.data(), Q, N, 2, d2, c.data(), g0, false);
}

// transpose of evaluate: out[node] += sum_q (phi val + dphi/dxi_f g_f)
static void integrate(const Basis &B, const double *val, const double *g0, const double *g1,
                      const double *g2, double *out) {
  const int N = B.N, Q = B.Q;
  int d2[3] = {Q, Q, Q};
  vector<double> b(Q * Q * N), bd(Q * Q * N), c(Q * Q * N), tmp(Q * Q * N);
  std::fill(b.begin(), b.end(), 0.0);
  if (val) apply1d(B.S.data(), Q, N, 2, d2, val, b.data(), true);
  if (g2) {
    apply1d(B.Sd.data(), Q, N, 2, d2, g2, tmp.data(), true);
    for (size_t i = 0; i < b.size(); ++i) b[i] += tmp[i];
  }
  apply1d(B.S.data(), Q, N, 2, d2, g1, bd.data(), true);
  apply1d(B.S.data(), Q, N, 2, d2, g0, c.data(), true);
  int d1[3] = {Q, Q, N};
  vector<double> a(Q * N * N), ad(Q * N * N), t2(Q * N * N);
  apply1d(B.S.data(), Q, N, 1, d1, b.data(), a.data(), true);
  apply1d(B.Sd.data(), Q, N, 1, d1, bd.data(), t2.data(), true);
  for (size_t i = 0; i < a.size(); ++i) a[i] += t2[i];
  apply1d(B.S.data(), Q, N, 1, d1, c.data(), ad.data(), true);
  int d0[3] = {Q, N, N};
  vector<double> r(N * N * N), r2(N * N * N);
  apply1d(B.S.data(), Q, N, 0, d0, a.data(), r.data(), true);
  apply1d(B.Sd.data(), Q, N, 0, d0, ad.data(), r2.data(), true);
  for (int i = 0; i < N * N * N; ++i) out[i] += r[i] + r2[i];
}

// ------------------------------------------------------------- geometry ---
struct Geometry {
  // reference-frame (MappingQ1) inverse Jacobian [f][e] = dxi_f/dX_e, JxW
  vector<double> jinv, jxw, jac;  // per (cell, q): 9, 1, 9
  void init(const Mesh &m, const Basis &B) {
    const int Q3 = B.Q * B.Q * B.Q;
    jinv.resize((size_t)m.n_cells * Q3 * 9);
    jac.resize((size_t)m.n_cells * Q3 * 9);
    jxw.resize((size_t)m.n_cells * Q3);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < m.n_cells; ++c)
      for (int q = 0; q < Q3; ++q) {
        int q0 = q % B.Q, q1 = (q / B.Q) % B.Q, q2 = q / (B.Q * B.Q);
        double xi[3] = {B.qx[q0], B.qx[q1], B.qx[q2]};
        double J[3][3], Ji[3][3];
        q1_jacobian(&m.vert[(size_t)c * 24], xi, J);
        inv3(J, Ji);
        size_t o = (size_t)c * Q3 + q;
        for (int i = 0; i < 9; ++i) {
          jinv[o * 9 + i] = Ji[i / 3][i % 3];
          jac[o * 9 + i] = J[i / 3][i % 3];
        }
        jxw[o] = std::abs(det3(J)) * B.qw[q0] * B.qw[q1] * B.qw[q2];
      }
  }
};

// ------------------------------------------------------------- solid op ---
// neo-Hookean constants of one material (material.h:78-94)
struct Material {
  double mu = 0, nu = 0, rho = 0, kappa = 0, c_1 = 0, lambda = 0, rho_alpha = 0;
};

struct Solid {
  Mesh *m = nullptr;
  Basis B;
  Geometry G;
  int Q3 = 0, form = 1, caching = 1;
  double alpha1 = 0;
  // mat[0] on cells of material id 0, mat[1] (the inclusion) on the others
  // (mf_nh_operator.h:404, 658); mat_id empty: every cell has id 0
  Material mat[2];
  vector<int32_t> mat_id;
  const Material &material(int cell) const { return !mat_id.empty() && mat_id[cell] != 0 ? mat[1] : mat[0]; }
  // caches, per (cell, q)
  vector<double> sc, sc2, t2, t4, jinv_cur, jxw_cur;
  vector<double> u_cell;  // scalar_referential: plain cell dof values
  vector<double> u;       // linearisation point (for brute force)
  bool deformed() const { return caching == 1 || caching == 2; }
};

struct Heat {
  Mesh *m = nullptr;
  Basis B;
  Geometry G;
  int Q3 = 0;
  double dt = 0;
  vector<double> coef;  // per (cell, q)
};

static Material material_constants(double mu, double nu, double rho, double alpha1) {
  // material.h:78-94
  Material M;
  M.mu = mu;
  M.nu = nu;
  M.rho = rho;
  M.kappa = (2.0 * mu * (1.0 + nu)) / (3.0 * (1.0 - 2.0 * nu));
  M.c_1 = mu / 2.0;
  M.rho_alpha = rho * alpha1;
  M.lambda = (2.0 * mu * nu) / (1.0 - 2.0 * nu);
  return M;
}

static void gather(const Mesh &m, int cell, const double *src, bool zero_constrained,
                   vector<double> &loc /* nc * N3, comp-major */) {
  const int N3 = m.N * m.N * m.N;
  loc.resize((size_t)m.nc * N3);
  for (int i = 0; i < N3; ++i) {
    int64_t node = m.cell_nodes[(size_t)cell * N3 + i];
    for (int c = 0; c < m.nc; ++c) {
      int64_t dof = node * m.nc + c;
      loc[(size_t)c * N3 + i] = (zero_constrained && m.is_con[dof]) ? 0.0 : src[dof];
    }
  }
}

// ref gradient [comp][dir] at all q of a vector field from cell dofs
static void eval_vector(const Basis &B, int nc, int N3, const double *loc, vector<double> &val,
                        vector<double> &gr /* [c][f][q] */) {
  const int Q3 = B.Q * B.Q * B.Q;
  val.resize((size_t)nc * Q3);
  gr.resize((size_t)nc * 3 * Q3);
  for (int c = 0; c < nc; ++c)
    evaluate(B, loc + (size_t)c * N3, &val[(size_t)c * Q3], &gr[(size_t)(c * 3 + 0) * Q3],
             &gr[(size_t)(c * 3 + 1) * Q3], &gr[(size_t)(c * 3 + 2) * Q3]);
}

// F = I + Grad_X u at one q from reference gradients and Jref^{-1}
static void deformation_gradient(const double *gxi /*[c][f] stride Q3*/, int Q3, int q,
                                 const double *jinv, double F[3][3]) {
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e) {
      double s = 0;
      for (int f = 0; f < 3; ++f) s += gxi[(size_t)(d * 3 + f) * Q3 + q] * jinv[f * 3 + e];
      F[d][e] = s + (d == e ? 1.0 : 0.0);
    }
}

// Kirchhoff stress tau (material.h:116-161)
static void get_tau(const Solid &s, const Material &M, const double F[3][3], double tau[3][3]) {
  const double J = det3(F);
  double b[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = 0;
      for (int k = 0; k < 3; ++k) t += F[i][k] * F[j][k];
      b[i][j] = t;
    }
  if (s.form == 1) {
    const double tmp = M.mu - 2.0 * M.lambda * std::log(J);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) tau[i][j] = M.mu * b[i][j] - (i == j ? tmp : 0.0);
  } else {
    const double Jm23 = std::pow(J, -2.0 / 3.0);
    const double dpsi = (M.kappa / 2.0) * (J - 1.0 / J);
    double tr = 0;
    for (int i = 0; i < 3; ++i) tr += 2.0 * M.c_1 * Jm23 * b[i][i];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        tau[i][j] = 2.0 * M.c_1 * Jm23 * b[i][j] + (i == j ? dpsi * J - tr / 3.0 : 0.0);
  }
}

static int solid_cache(Solid &s, const double *u) {
  const Mesh &m = *s.m;
  const int N3 = m.N * m.N * m.N, Q3 = s.Q3;
  const size_t nq = (size_t)m.n_cells * Q3;
  s.u.assign(u, u + m.n_dofs);
  s.sc.assign(nq, 0.0);
  s.sc2.assign(nq, 0.0);
  if (s.caching == 1 || s.caching == 2) s.t2.assign(nq * 9, 0.0);
  if (s.caching == 2 || s.caching == 3) s.t4.assign(nq * 81, 0.0);
  if (s.caching == 0) s.u_cell.assign((size_t)m.n_cells * N3 * 3, 0.0);
  if (s.deformed()) {
    s.jinv_cur.assign(nq * 9, 0.0);
    s.jxw_cur.assign(nq, 0.0);
  }
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int cell = 0; cell < m.n_cells; ++cell) {
#ifdef ORC_PLANT_INCL_CACHE_MAT0
    const Material &M = s.mat[0];  // planted: the cache ignores the cell's material
#else
    const Material &M = s.material(cell);
#endif
    vector<double> loc, val, gr, xloc, xval, xgr;
    gather(m, cell, u, false, loc);  // read_dof_values_plain
    eval_vector(s.B, 3, N3, loc.data(), val, gr);
    if (s.caching == 0)
      for (int i = 0; i < 3 * N3; ++i) s.u_cell[(size_t)cell * 3 * N3 + i] = loc[i];
    if (s.deformed()) {
      // MappingQEulerian(p): support points x = X(support) + u(support); the
      // Jacobian of the isoparametric map from the GLL-node positions.
      xloc.resize(3 * N3);
      for (int i = 0; i < N3; ++i) {
        int64_t node = m.cell_nodes[(size_t)cell * N3 + i];
        for (int d = 0; d < 3; ++d) xloc[d * N3 + i] = m.xnode[node * 3 + d] + loc[d * N3 + i];
      }
      eval_vector(s.B, 3, N3, xloc.data(), xval, xgr);
    }
    for (int q = 0; q < Q3; ++q) {
      const size_t o = (size_t)cell * Q3 + q;
      double F[3][3];
      deformation_gradient(gr.data(), Q3, q, &s.G.jinv[o * 9], F);
      const double J = det3(F);
      if (!(J > 0)) { bad += 1; continue; }
      if (s.deformed()) {
        double Jc[3][3], Jci[3][3];
        for (int e = 0; e < 3; ++e)
          for (int f = 0; f < 3; ++f) Jc[e][f] = xgr[(size_t)(e * 3 + f) * Q3 + q];
        inv3(Jc, Jci);
        for (int i = 0; i < 9; ++i) s.jinv_cur[o * 9 + i] = Jci[i / 3][i % 3];
        int q0 = q % s.B.Q, q1 = (q / s.B.Q) % s.B.Q, q2 = q / (s.B.Q * s.B.Q);
        s.jxw_cur[o] = std::abs(det3(Jc)) * s.B.qw[q0] * s.B.qw[q1] * s.B.qw[q2];
      }
      if (s.form == 0) {
        // mf_nh_operator.h:419-440
        s.sc[o] = std::pow(J, -1.0 / 3.0);
        for (int i = 0; i < 9; ++i) s.t2[o * 9 + i] = F[i / 3][i % 3];
        continue;
      }
      const double scalar = M.mu - 2.0 * M.lambda * std::log(J);
      double tau[3][3];
      get_tau(s, M, F, tau);
      switch (s.caching) {
        case 0:
          s.sc[o] = scalar;
          break;
        case 1:  // mf_nh_operator.h:462-477
          s.sc[o] = scalar * 2.0 / J;
          s.sc2[o] = 1.0 / J;
          for (int i = 0; i < 9; ++i) s.t2[o * 9 + i] = tau[i / 3][i % 3] / J;
          break;
        case 2:  // mf_nh_operator.h:478-494
          s.sc2[o] = 1.0 / J;
          for (int i = 0; i < 9; ++i) s.t2[o * 9 + i] = tau[i / 3][i % 3] / J;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                  // S_ijkl = (d_ik d_jl + d_il d_jk)/2, IxI = d_ij d_kl
                  double S = 0.5 * ((i == k && j == l) + (i == l && j == k));
                  double IxI = (i == j && k == l) ? 1.0 : 0.0;
                  s.t4[o * 81 + ((i * 3 + j) * 3 + k) * 3 + l] =
                      (scalar * 2.0 / J) * S + (M.lambda * 2.0 / J) * IxI;
                }
          break;
        case 3: {  // mf_nh_operator.h:495-516
          double Fi[3][3];
          inv3(F, Fi);
          const double lnJ = std::log(J);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                  // F^-T (x) F^-T : A_ij B_kl ; iljk: A_il B_jk with A=F^-T, B=F^-1;
                  // ikjl of (I, I): d_ik d_jl
                  double FiT_ij = Fi[j][i], FiT_kl = Fi[l][k], FiT_il = Fi[l][i], Fi_jk = Fi[j][k];
                  s.t4[o * 81 + ((i * 3 + j) * 3 + k) * 3 + l] =
                      (2.0 * M.lambda) * FiT_ij * FiT_kl +
                      (M.mu - 2.0 * M.lambda * lnJ) * FiT_il * Fi_jk +
                      M.mu * ((i == k && j == l) ? 1.0 : 0.0);
                }
          break;
        }
      }
    }
  }
  return bad ? -1 : 0;
}

// The per-cell operator: local in (comp-major nc*N3), local out (+=).
// mf_nh_operator.h:651-1019.
static void solid_cell(const Solid &s, int cell, const double *loc, double *out) {
  const Mesh &m = *s.m;
  const Material &M = s.material(cell);
  const int N3 = m.N * m.N * m.N, Q3 = s.Q3;
  vector<double> val, gr;
  eval_vector(s.B, 3, N3, loc, val, gr);
  vector<double> rv(3 * Q3), rg(9 * Q3);
  vector<double> ugr, uval;
  if (s.caching == 0) {
    // gradient of the cached nodal displacement (phi_grad.evaluate)
    eval_vector(s.B, 3, N3, &s.u_cell[(size_t)cell * 3 * N3], uval, ugr);
  }
  for (int q = 0; q < Q3; ++q) {
    const size_t o = (size_t)cell * Q3 + q;
    const double *jinv = s.deformed() ? &s.jinv_cur[o * 9] : &s.G.jinv[o * 9];
    const double jxw = s.deformed() ? s.jxw_cur[o] : s.G.jxw[o];
    double g[3][3], v[3];
    for (int d = 0; d < 3; ++d) {
      v[d] = val[(size_t)d * Q3 + q];
      for (int e = 0; e < 3; ++e) {
        double t = 0;
        for (int f = 0; f < 3; ++f) t += gr[(size_t)(d * 3 + f) * Q3 + q] * jinv[f * 3 + e];
        g[d][e] = t;
      }
    }
    double P[3][3], vs[3];
    if (s.form == 0) {
      // mf_nh_operator.h:701-840 (tensor2 caching, deformed frame)
      double F[3][3];
      for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = s.t2[o * 9 + i];
      const double det_F = det3(F);
      double Fb[3][3], bb[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Fb[i][j] = F[i][j] * s.sc[o];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0;
          for (int k = 0; k < 3; ++k) t += Fb[i][k] * Fb[j][k];
          bb[i][j] = t;
        }
      double sym[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sym[i][j] = 0.5 * (g[i][j] + g[j][i]);
      double tau_bar[3][3], tr_tau_bar = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau_bar[i][j] = bb[i][j] * (2.0 * M.c_1);
      for (int i = 0; i < 3; ++i) tr_tau_bar += tau_bar[i][i];
      const double tr_tau_bar_dim = tr_tau_bar / 3.0;
      const double dPsi = (M.kappa / 2.0) * (det_F - 1.0 / det_F);
      const double dPsi_J = dPsi * det_F;
      const double d2Psi = (M.kappa / 2.0) * (1.0 + 1.0 / (det_F * det_F));
      double tau[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau[i][j] = tau_bar[i][j] + (i == j ? dPsi_J - tr_tau_bar_dim : 0.0);
      const double tr = sym[0][0] + sym[1][1] + sym[2][2];
      double jc[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) jc[i][j] = sym[i][j] * (-dPsi_J * 2.0);
      const double tmp = det_F * (dPsi + det_F * d2Psi) * tr;
      for (int i = 0; i < 3; ++i) jc[i][i] += tmp;
      double tau_iso[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau_iso[i][j] = tau_bar[i][j] - (i == j ? tr_tau_bar_dim : 0.0);
      double dev[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dev[i][j] = sym[i][j] - (i == j ? tr / 3.0 : 0.0);
      double tau_iso_src = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau_iso_src += tau_iso[i][j] * sym[i][j];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          jc[i][j] += (2.0 / 3.0 * tr_tau_bar) * dev[i][j];
          jc[i][j] -= (2.0 / 3.0 * tr) * tau_iso[i][j];
        }
      for (int i = 0; i < 3; ++i) jc[i][i] -= 2.0 / 3.0 * tau_iso_src;
      const double inv_det_F = 1.0 / det_F;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double geo = 0;
          for (int k = 0; k < 3; ++k) geo += g[i][k] * tau[k][j];
          P[i][j] = (jc[i][j] + geo) * inv_det_F;
        }
      for (int d = 0; d < 3; ++d) vs[d] = v[d] * M.rho_alpha * inv_det_F;
    } else if (s.caching == 1) {
      // mf_nh_operator.h:945-974
      double sym[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sym[i][j] = 0.5 * (g[i][j] + g[j][i]);
      const double tr = sym[0][0] + sym[1][1] + sym[2][2];
      const double tmp = 2.0 * M.lambda * s.sc2[o] * tr;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double geo = 0;
          for (int k = 0; k < 3; ++k) geo += g[i][k] * s.t2[o * 9 + k * 3 + j];
          P[i][j] = s.sc[o] * sym[i][j] + (i == j ? tmp : 0.0) + geo;
        }
      for (int d = 0; d < 3; ++d) vs[d] = v[d] * M.rho_alpha * s.sc2[o];
    } else if (s.caching == 2) {
      // mf_nh_operator.h:976-994
      double sym[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sym[i][j] = 0.5 * (g[i][j] + g[j][i]);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double geo = 0, c4 = 0;
          for (int k = 0; k < 3; ++k) geo += g[i][k] * s.t2[o * 9 + k * 3 + j];
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) c4 += s.t4[o * 81 + ((i * 3 + j) * 3 + k) * 3 + l] * sym[k][l];
          P[i][j] = geo + c4;
        }
      for (int d = 0; d < 3; ++d) vs[d] = v[d] * M.rho_alpha * s.sc2[o];
    } else if (s.caching == 3) {
      // mf_nh_operator.h:996-1010: double_contract<2,0,3,1>(A, g)_ij = A_ijkl g_kl
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0;
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) t += s.t4[o * 81 + ((i * 3 + j) * 3 + k) * 3 + l] * g[k][l];
          P[i][j] = t;
        }
      for (int d = 0; d < 3; ++d) vs[d] = v[d] * M.rho_alpha;
    } else {
      // scalar_referential, mf_nh_operator.h:848-943
      double F[3][3], gx[3][3];
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
          double t = 0;
          for (int f = 0; f < 3; ++f) t += jinv[f * 3 + e] * ugr[(size_t)(d * 3 + f) * Q3 + q];
          F[d][e] = t + (d == e ? 1.0 : 0.0);
        }
      double b[3][3], tau[3][3], Fi[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0;
          for (int k = 0; k < 3; ++k) t += F[i][k] * F[j][k];
          b[i][j] = t;
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tau[i][j] = M.mu * b[i][j] - (i == j ? s.sc[o] : 0.0);
      inv3(F, Fi);
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
          double t = 0;
          for (int f = 0; f < 3; ++f) t += Fi[f][e] * g[d][f];
          gx[d][e] = t;
        }
      double jc[3][3];
      const double trg = gx[0][0] + gx[1][1] + gx[2][2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          jc[i][j] = 2.0 * s.sc[o] * 0.5 * (gx[i][j] + gx[j][i]) + (i == j ? 2.0 * M.lambda * trg : 0.0);
      double queued[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0;
          for (int k = 0; k < 3; ++k) t += gx[i][k] * tau[k][j];
          queued[i][j] = jc[i][j] + t;
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double t = 0;
          for (int k = 0; k < 3; ++k) t += queued[i][k] * Fi[j][k];
          P[i][j] = t;
        }
      for (int d = 0; d < 3; ++d) vs[d] = v[d] * M.rho_alpha;
    }
    // submit: P -> reference gradient  sum_e P[d][e] jinv[f][e] * JxW
    for (int d = 0; d < 3; ++d) {
      rv[(size_t)d * Q3 + q] = vs[d] * jxw;
      for (int f = 0; f < 3; ++f) {
        double t = 0;
        for (int e = 0; e < 3; ++e) t += P[d][e] * jinv[f * 3 + e];
        rg[(size_t)(d * 3 + f) * Q3 + q] = t * jxw;
      }
    }
  }
  for (int d = 0; d < 3; ++d)
    integrate(s.B, &rv[(size_t)d * Q3], &rg[(size_t)(d * 3 + 0) * Q3], &rg[(size_t)(d * 3 + 1) * Q3],
              &rg[(size_t)(d * 3 + 2) * Q3], out + (size_t)d * N3);
}

static void heat_cell(const Heat &h, int cell, const double *loc, double *out) {
  const int Q3 = h.Q3;
  vector<double> val(Q3), g0(Q3), g1(Q3), g2(Q3);
  evaluate(h.B, loc, val.data(), g0.data(), g1.data(), g2.data());
  vector<double> rv(Q3), r0(Q3), r1(Q3), r2(Q3);
  for (int q = 0; q < Q3; ++q) {
    const size_t o = (size_t)cell * Q3 + q;
    const double *ji = &h.G.jinv[o * 9];
    const double jxw = h.G.jxw[o];
    double gx[3];
    for (int e = 0; e < 3; ++e) gx[e] = g0[q] * ji[0 * 3 + e] + g1[q] * ji[1 * 3 + e] + g2[q] * ji[2 * 3 + e];
    // heat_transfer.h:242-248: value v, gradient c*dt*grad v
    const double c = h.coef.empty() ? 1.0 : h.coef[o];
    double sg[3];
    for (int e = 0; e < 3; ++e) sg[e] = c * h.dt * gx[e];
    rv[q] = val[q] * jxw;
    r0[q] = (sg[0] * ji[0] + sg[1] * ji[1] + sg[2] * ji[2]) * jxw;
    r1[q] = (sg[0] * ji[3] + sg[1] * ji[4] + sg[2] * ji[5]) * jxw;
    r2[q] = (sg[0] * ji[6] + sg[1] * ji[7] + sg[2] * ji[8]) * jxw;
  }
  integrate(h.B, rv.data(), r0.data(), r1.data(), r2.data(), out);
}

// cell loop over 8 colours (cell parity) so that the scatter needs no atomics
template <class CellFn>
static void cell_loop(const Mesh &m, const double *src, double *dst, bool constrained_identity,
                      bool zero_constrained, CellFn fn) {
  const int N3 = m.N * m.N * m.N;
  for (int64_t i = 0; i < m.n_dofs; ++i) dst[i] = 0.0;
  for (int colour = 0; colour < 8; ++colour) {
#pragma omp parallel
    {
      vector<double> loc, out;
#pragma omp for schedule(static)
      for (int cell = 0; cell < m.n_cells; ++cell) {
        int cx = cell % m.n[0], cy = (cell / m.n[0]) % m.n[1], cz = cell / (m.n[0] * m.n[1]);
        if (((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) != colour) continue;
        gather(m, cell, src, zero_constrained, loc);
        out.assign(loc.size(), 0.0);
        fn(cell, loc.data(), out.data());
        for (int i = 0; i < N3; ++i) {
          int64_t node = m.cell_nodes[(size_t)cell * N3 + i];
          for (int c = 0; c < m.nc; ++c) {
            int64_t dof = node * m.nc + c;
            if (zero_constrained && m.is_con[dof]) continue;
            dst[dof] += out[(size_t)c * N3 + i];
          }
        }
      }
    }
  }
  if (constrained_identity)
    for (int32_t c : m.constrained) dst[c] += src[c];
}

// MatrixFreeTools::compute_diagonal restated: local unit vectors per cell.
template <class CellFn>
static void diagonal_loop(const Mesh &m, double *diag, CellFn fn) {
  const int N3 = m.N * m.N * m.N, nd = m.nc * N3;
  for (int64_t i = 0; i < m.n_dofs; ++i) diag[i] = 0.0;
  for (int colour = 0; colour < 8; ++colour) {
#pragma omp parallel
    {
      vector<double> e(nd), out(nd);
#pragma omp for schedule(static)
      for (int cell = 0; cell < m.n_cells; ++cell) {
        int cx = cell % m.n[0], cy = (cell / m.n[0]) % m.n[1], cz = cell / (m.n[0] * m.n[1]);
        if (((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) != colour) continue;
        for (int j = 0; j < nd; ++j) {
          int64_t dof = (int64_t)m.cell_nodes[(size_t)cell * N3 + (j % N3)] * m.nc + j / N3;
          if (m.is_con[dof]) continue;
          std::fill(e.begin(), e.end(), 0.0);
          e[j] = 1.0;
          std::fill(out.begin(), out.end(), 0.0);
          fn(cell, e.data(), out.data());
          diag[dof] += out[j];
        }
      }
    }
  }
  for (int32_t c : m.constrained) diag[c] = 1.0;
}

// ---------------------------------------------------------- brute force ---
// Dense (non-factorised) shape gradients at one q of one cell, reference frame.
static void shape_at_q(const Basis &B, int q, const double *jinv, vector<double> &phi,
                       vector<double> &dphi /* N3 x 3, Grad_X */) {
  const int N = B.N, N3 = N * N * N, Q = B.Q;
  const int q0 = q % Q, q1 = (q / Q) % Q, q2 = q / (Q * Q);
  phi.resize(N3);
  dphi.resize(N3 * 3);
  for (int c = 0; c < N; ++c)
    for (int b = 0; b < N; ++b)
      for (int a = 0; a < N; ++a) {
        const int i = a + N * (b + N * c);
        double sa = B.S[q0 * N + a], sb = B.S[q1 * N + b], sc = B.S[q2 * N + c];
        double da = B.Sd[q0 * N + a], db = B.Sd[q1 * N + b], dc = B.Sd[q2 * N + c];
        phi[i] = sa * sb * sc;
        double gx[3] = {da * sb * sc, sa * db * sc, sa * sb * dc};
        for (int e = 0; e < 3; ++e)
          dphi[i * 3 + e] = gx[0] * jinv[0 * 3 + e] + gx[1] * jinv[1 * 3 + e] + gx[2] * jinv[2 * 3 + e];
      }
}

static void solid_bruteforce_cell(const Solid &s, int cell, const double *loc, double *out) {
  const Material &M = s.material(cell);
  // First Piola stress derivative of the one-field neo-Hookean (formulation 1):
  // dP_ij/dF_kl = mu d_ik d_jl + (mu - 2 lambda lnJ) Fi_li Fi_jk + 2 lambda Fi_ji Fi_lk
  const Mesh &m = *s.m;
  const int N3 = m.N * m.N * m.N, Q3 = s.Q3;
  vector<double> phi, dphi, ucell;
  gather(m, cell, s.u.data(), false, ucell);
  for (int q = 0; q < Q3; ++q) {
    const size_t o = (size_t)cell * Q3 + q;
    shape_at_q(s.B, q, &s.G.jinv[o * 9], phi, dphi);
    double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, gv[3][3] = {}, v[3] = {};
    for (int i = 0; i < N3; ++i)
      for (int d = 0; d < 3; ++d) {
        v[d] += loc[d * N3 + i] * phi[i];
        for (int e = 0; e < 3; ++e) {
          F[d][e] += ucell[d * N3 + i] * dphi[i * 3 + e];
          gv[d][e] += loc[d * N3 + i] * dphi[i * 3 + e];
        }
      }
    const double J = det3(F);
    double Fi[3][3];
    inv3(F, Fi);
    const double sc = M.mu - 2.0 * M.lambda * std::log(J);
    double trFig = 0;
    for (int l = 0; l < 3; ++l)
      for (int k = 0; k < 3; ++k) trFig += Fi[l][k] * gv[k][l];
    double dP[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = M.mu * gv[i][j] + 2.0 * M.lambda * Fi[j][i] * trFig;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) t += sc * Fi[l][i] * Fi[j][k] * gv[k][l];
        dP[i][j] = t;
      }
    const double w = s.G.jxw[o];
    for (int i = 0; i < N3; ++i)
      for (int d = 0; d < 3; ++d) {
        double t = M.rho_alpha * v[d] * phi[i];
        for (int e = 0; e < 3; ++e) t += dP[d][e] * dphi[i * 3 + e];
        out[d * N3 + i] += t * w;
      }
  }
}

static void heat_bruteforce_cell(const Heat &h, int cell, const double *loc, double *out) {
  const int N3 = h.m->N * h.m->N * h.m->N, Q3 = h.Q3;
  vector<double> phi, dphi;
  for (int q = 0; q < Q3; ++q) {
    const size_t o = (size_t)cell * Q3 + q;
    shape_at_q(h.B, q, &h.G.jinv[o * 9], phi, dphi);
    double v = 0, g[3] = {0, 0, 0};
    for (int i = 0; i < N3; ++i) {
      v += loc[i] * phi[i];
      for (int e = 0; e < 3; ++e) g[e] += loc[i] * dphi[i * 3 + e];
    }
    const double c = h.coef.empty() ? 1.0 : h.coef[o];
    for (int i = 0; i < N3; ++i)
      out[i] += (v * phi[i] + h.dt * c * (g[0] * dphi[i * 3] + g[1] * dphi[i * 3 + 1] + g[2] * dphi[i * 3 + 2])) *
                h.G.jxw[o];
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

void orc_gll_points(int p, double *x) { gll_points(p, x); }
void orc_gauss_rule(int n, double *x, double *w) { gauss_rule(n, x, w); }
void orc_lagrange(int m, const double *z, int nx, const double *x, double *val, double *der) {
  for (int i = 0; i < nx; ++i) lagrange_eval(m, z, x[i], val + (size_t)i * m, der + (size_t)i * m);
}

void *orc_mesh_create(const double *lo, const double *hi, const int *n, int p, int nc, int face_mask,
                      int comp_mask) {
  int face_comp[6];
  for (int f = 0; f < 6; ++f) face_comp[f] = ((face_mask >> f) & 1) ? comp_mask : 0;
  return orc_mesh_create_bc(lo, hi, n, p, nc, face_comp);
}
void *orc_mesh_create_bc(const double *lo, const double *hi, const int *n, int p, int nc, const int *face_comp) {
  Mesh *m = new Mesh();
  build_mesh(*m, lo, hi, n, p, nc, face_comp);
  return m;
}
void orc_mesh_destroy(void *m) { delete (Mesh *)m; }
void orc_mesh_sizes(void *mp, int64_t *out) {
  Mesh &m = *(Mesh *)mp;
  out[0] = m.n_cells; out[1] = m.n_nodes; out[2] = m.n_dofs;
  out[3] = (int64_t)m.constrained.size(); out[4] = m.N * m.N * m.N; out[5] = m.nc;
}
void orc_mesh_cell_nodes(void *mp, int32_t *out) {
  Mesh &m = *(Mesh *)mp;
  std::memcpy(out, m.cell_nodes.data(), m.cell_nodes.size() * sizeof(int32_t));
}
void orc_mesh_constrained(void *mp, int32_t *out) {
  Mesh &m = *(Mesh *)mp;
  std::memcpy(out, m.constrained.data(), m.constrained.size() * sizeof(int32_t));
}
void orc_mesh_node_coords(void *mp, double *out) {
  Mesh &m = *(Mesh *)mp;
  std::memcpy(out, m.xnode.data(), m.xnode.size() * sizeof(double));
}
void orc_mesh_cell_vertices(void *mp, double *out) {
  Mesh &m = *(Mesh *)mp;
  std::memcpy(out, m.vert.data(), m.vert.size() * sizeof(double));
}

// Replaces the cell vertices (n_cells*8*3) and moves every node to the image
// of its cell-local Gauss-Lobatto point under the cell's trilinear map, so the
// support points stay the nodal interpolant of that map.  A conforming vertex
// field (shared vertices equal) gives shared nodes one image.
void orc_mesh_set_cell_vertices(void *mp, const double *v) {
  Mesh &m = *(Mesh *)mp;
  std::memcpy(m.vert.data(), v, m.vert.size() * sizeof(double));
  const int N = m.N, N3 = N * N * N;
  for (int c = 0; c < m.n_cells; ++c)
    for (int l = 0; l < N3; ++l) {
      const double xi[3] = {m.gll[l % N], m.gll[(l / N) % N], m.gll[l / (N * N)]};
      double x[3] = {0, 0, 0};
      for (int w = 0; w < 8; ++w) {
        double phi = 1;
        for (int d = 0; d < 3; ++d) phi *= ((w >> d) & 1) ? xi[d] : 1 - xi[d];
        for (int e = 0; e < 3; ++e) x[e] += phi * v[(size_t)c * 24 + w * 3 + e];
      }
      const int64_t node = m.cell_nodes[(size_t)c * N3 + l];
      for (int e = 0; e < 3; ++e) m.xnode[node * 3 + e] = x[e];
    }
}

void orc_fill_u0(void *mp, double amp, double noise, uint64_t seed, double *u) {
  Mesh &m = *(Mesh *)mp;
  for (int64_t node = 0; node < m.n_nodes; ++node) {
    const double *x = &m.xnode[node * 3];
    const double base = amp * std::sin(M_PI * x[0]) * std::sin(M_PI * x[1]) * std::sin(M_PI * x[2]);
    for (int c = 0; c < m.nc; ++c) {
      int64_t dof = node * m.nc + c;
      u[dof] = m.is_con[dof] ? 0.0 : base + noise * uniform_pm1(seed, (uint64_t)dof);
    }
  }
}
void orc_fill_uniform(uint64_t seed, int64_t n, double *v) {
  for (int64_t i = 0; i < n; ++i) v[i] = uniform_pm1(seed, (uint64_t)i);
}

void *orc_solid_create(void *mesh, int n_q, double mu, double nu, double rho, double alpha_1, int formulation,
                       int caching) {
  Solid *s = new Solid();
  s->m = (Mesh *)mesh;
  if (s->m->nc != 3) { delete s; return nullptr; }
  if (formulation == 0 && caching != 1) { delete s; return nullptr; }  // mf_nh_operator.h:706
  s->B.init(*s->m, n_q);
  s->G.init(*s->m, s->B);
  s->Q3 = n_q * n_q * n_q;
  s->alpha1 = alpha_1;
  s->form = formulation; s->caching = caching;
  s->mat[0] = s->mat[1] = material_constants(mu, nu, rho, alpha_1);
  return s;
}
void orc_solid_set_inclusion(void *op, const int32_t *material_ids, double mu, double nu, double rho) {
  Solid &s = *(Solid *)op;
  if (material_ids) {
    s.mat_id.assign(material_ids, material_ids + s.m->n_cells);
    s.mat[1] = material_constants(mu, nu, rho, s.alpha1);
  } else {
    s.mat_id.clear();
    s.mat[1] = s.mat[0];
  }
}
void orc_solid_destroy(void *op) { delete (Solid *)op; }
int orc_solid_cache(void *op, const double *u) { return solid_cache(*(Solid *)op, u); }
void orc_solid_vmult(void *op, double *dst, const double *src) {
  Solid &s = *(Solid *)op;
  cell_loop(*s.m, src, dst, true, true,
            [&](int cell, const double *loc, double *out) { solid_cell(s, cell, loc, out); });
}
void orc_solid_vmult_bruteforce(void *op, double *dst, const double *src) {
  Solid &s = *(Solid *)op;
  cell_loop(*s.m, src, dst, true, true,
            [&](int cell, const double *loc, double *out) { solid_bruteforce_cell(s, cell, loc, out); });
}
void orc_solid_diagonal(void *op, double *diag) {
  Solid &s = *(Solid *)op;
  diagonal_loop(*s.m, diag, [&](int cell, const double *loc, double *out) { solid_cell(s, cell, loc, out); });
}
void orc_solid_tensor2_data(void *op, double *out) {
  Solid &s = *(Solid *)op;
  const size_t nq = (size_t)s.m->n_cells * s.Q3;
  for (size_t o = 0; o < nq; ++o) {
    for (int i = 0; i < 9; ++i) out[i * nq + o] = s.jinv_cur[o * 9 + i];
    out[9 * nq + o] = s.jxw_cur[o];
    out[10 * nq + o] = s.sc[o];
    out[11 * nq + o] = s.sc2[o];
    const double *T = &s.t2[o * 9];
    out[12 * nq + o] = T[0]; out[13 * nq + o] = T[4]; out[14 * nq + o] = T[8];
    out[15 * nq + o] = T[1]; out[16 * nq + o] = T[2]; out[17 * nq + o] = T[5];
  }
}

void orc_interface_traction(void *mp, int64_t n_faces, const int32_t *faces, const double *u, const double *t,
                            double *rhs) {
  // mf_elasticity.h:1849-1897: t0 = |det F F^{-T} N| t on the reference face.
  // Restated through Nanson's formula, |det F F^{-T} N| dA = |x_a x x_b| dxi_a
  // dxi_b with x = X + u and a, b the face's tangential reference directions:
  // the deformed area element straight from the deformed tangents, with no
  // F^{-1} and no normal.
  const Mesh &m = *(const Mesh *)mp;
  const int N = m.N, N3 = N * N * N;
  vector<double> gq(N), gw(N), val(N), der(N);
  gauss_rule(N, gq.data(), gw.data());
  for (int64_t f = 0; f < n_faces; ++f) {
    const int cell = faces[2 * f], face = faces[2 * f + 1];
    const int axis = face / 2, side = face % 2;
    const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
    const int32_t *cn = &m.cell_nodes[(size_t)cell * N3];
    for (int qb = 0; qb < N; ++qb)
      for (int qa = 0; qa < N; ++qa) {
        double xi[3];
        xi[axis] = side;
        xi[ta] = gq[qa];
        xi[tb] = gq[qb];
        double v[3][16], d[3][16];
        for (int e = 0; e < 3; ++e) {
          lagrange_eval(N, m.gll.data(), xi[e], val.data(), der.data());
          for (int a = 0; a < N; ++a) {
            v[e][a] = val[a];
            d[e][a] = der[a];
          }
        }
        double J[3][3];
        q1_jacobian(&m.vert[(size_t)cell * 24], xi, J);
        double xa[3], xb[3];
        for (int c = 0; c < 3; ++c) {
          xa[c] = J[c][ta];
          xb[c] = J[c][tb];
        }
        for (int k = 0; k < N; ++k)
          for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
              const int ijk[3] = {i, j, k};
              double da = 1, db = 1;
              for (int e = 0; e < 3; ++e) {
                da *= e == ta ? d[e][ijk[e]] : v[e][ijk[e]];
                db *= e == tb ? d[e][ijk[e]] : v[e][ijk[e]];
              }
              const int64_t node = cn[i + N * (j + N * k)];
              for (int c = 0; c < 3; ++c) {
                xa[c] += u[node * 3 + c] * da;
                xb[c] += u[node * 3 + c] * db;
              }
            }
        const double cr[3] = {xa[1] * xb[2] - xa[2] * xb[1], xa[2] * xb[0] - xa[0] * xb[2], xa[0] * xb[1] - xa[1] * xb[0]};
        const double area = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) * gw[qa] * gw[qb];
        const double *tq = t + ((size_t)f * N * N + qa + (size_t)N * qb) * 3;
        for (int k = 0; k < N; ++k)
          for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
              const double phi = v[0][i] * v[1][j] * v[2][k];
              if (phi == 0.0) continue;
              const int64_t node = cn[i + N * (j + N * k)];
              for (int c = 0; c < 3; ++c)
                if (!m.is_con[node * 3 + c]) rhs[node * 3 + c] += tq[c] * phi * area;
            }
      }
  }
}

// Face points of (cell, face = 2 axis + side): reference coordinates xi of
// point qa + N qb over the Gauss points of the two tangential directions.
static void face_point_xi(int face, int qa, int qb, const vector<double> &gq, double xi[3], int &ta, int &tb) {
  const int axis = face / 2, side = face % 2;
  ta = axis == 0 ? 1 : 0;
  tb = axis == 2 ? 1 : 2;
  xi[axis] = side;
  xi[ta] = gq[qa];
  xi[tb] = gq[qb];
}

void orc_face_values(void *mp, int64_t n_faces, const int32_t *faces, const double *u, double *out) {
  // full tensor-product evaluation over all cell nodes (FEFaceEvaluation
  // values at the interface quadrature points, quad_interface.h:198-340)
  const Mesh &m = *(const Mesh *)mp;
  const int N = m.N, N3 = N * N * N, nc = m.nc;
  vector<double> gq(N), gw(N), val(N), der(N);
  gauss_rule(N, gq.data(), gw.data());
  for (int64_t f = 0; f < n_faces; ++f) {
    const int32_t *cn = &m.cell_nodes[(size_t)faces[2 * f] * N3];
    for (int qb = 0; qb < N; ++qb)
      for (int qa = 0; qa < N; ++qa) {
        double xi[3];
        int ta, tb;
        face_point_xi(faces[2 * f + 1], qa, qb, gq, xi, ta, tb);
        double v[3][16];
        for (int e = 0; e < 3; ++e) {
          lagrange_eval(N, m.gll.data(), xi[e], val.data(), der.data());
          for (int a = 0; a < N; ++a) v[e][a] = val[a];
        }
        double *o = out + ((size_t)f * N * N + qa + (size_t)N * qb) * nc;
        for (int c = 0; c < nc; ++c) o[c] = 0;
        for (int k = 0; k < N; ++k)
          for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
              const double phi = v[0][i] * v[1][j] * v[2][k];
              const int64_t node = cn[i + N * (j + N * k)];
              for (int c = 0; c < nc; ++c) o[c] += phi * u[node * nc + c];
            }
      }
  }
}

void orc_interface_flux(void *mp, int64_t n_faces, const int32_t *faces, const double *flux, double dt, double *rhs) {
  // heat_transfer.h:618-652: submit -q dt on the face, integrate; the area
  // element from the cross product of the face's reference tangents
  const Mesh &m = *(const Mesh *)mp;
  const int N = m.N, N3 = N * N * N;
  vector<double> gq(N), gw(N), val(N), der(N);
  gauss_rule(N, gq.data(), gw.data());
  for (int64_t f = 0; f < n_faces; ++f) {
    const int cell = faces[2 * f];
    const int32_t *cn = &m.cell_nodes[(size_t)cell * N3];
    for (int qb = 0; qb < N; ++qb)
      for (int qa = 0; qa < N; ++qa) {
        double xi[3];
        int ta, tb;
        face_point_xi(faces[2 * f + 1], qa, qb, gq, xi, ta, tb);
        double J[3][3];
        q1_jacobian(&m.vert[(size_t)cell * 24], xi, J);
        const double a[3] = {J[0][ta], J[1][ta], J[2][ta]}, b[3] = {J[0][tb], J[1][tb], J[2][tb]};
        const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double w = -dt * flux[(size_t)f * N * N + qa + (size_t)N * qb] *
                         std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) * gw[qa] * gw[qb];
        double v[3][16];
        for (int e = 0; e < 3; ++e) {
          lagrange_eval(N, m.gll.data(), xi[e], val.data(), der.data());
          for (int k = 0; k < N; ++k) v[e][k] = val[k];
        }
        for (int k = 0; k < N; ++k)
          for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
              const double phi = v[0][i] * v[1][j] * v[2][k];
              const int64_t node = cn[i + N * (j + N * k)];
              if (phi != 0.0 && !m.is_con[node * m.nc]) rhs[node * m.nc] += w * phi;
            }
      }
  }
}

void orc_solid_residual(void *op, const double *u, const double *acc, const double *body, double *R) {
  // mf_elasticity.h:1764-1846 (fast FEEvaluation path) + set_zero (1905)
  Solid &s = *(Solid *)op;
  const Mesh &m = *s.m;
  const int N3 = m.N * m.N * m.N, Q3 = s.Q3;
  for (int64_t i = 0; i < m.n_dofs; ++i) R[i] = 0.0;
  for (int colour = 0; colour < 8; ++colour) {
#pragma omp parallel
    {
      vector<double> lu, la, uval, ugr, aval, agr, rv(3 * Q3), rg(9 * Q3), out;
#pragma omp for schedule(static)
      for (int cell = 0; cell < m.n_cells; ++cell) {
        int cx = cell % m.n[0], cy = (cell / m.n[0]) % m.n[1], cz = cell / (m.n[0] * m.n[1]);
        if (((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)) != colour) continue;
#ifdef ORC_PLANT_INCL_RESIDUAL_MAT0
        const Material &M = s.mat[0];  // planted: the residual ignores the cell's material
#else
        const Material &M = s.material(cell);  // mf_elasticity.h:1781-1784
#endif
        gather(m, cell, u, false, lu);
        gather(m, cell, acc, false, la);
        eval_vector(s.B, 3, N3, lu.data(), uval, ugr);
        eval_vector(s.B, 3, N3, la.data(), aval, agr);
        for (int q = 0; q < Q3; ++q) {
          const size_t o = (size_t)cell * Q3 + q;
          const double *jinv = &s.G.jinv[o * 9];
          double F[3][3], Fi[3][3], tau[3][3];
          deformation_gradient(ugr.data(), Q3, q, jinv, F);
          inv3(F, Fi);
          get_tau(s, M, F, tau);
          for (int d = 0; d < 3; ++d) {
            rv[(size_t)d * Q3 + q] = (-aval[(size_t)d * Q3 + q] + body[d]) * M.rho * s.G.jxw[o];
            for (int f = 0; f < 3; ++f) {
              // -(tau F^-T)[d][e] mapped to the reference gradient
              double t = 0;
              for (int e = 0; e < 3; ++e) {
                double P = 0;
                for (int k = 0; k < 3; ++k) P += tau[d][k] * Fi[e][k];
                t += -P * jinv[f * 3 + e];
              }
              rg[(size_t)(d * 3 + f) * Q3 + q] = t * s.G.jxw[o];
            }
          }
        }
        out.assign(3 * N3, 0.0);
        for (int d = 0; d < 3; ++d)
          integrate(s.B, &rv[(size_t)d * Q3], &rg[(size_t)(d * 3) * Q3], &rg[(size_t)(d * 3 + 1) * Q3],
                    &rg[(size_t)(d * 3 + 2) * Q3], out.data() + (size_t)d * N3);
        for (int i = 0; i < N3; ++i) {
          int64_t node = m.cell_nodes[(size_t)cell * N3 + i];
          for (int d = 0; d < 3; ++d) R[node * 3 + d] += out[(size_t)d * N3 + i];
        }
      }
    }
  }
  for (int32_t c : m.constrained) R[c] = 0.0;
}

void *orc_heat_create(void *mesh, int n_q, double dt, const double *coef) {
  Heat *h = new Heat();
  h->m = (Mesh *)mesh;
  if (h->m->nc != 1) { delete h; return nullptr; }
  h->B.init(*h->m, n_q);
  h->G.init(*h->m, h->B);
  h->Q3 = n_q * n_q * n_q;
  h->dt = dt;
  if (coef) h->coef.assign(coef, coef + (size_t)h->m->n_cells * h->Q3);
  return h;
}
void orc_heat_destroy(void *op) { delete (Heat *)op; }
void orc_heat_vmult(void *op, double *dst, const double *src) {
  Heat &h = *(Heat *)op;
  cell_loop(*h.m, src, dst, true, true, [&](int cell, const double *loc, double *out) { heat_cell(h, cell, loc, out); });
}
void orc_heat_vmult_noconstraints(void *op, double *dst, const double *src) {
  Heat &h = *(Heat *)op;
  cell_loop(*h.m, src, dst, false, false,
            [&](int cell, const double *loc, double *out) { heat_cell(h, cell, loc, out); });
}
void orc_heat_vmult_bruteforce(void *op, double *dst, const double *src) {
  Heat &h = *(Heat *)op;
  cell_loop(*h.m, src, dst, true, true,
            [&](int cell, const double *loc, double *out) { heat_bruteforce_cell(h, cell, loc, out); });
}
void orc_heat_diagonal(void *op, double *diag) {
  Heat &h = *(Heat *)op;
  diagonal_loop(*h.m, diag, [&](int cell, const double *loc, double *out) { heat_cell(h, cell, loc, out); });
}
void *orc_operator_mesh(int kind, void *op) {
  return kind == 0 ? (void *)((Solid *)op)->m : (void *)((Heat *)op)->m;
}
void orc_heat_load(void *op, double f, double *b) {
  Heat &h = *(Heat *)op;
  const Mesh &m = *h.m;
  const int N3 = m.N * m.N * m.N, Q3 = h.Q3;
  for (int64_t i = 0; i < m.n_dofs; ++i) b[i] = 0;
  vector<double> rv(Q3), z(Q3, 0.0), out(N3);
  for (int cell = 0; cell < m.n_cells; ++cell) {
    for (int q = 0; q < Q3; ++q) rv[q] = f * h.G.jxw[(size_t)cell * Q3 + q];
    std::fill(out.begin(), out.end(), 0.0);
    integrate(h.B, rv.data(), z.data(), z.data(), z.data(), out.data());
    for (int i = 0; i < N3; ++i) b[m.cell_nodes[(size_t)cell * N3 + i]] += out[i];
  }
}

}  // extern "C"
