// CPU oracle (test infrastructure only; see oracle.h): multigrid transfer,
// Chebyshev smoother, CG and the V-cycle, restating the solver settings of
// mf_elasticity.h:1140-1279, 1981-2070 and heat_transfer.h:757-812.  The
// deal.II solver internals these settings configure (PreconditionChebyshev,
// SolverCG, Multigrid) are not part of the reference sources; their
// behaviour is written here from the library's documented semantics.
#include <omp.h>

#include <functional>

#include "fe.hpp"
#include "oracle.h"

namespace orc {

// Which coarse cell (along one direction) holds the fine node index I of a
// once-refined grid, and the local coordinate in that coarse cell.
static void locate_fine_node(const Mesh &c, const Mesh &f, int d, int I, int &cc, double &xi) {
  const int p = f.p;
  int fc = std::min(I / p, f.n[d] - 1);
  int a = I - fc * p;
  double t = 0.5 * (fc + f.gll[a]);  // in coarse-cell units
  cc = std::min((int)std::floor(t), c.n[d] - 1);
  xi = t - cc;
  if (xi < 0) xi = 0;
}

// P[F][A] = phi_A(x_F): value of coarse basis function A at fine node F.
// visit(F, A, weight) over all non-zero entries with coarse constrained
// columns and fine constrained rows removed.
static void transfer_entries(const Mesh &c, const Mesh &f,
                             const std::function<void(int64_t, int64_t, double)> &visit) {
  const int N = c.N;
  vector<double> val(N), der(N);
  vector<vector<double>> w1[3];
  vector<int> cc1[3];
  for (int d = 0; d < 3; ++d) {
    w1[d].resize(f.NN[d]);
    cc1[d].resize(f.NN[d]);
    for (int I = 0; I < f.NN[d]; ++I) {
      double xi;
      locate_fine_node(c, f, d, I, cc1[d][I], xi);
      lagrange_eval(N, c.gll.data(), xi, val.data(), der.data());
      w1[d][I] = val;
    }
  }
  for (int K = 0; K < f.NN[2]; ++K)
    for (int J = 0; J < f.NN[1]; ++J)
      for (int I = 0; I < f.NN[0]; ++I) {
        const int64_t F = f.node_index(I, J, K);
        for (int kc = 0; kc < N; ++kc)
          for (int jc = 0; jc < N; ++jc)
            for (int ic = 0; ic < N; ++ic) {
              double w = w1[0][I][ic] * w1[1][J][jc] * w1[2][K][kc];
              if (w == 0.0) continue;
              const int64_t A = c.node_index(cc1[0][I] * c.p + ic, cc1[1][J] * c.p + jc, cc1[2][K] * c.p + kc);
              visit(F, A, w);
            }
      }
}

void mg_prolongate_add(const Mesh &c, const Mesh &f, const double *cv, double *fv) {
  const int nc = c.nc;
  transfer_entries(c, f, [&](int64_t F, int64_t A, double w) {
    for (int k = 0; k < nc; ++k) {
      if (f.is_con[F * nc + k] || c.is_con[A * nc + k]) continue;
      fv[F * nc + k] += w * cv[A * nc + k];
    }
  });
}
void mg_restrict_add(const Mesh &c, const Mesh &f, const double *fv, double *cv) {
  const int nc = c.nc;
  transfer_entries(c, f, [&](int64_t F, int64_t A, double w) {
    for (int k = 0; k < nc; ++k) {
      if (f.is_con[F * nc + k] || c.is_con[A * nc + k]) continue;
      cv[A * nc + k] += w * fv[F * nc + k];
    }
  });
}
// coarse nodal values = fine FE function at the coarse support points
void mg_interpolate(const Mesh &c, const Mesh &f, const double *fv, double *cv) {
  const int N = f.N, nc = c.nc;
  vector<double> val[3], der(N);
  for (int d = 0; d < 3; ++d) val[d].resize(N);
  for (int K = 0; K < c.NN[2]; ++K)
    for (int J = 0; J < c.NN[1]; ++J)
      for (int I = 0; I < c.NN[0]; ++I) {
        const int idx[3] = {I, J, K};
        int fc[3];
        for (int d = 0; d < 3; ++d) {
          // coordinate of the coarse node in fine-cell units
          int ccell = std::min(idx[d] / c.p, c.n[d] - 1);
          int a = idx[d] - ccell * c.p;
          double t = 2.0 * (ccell + c.gll[a]);
          fc[d] = std::min((int)std::floor(t), f.n[d] - 1);
          double xi = t - fc[d];
          lagrange_eval(N, f.gll.data(), xi, val[d].data(), der.data());
        }
        const int64_t A = c.node_index(I, J, K);
        for (int k = 0; k < nc; ++k) {
          double s = 0;
          for (int kc = 0; kc < N; ++kc)
            for (int jc = 0; jc < N; ++jc)
              for (int ic = 0; ic < N; ++ic) {
                const int64_t F = f.node_index(fc[0] * f.p + ic, fc[1] * f.p + jc, fc[2] * f.p + kc);
                s += val[0][ic] * val[1][jc] * val[2][kc] * fv[F * nc + k];
              }
          cv[A * nc + k] = s;
        }
      }
}

// ------------------------------------------------------------- algebra ---
static double dot(const vector<double> &a, const vector<double> &b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// A level operator: a solid (kind 0) or heat (kind 1) oracle operator on a
// mesh, or (kind 2) an explicit dense row-major matrix, which lets tests
// check the smoother against closed forms on a small spectrum.
struct LevelOp {
  int kind = 0;
  void *op = nullptr;
  Mesh *m = nullptr;
  const double *dense = nullptr;
  size_t dense_n = 0;
  vector<double> inv_diag;
  size_t n() const { return kind == 2 ? dense_n : (size_t)m->n_dofs; }
  void vmult(vector<double> &dst, const vector<double> &src) const {
    if (kind == 0) orc_solid_vmult(op, dst.data(), src.data());
    else if (kind == 1) orc_heat_vmult(op, dst.data(), src.data());
    else
      for (size_t i = 0; i < dense_n; ++i) {
        double s = 0;
        for (size_t j = 0; j < dense_n; ++j) s += dense[i * dense_n + j] * src[j];
        dst[i] = s;
      }
  }
  void diagonal() {
    inv_diag.resize(n());
    if (kind == 0) orc_solid_diagonal(op, inv_diag.data());
    else if (kind == 1) orc_heat_diagonal(op, inv_diag.data());
    else
      for (size_t i = 0; i < dense_n; ++i) inv_diag[i] = dense[i * dense_n + i];
    for (auto &d : inv_diag) d = 1.0 / d;
  }
};

// Eigenvalues of the symmetric tridiagonal Lanczos matrix by bisection.
static void tridiag_eigs(const vector<double> &dg, const vector<double> &off, double &lmin, double &lmax) {
  const int n = (int)dg.size();
  if (n == 0) { lmin = lmax = 0; return; }
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::abs(off[i - 1]) : 0) + (i < n - 1 ? std::abs(off[i]) : 0);
    lo = std::min(lo, dg[i] - r);
    hi = std::max(hi, dg[i] + r);
  }
  auto count_below = [&](double x) {
    int cnt = 0;
    double q = dg[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < n; ++i) {
      double qq = (q == 0) ? 1e-300 : q;
      q = dg[i] - x - off[i - 1] * off[i - 1] / qq;
      if (q < 0) cnt++;
    }
    return cnt;
  };
  auto kth = [&](int k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (a + b);
      if (count_below(mid) > k) b = mid; else a = mid;
    }
    return 0.5 * (a + b);
  };
  lmin = kth(0);
  lmax = kth(n - 1);
}

// Initial vector of the eigenvalue-estimation CG: 1 except every 11th
// global index, constrained entries 0 (a fixed, partition-independent choice).
static void eig_initial(const LevelOp &A, vector<double> &v) {
  const size_t n = A.n();
  v.resize(n);
  for (size_t i = 0; i < n; ++i) v[i] = (i % 11 == 0) ? 0.0 : 1.0;
  if (A.m)
    for (int32_t c : A.m->constrained) v[c] = 0.0;
}

// Preconditioned CG that records the Lanczos coefficients (SolverCG with
// condition-number tracking): n_it iterations on A x = b, x0 = 0.
static void cg_eigen(const LevelOp &A, const vector<double> &b, int n_it, double &lmin, double &lmax) {
  const size_t n = A.n();
  vector<double> x(n, 0.0), r(b), d(n), p(n), h(n);
  for (size_t i = 0; i < n; ++i) d[i] = A.inv_diag[i] * r[i];
  double gh = dot(r, d);
  p = d;
  vector<double> alphas, betas, dg, off;
  double res0 = std::sqrt(dot(r, r));
  for (int it = 0; it < n_it; ++it) {
    A.vmult(h, p);
    double ph = dot(p, h);
    double alpha = gh / ph;
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * h[i]; }
    alphas.push_back(alpha);
    double res = std::sqrt(dot(r, r));
    if (res <= 1e-15 * res0) break;  // converged: stop collecting
    for (size_t i = 0; i < n; ++i) d[i] = A.inv_diag[i] * r[i];
    double ghn = dot(r, d);
    double beta = ghn / gh;
    gh = ghn;
    betas.push_back(beta);
    for (size_t i = 0; i < n; ++i) p[i] = d[i] + beta * p[i];
  }
  const int k = (int)alphas.size();
  dg.resize(k);
  off.resize(k > 0 ? k - 1 : 0);
  for (int i = 0; i < k; ++i) {
    dg[i] = 1.0 / alphas[i] + (i > 0 ? betas[i - 1] / alphas[i - 1] : 0.0);
    if (i < k - 1) off[i] = std::sqrt(betas[i]) / alphas[i];
  }
  tridiag_eigs(dg, off, lmin, lmax);
}

struct Chebyshev {
  const LevelOp *A = nullptr;
  int degree = 5;
  double theta = 1, delta = 0;
  double lmin = 0, lmax = 0;
  // smoothing_range > 1: [1.2 lmax / range, 1.2 lmax] (mf_elasticity.h:1167-1171)
  // smoothing_range < 1 ("solver" mode, degree chosen from the estimate)
  void setup(const LevelOp &op, int deg, double range, int eig_it) {
    A = &op;
    vector<double> v;
    eig_initial(op, v);
    cg_eigen(op, v, eig_it, lmin, lmax);
    const double max_ev = 1.2 * lmax;
    const double alpha = range > 1.0 ? max_ev / range : std::min(0.9 * lmax, lmin);
    theta = 0.5 * (max_ev + alpha);
    delta = 0.5 * (max_ev - alpha);
    degree = deg;
    if (deg <= 0) {
      const double actual_range = max_ev / alpha;
      const double sigma = (1.0 - std::sqrt(1.0 / actual_range)) / (1.0 + std::sqrt(1.0 / actual_range));
      const double eps = range;
      degree = 1 + (int)(std::log(1.0 / eps + std::sqrt(1.0 / eps / eps - 1.0)) / std::log(1.0 / sigma));
    }
  }
  // x = p(A) b, zero initial guess: degree-1 operator applications
  void vmult(vector<double> &x, const vector<double> &b) const { run(x, b, true); }
  // x <- x + p(A)(b - A x): degree operator applications
  void step(vector<double> &x, const vector<double> &b) const { run(x, b, false); }
  void run(vector<double> &x, const vector<double> &b, bool zero_guess) const {
    const size_t n = A->n();
    const vector<double> &D = A->inv_diag;
    vector<double> r(n), d(n), t(n);
    if (zero_guess) {
      for (size_t i = 0; i < n; ++i) { d[i] = D[i] * b[i] / theta; x[i] = d[i]; }
    } else {
      A->vmult(t, x);
      for (size_t i = 0; i < n; ++i) { d[i] = D[i] * (b[i] - t[i]) / theta; x[i] += d[i]; }
    }
    if (degree < 2 || std::abs(delta) < 1e-40) return;
    double rhok = delta / theta;
    const double sigma = theta / delta;
    for (int k = 0; k < degree - 1; ++k) {
      A->vmult(t, x);
      const double rhokp = 1.0 / (2.0 * sigma - rhok);
#ifdef ORC_PLANT_CHEB_F2
      const double f1 = rhokp * rhok, f2 = rhokp / delta;  // planted mistake (mutation tests only)
#else
      const double f1 = rhokp * rhok, f2 = 2.0 * rhokp / delta;
#endif
      rhok = rhokp;
      for (size_t i = 0; i < n; ++i) {
        d[i] = f1 * d[i] + f2 * D[i] * (b[i] - t[i]);
        x[i] += d[i];
      }
    }
  }
};

struct MGSolver {
  int L = 0;  // finest level index
  vector<Mesh *> meshes;
  vector<LevelOp> ops;
  vector<Chebyshev> cheb;
  int coarse_kind = 0;
  mutable vector<vector<double>> sol, def, tmp;

  void coarse(vector<double> &x, const vector<double> &b) const {
    if (coarse_kind == 0) {
      cheb[0].vmult(x, b);
      return;
    }
    // CG to relative 1e-3 (abs 1e-10), max(m,100) iterations, preconditioned
    // by the level-0 Chebyshev (mf_elasticity.h:1183-1198)
    const LevelOp &A = ops[0];
    const size_t n = A.n();
    vector<double> r(b), d(n), p(n), h(n);
    std::fill(x.begin(), x.end(), 0.0);
    double res0 = std::sqrt(dot(r, r)), res = res0;
    const double tol = std::max(1e-10, 1e-3 * res0);
    if (res <= tol) return;
    cheb[0].vmult(d, r);
    double gh = dot(r, d);
    p = d;
    const int maxit = std::max((int)n, 100);
    for (int it = 0; it < maxit; ++it) {
      A.vmult(h, p);
      double alpha = gh / dot(p, h);
      for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * h[i]; }
      res = std::sqrt(dot(r, r));
      if (res <= tol) return;
      cheb[0].vmult(d, r);
      double ghn = dot(r, d);
      double beta = ghn / gh;
      gh = ghn;
      for (size_t i = 0; i < n; ++i) p[i] = d[i] + beta * p[i];
    }
  }

  void v_cycle(int l) const {
    if (l == 0) { coarse(sol[0], def[0]); return; }
    cheb[l].vmult(sol[l], def[l]);  // pre-smoothing, zero initial guess
    ops[l].vmult(tmp[l], sol[l]);
    for (size_t i = 0; i < tmp[l].size(); ++i) tmp[l][i] = def[l][i] - tmp[l][i];
    std::fill(def[l - 1].begin(), def[l - 1].end(), 0.0);
    mg_restrict_add(*meshes[l - 1], *meshes[l], tmp[l].data(), def[l - 1].data());
    v_cycle(l - 1);
    mg_prolongate_add(*meshes[l - 1], *meshes[l], sol[l - 1].data(), sol[l].data());
#if defined(ORC_PLANT_POST_ZERO_GUESS)
    cheb[l].vmult(sol[l], def[l]);  // planted mistake (mutation tests only): drops the coarse correction
#elif defined(ORC_PLANT_NO_POST_SMOOTH)
    // planted mistake (mutation tests only): no post-smoothing
#else
    cheb[l].step(sol[l], def[l]);  // post-smoothing
#endif
  }
  // PreconditionMG::vmult: copy_to_mg, V-cycle, copy_from_mg
  void precondition(vector<double> &dst, const vector<double> &src) const {
    def[L] = src;
    std::fill(sol[L].begin(), sol[L].end(), 0.0);
    v_cycle(L);
    dst = sol[L];
  }
};

}  // namespace orc

using namespace orc;

extern "C" {

void orc_prolongate_add(void *cm, void *fm, const double *c, double *f) {
  mg_prolongate_add(*(Mesh *)cm, *(Mesh *)fm, c, f);
}
void orc_restrict_add(void *cm, void *fm, const double *f, double *c) {
  mg_restrict_add(*(Mesh *)cm, *(Mesh *)fm, f, c);
}
void orc_interpolate(void *cm, void *fm, const double *f, double *c) {
  mg_interpolate(*(Mesh *)cm, *(Mesh *)fm, f, c);
}

void orc_cheb_dense(int n, const double *Adense, int degree, double range, int eig_it, int zero_guess,
                    const double *b, double *x, double *info) {
  LevelOp A;
  A.kind = 2;
  A.dense = Adense;
  A.dense_n = (size_t)n;
  A.diagonal();
  Chebyshev c;
  c.setup(A, degree, range, eig_it);
  vector<double> bb(b, b + n), xx(x, x + n);
  c.run(xx, bb, zero_guess != 0);
  std::copy(xx.begin(), xx.end(), x);
  if (info) {
    info[0] = c.lmax;
    info[1] = c.lmin;
    info[2] = c.theta;
    info[3] = c.delta;
    info[4] = c.degree;
  }
}

void orc_cheb_eigenvalues(int kind, void *op, void *mesh, int n_it, double *lmax, double *lmin) {
  LevelOp A;
  A.kind = kind; A.op = op; A.m = (Mesh *)mesh;
  A.diagonal();
  vector<double> v;
  eig_initial(A, v);
  cg_eigen(A, v, n_it, *lmin, *lmax);
}

static void mg_build(MGSolver &mg, int kind, int n_levels, void **meshes, void **ops, const orc_mg_params *prm) {
  mg.L = n_levels - 1;
  // a single level is its own coarse solve: Chebyshev solver mode keeps the
  // preconditioner linear (an inner CG would not)
  mg.coarse_kind = n_levels == 1 ? 0 : prm->coarse_kind;
  mg.meshes.resize(n_levels);
  mg.ops.resize(n_levels);
  mg.cheb.resize(n_levels);
  mg.sol.resize(n_levels); mg.def.resize(n_levels); mg.tmp.resize(n_levels);
  for (int l = 0; l < n_levels; ++l) {
    mg.meshes[l] = (Mesh *)meshes[l];
    mg.ops[l].kind = kind; mg.ops[l].op = ops[l]; mg.ops[l].m = mg.meshes[l];
    mg.ops[l].diagonal();
    size_t n = mg.ops[l].n();
    mg.sol[l].assign(n, 0.0); mg.def[l].assign(n, 0.0); mg.tmp[l].assign(n, 0.0);
  }
  for (int l = 0; l < n_levels; ++l) {
    if (l == 0 && mg.coarse_kind == 0)
      mg.cheb[0].setup(mg.ops[0], 0, 1e-3, (int)mg.ops[0].n());
    else
      mg.cheb[l].setup(mg.ops[l], prm->cheb_degree, prm->smoothing_range, prm->eig_cg_iterations);
  }
}

void orc_mg_vmult(int kind, int n_levels, void **meshes, void **ops, const orc_mg_params *prm, const double *src,
                  double *dst, double *eig /* 3 per level: lmax, lmin, degree */) {
  MGSolver mg;
  mg_build(mg, kind, n_levels, meshes, ops, prm);
  const size_t n = mg.ops[mg.L].n();
  vector<double> s(src, src + n), d(n);
  mg.precondition(d, s);
  std::copy(d.begin(), d.end(), dst);
  if (eig)
    for (int l = 0; l < n_levels; ++l) {
      eig[3 * l] = mg.cheb[l].lmax;
      eig[3 * l + 1] = mg.cheb[l].lmin;
      eig[3 * l + 2] = mg.cheb[l].degree;
    }
}

int orc_mg_cg_solve(int kind, int n_levels, void **meshes, void **ops, const orc_mg_params *prm, const double *bp,
                    double *xp, double *final_rel) {
  MGSolver mg;
  mg_build(mg, kind, n_levels, meshes, ops, prm);
  const LevelOp &A = mg.ops[mg.L];
  const size_t n = A.n();
  vector<double> b(bp, bp + n), x(xp, xp + n), r(n), d(n), p(n), h(n);
  A.vmult(h, x);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - h[i];
  const double bnorm = std::sqrt(dot(b, b));
  const double tol = prm->tol_rel * bnorm;
  double res = std::sqrt(dot(r, r));
  int it = 0;
  if (res > tol) {
    mg.precondition(d, r);
    double gh = dot(r, d);
    p = d;
    for (it = 1; it <= prm->max_iter; ++it) {
      A.vmult(h, p);
      double alpha = gh / dot(p, h);
      for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * h[i]; }
      res = std::sqrt(dot(r, r));
      if (res <= tol) break;
      mg.precondition(d, r);
      double ghn = dot(r, d);
      double beta = ghn / gh;
      gh = ghn;
      for (size_t i = 0; i < n; ++i) p[i] = d[i] + beta * p[i];
    }
  }
  std::copy(x.begin(), x.end(), xp);
  if (final_rel) *final_rel = bnorm > 0 ? res / bnorm : 0.0;
  return (res <= tol) ? it : -1;
}

// One Newmark time step with full Newton-Raphson and MG-CG linear solves,
// restating Solid::solve_nonlinear_timestep (mf_elasticity.h:1400-1491),
// check_convergence (:1357-1395), update_acceleration/velocity (:1600-1620)
// and update_matrix_free (:1289-1353: every level re-linearised at the
// interpolated displacement, then the GMG set-up).
// nm = {beta, gamma, dt, bx, by, bz, tol_f, tol_u, tol_f_abs, tol_u_abs, tol_lin}
int orc_newmark_step(int n_levels, void **meshes, void **ops, const orc_mg_params *prm, const double *nm,
                     int max_newton, int min_newton, double *u, double *v_old, double *a_old, int *lin_its) {
  return orc_newmark_step_iface(n_levels, meshes, ops, prm, nm, max_newton, min_newton, u, v_old, a_old, lin_its, 0,
                                nullptr, nullptr);
}

int orc_newmark_step_iface(int n_levels, void **meshes, void **ops, const orc_mg_params *prm, const double *nm,
                           int max_newton, int min_newton, double *u, double *v_old, double *a_old, int *lin_its,
                           int64_t n_faces, const int32_t *faces, const double *traction) {
  const double beta = nm[0], gamma = nm[1], dt = nm[2];
  const double a1 = 1.0 / (beta * dt * dt), a2 = 1.0 / (beta * dt), a3 = (1 - 2 * beta) / (2 * beta);
  const double a4 = gamma / (beta * dt), a5 = 1 - gamma / beta, a6 = (1 - gamma / (2 * beta)) * dt;
  const double body[3] = {nm[3], nm[4], nm[5]};
  const double tol_f = nm[6], tol_u = nm[7], tol_f_abs = nm[8], tol_u_abs = nm[9], tol_lin = nm[10];
  Mesh &fm = *(Mesh *)meshes[n_levels - 1];
  const size_t n = (size_t)fm.n_dofs;
  vector<double> du(n, 0.0), delta(n, 0.0), acc(n), R(n);
  vector<vector<double>> ul(n_levels);
  double res0 = 1, upd = 1, upd0 = 1;
  int it = 0, total_lin = 0;
  bool converged = false;
  for (; it < max_newton; ++it) {
    for (size_t i = 0; i < n; ++i) acc[i] = a1 * delta[i] - a2 * v_old[i] - a3 * a_old[i];
    orc_solid_residual(ops[n_levels - 1], u, acc.data(), body, R.data());
    // mf_elasticity.h:1849-1897: the interface traction, pulled back at u
    if (n_faces) orc_interface_traction(&fm, n_faces, faces, u, traction, R.data());
    const double res = std::sqrt(dot(R, R));
    if (it == 0) res0 = res;
    if (it >= min_newton) {
      const double rn = res0 != 0 ? res / res0 : res, un = upd0 != 0 ? upd / upd0 : upd;
      if (res <= tol_f_abs || upd <= tol_u_abs || (rn <= tol_f && un <= tol_u)) {
        converged = true;
        break;
      }
    }
    ul[n_levels - 1].assign(u, u + n);
    for (int l = n_levels - 1; l >= 1; --l) {
      ul[l - 1].assign((size_t)((Mesh *)meshes[l - 1])->n_dofs, 0.0);
      mg_interpolate(*(Mesh *)meshes[l - 1], *(Mesh *)meshes[l], ul[l].data(), ul[l - 1].data());
    }
    for (int l = 0; l < n_levels; ++l)
      if (orc_solid_cache(ops[l], ul[l].data()) != 0) return -2;
    orc_mg_params lp = *prm;
    lp.tol_rel = tol_lin;
    std::fill(du.begin(), du.end(), 0.0);
    double rel;
    const int li = orc_mg_cg_solve(0, n_levels, meshes, ops, &lp, R.data(), du.data(), &rel);
    if (li < 0) return -3;
    total_lin += li;
    upd = std::sqrt(dot(du, du));
    if (it == 0) upd0 = upd;
    for (size_t i = 0; i < n; ++i) {
      delta[i] += du[i];
      u[i] += du[i];
    }
  }
  if (lin_its) *lin_its = total_lin;
  if (!converged) return -1;
  for (size_t i = 0; i < n; ++i) {
    const double vn = a4 * delta[i] + a5 * v_old[i] + a6 * a_old[i];
    v_old[i] = vn;
    a_old[i] = acc[i];
  }
  return it;
}

int orc_cg_solve(int kind, void *op, int precond, double omega, const double *bp, double *xp, double tol_rel,
                 int max_iter, double *final_rel) {
  LevelOp A;
  A.kind = kind; A.op = op;
  A.m = (Mesh *)orc_operator_mesh(kind, op);
  const size_t n = A.n();
  if (precond) {
    A.diagonal();
    for (auto &v : A.inv_diag) v *= omega;
  } else {
    A.inv_diag.assign(n, 1.0);
  }
  vector<double> b(bp, bp + n), x(xp, xp + n), r(n), d(n), p(n), h(n);
  A.vmult(h, x);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - h[i];
  const double bnorm = std::sqrt(dot(b, b)), tol = tol_rel * bnorm;
  double res = std::sqrt(dot(r, r));
  int it = 0;
  if (res > tol) {
    for (size_t i = 0; i < n; ++i) d[i] = A.inv_diag[i] * r[i];
    double gh = dot(r, d);
    p = d;
    for (it = 1; it <= max_iter; ++it) {
      A.vmult(h, p);
      double alpha = gh / dot(p, h);
      for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * h[i]; }
      res = std::sqrt(dot(r, r));
      if (res <= tol) break;
      for (size_t i = 0; i < n; ++i) d[i] = A.inv_diag[i] * r[i];
      double ghn = dot(r, d);
      double beta = ghn / gh;
      gh = ghn;
      for (size_t i = 0; i < n; ++i) p[i] = d[i] + beta * p[i];
    }
  }
  std::copy(x.begin(), x.end(), xp);
  if (final_rel) *final_rel = bnorm > 0 ? res / bnorm : 0.0;
  return (res <= tol) ? it : -1;
}

}  // extern "C"
