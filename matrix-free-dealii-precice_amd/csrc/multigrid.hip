// Multigrid transfer objects, Chebyshev smoothing, the V-cycle and CG on the
// device.  Host code orchestrates; every vector operation is a HIP kernel on
// the operators' stream.  Solver settings restate mf_elasticity.h:1140-1279,
// 1981-2070 and heat_transfer.h:757-812; the semantics of the deal.II solver
// classes they configure (PreconditionChebyshev, SolverCG, Multigrid) follow
// the library's documented behaviour.
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "operator.hpp"
#include "solver.hpp"
#include "transfer_kernels.hpp"
#include "vector_kernels.hpp"

namespace mfg {

// ------------------------------------------------------------- vectors ---
static int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

Reducer::Reducer() {
  own.allocate((kRedBlocks + kRedScalars) * sizeof(double));
  partials = own.as<double>();
  scal = partials + kRedBlocks;
}

Reducer::Reducer(double *p, double *sc) : partials(p), scal(sc) {}

template <typename T>
static void dot_partials(const void *a, const void *b, int64_t n, double *partials, hipStream_t s) {
  dot_partial_kernel<T><<<kRedBlocks, kRedThreads, 0, s>>>((const T *)a, (const T *)b, n, partials);
}

void Reducer::epilogue(int mode, int slot, double tol, double rel, bool record, hipStream_t s) {
  sum_partials_epilogue_kernel<<<1, kRedThreads, 0, s>>>(partials, kRedBlocks, scal, mode, slot, tol, rel,
                                                         record ? hist_buf.as<double>() : nullptr, hist_cap);
  MFG_HIP(hipGetLastError());
}

double Reducer::dot(int prec, const void *a, const void *b, int64_t n, hipStream_t s) {
  if (prec == MFG_FP64) dot_partials<double>(a, b, n, partials, s);
  else dot_partials<float>(a, b, n, partials, s);
  epilogue(kRedSet, kSlotHost, 0, 0, false, s);
  double h = 0;
  MFG_HIP(hipMemcpyAsync(&h, scal + kSlotHost, sizeof(double), hipMemcpyDeviceToHost, s));
  MFG_HIP(hipStreamSynchronize(s));
  return h;
}

void Reducer::dot_dev(int prec, const void *a, const void *b, int64_t n, int mode, int slot, hipStream_t s,
                      double tol, double rel, bool record) {
  if (prec == MFG_FP64) dot_partials<double>(a, b, n, partials, s);
  else dot_partials<float>(a, b, n, partials, s);
  epilogue(mode, slot, tol, rel, record, s);
}

void Reducer::cg_update_dev(int prec, void *x, void *r, const void *p, const void *h, int64_t n, int mode,
                            hipStream_t s) {
  if (prec == MFG_FP64)
    cg_update_dev_kernel<double><<<kRedBlocks, kRedThreads, 0, s>>>((double *)x, (double *)r, (const double *)p,
                                                                    (const double *)h, scal, n, partials);
  else
    cg_update_dev_kernel<float><<<kRedBlocks, kRedThreads, 0, s>>>((float *)x, (float *)r, (const float *)p,
                                                                   (const float *)h, scal, n, partials);
  epilogue(mode, kSlotRR, 0, 0, false, s);
}

void Reducer::jacobi_dot_dev(int prec, void *z, const void *invD, const void *r, double omega, int64_t n, int mode,
                             hipStream_t s, bool record) {
  if (prec == MFG_FP64)
    jacobi_dot_kernel<double><<<kRedBlocks, kRedThreads, 0, s>>>((double *)z, (const double *)invD,
                                                                 (const double *)r, omega, n, partials);
  else
    jacobi_dot_kernel<float><<<kRedBlocks, kRedThreads, 0, s>>>((float *)z, (const float *)invD, (const float *)r,
                                                                (float)omega, n, partials);
  epilogue(mode, 0, 0, 0, record, s);
}

void Reducer::repeat_epilogue(int mode, int slot, hipStream_t s, double tol, double rel, bool record) {
  epilogue(mode, slot, tol, rel, record, s);
}

void Reducer::read(double *out, hipStream_t s) {
  MFG_HIP(hipMemcpyAsync(out, scal, kRedScalars * sizeof(double), hipMemcpyDeviceToHost, s));
  MFG_HIP(hipStreamSynchronize(s));
}

void Reducer::ensure_hist(int cap) {
  if (cap <= hist_cap) return;
  hist_buf.allocate((size_t)2 * cap * sizeof(double));
  hist_cap = cap;
}

void Reducer::read_hist(const double *sc, std::vector<double> &alphas, std::vector<double> &betas, hipStream_t s) {
  const int na = (int)sc[kSlotNAlpha], nb = (int)sc[kSlotNBeta];
  alphas.assign(na, 0.0);
  betas.assign(nb, 0.0);
  if (na) MFG_HIP(hipMemcpyAsync(alphas.data(), hist_buf.ptr, na * sizeof(double), hipMemcpyDeviceToHost, s));
  if (nb)
    MFG_HIP(hipMemcpyAsync(betas.data(), hist_buf.as<double>() + hist_cap, nb * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  MFG_HIP(hipStreamSynchronize(s));
}

void vec_fill(int prec, void *x, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  if (prec == MFG_FP64) fill_kernel<double><<<grid_for(n), 256, 0, s>>>((double *)x, n, v);
  else fill_kernel<float><<<grid_for(n), 256, 0, s>>>((float *)x, n, (float)v);
  MFG_HIP(hipGetLastError());
}

void vec_convert(int prec_dst, void *dst, int prec_src, const void *src, int64_t n, hipStream_t s) {
  if (prec_dst == prec_src) {
    MFG_HIP(hipMemcpyAsync(dst, src, n * (prec_dst == MFG_FP64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
    return;
  }
  if (prec_dst == MFG_FP64)
    convert_kernel<double, float><<<grid_for(n), 256, 0, s>>>((double *)dst, (const float *)src, n);
  else
    convert_kernel<float, double><<<grid_for(n), 256, 0, s>>>((float *)dst, (const double *)src, n);
  MFG_HIP(hipGetLastError());
}

#define MFG_PREC_DISPATCH(prec, KERNEL, ...)                                                   \
  do {                                                                                         \
    if ((prec) == MFG_FP64) KERNEL<double><<<grid_for(n), 256, 0, s>>>(__VA_ARGS__);            \
    else KERNEL<float><<<grid_for(n), 256, 0, s>>>(__VA_ARGS__);                                \
    MFG_HIP(hipGetLastError());                                                                \
  } while (0)

template <typename T>
static void xpby_t(void *p, const void *d, double beta, int64_t n, hipStream_t s) {
  xpby_kernel<T><<<grid_for(n), 256, 0, s>>>((T *)p, (const T *)d, beta, n);
}
void vec_xpby(int prec, void *p, const void *d, double beta, int64_t n, hipStream_t s) {
  if (prec == MFG_FP64) xpby_t<double>(p, d, beta, n, s);
  else xpby_t<float>(p, d, beta, n, s);
  MFG_HIP(hipGetLastError());
}
void Reducer::xpby_dev(int prec, void *p, const void *z, int64_t n, hipStream_t s) {
  if (prec == MFG_FP64) xpby_dev_kernel<double><<<grid_for(n), 256, 0, s>>>((double *)p, (const double *)z, scal, n);
  else xpby_dev_kernel<float><<<grid_for(n), 256, 0, s>>>((float *)p, (const float *)z, scal, n);
  MFG_HIP(hipGetLastError());
}
template <typename T>
static void residual_t(void *t, const void *b, int64_t n, hipStream_t s) {
  residual_kernel<T><<<grid_for(n), 256, 0, s>>>((T *)t, (const T *)b, n);
}
void vec_residual(int prec, void *t, const void *b, int64_t n, hipStream_t s) {
  if (prec == MFG_FP64) residual_t<double>(t, b, n, s);
  else residual_t<float>(t, b, n, s);
  MFG_HIP(hipGetLastError());
}
template <typename T>
static void add_scaled_t(void *y, const void *a, double sc, const void *b, int64_t n, hipStream_t s) {
  add_scaled_kernel<T><<<grid_for(n), 256, 0, s>>>((T *)y, (const T *)a, sc, (const T *)b, n);
}
void vec_add_scaled(int prec, void *y, const void *a, double sc, const void *b, int64_t n, hipStream_t s) {
  if (prec == MFG_FP64) add_scaled_t<double>(y, a, sc, b, n, s);
  else add_scaled_t<float>(y, a, sc, b, n, s);
  MFG_HIP(hipGetLastError());
}

template <typename T>
__global__ void eig_initial_kernel(T *v, int64_t n) {
  MFG_GRID_STRIDE(i, n) v[i] = (i % 11 == 0) ? T(0) : T(1);
}

// ---------------------------------------------------------- transfers ---
template <int P, int NC, typename CN, typename FN>
static void transfer_launch(TransferData &t, int kind, void *fine, const void *coarse_in, void *coarse_out,
                            const void *fine_in) {
  const OperatorData &C = *t.coarse;
  TransferArgs<CN, FN> a;
  a.coarse_nodes = C.cell_nodes.as<uint32_t>();
  a.lattice = t.lattice.as<uint32_t>();
  a.coarse_in = (const CN *)coarse_in;
  a.coarse_out = (CN *)coarse_out;
  a.fine_in = (const FN *)fine_in;
  a.fine_out = (FN *)fine;
  hipStream_t s = C.stream;
  if (kind == 0) {  // prolongate_add: every fine node written once by its owner
    a.E = t.E.as<double>();
    a.cell_begin = 0;
    a.cell_end = C.n_cells;
    prolongate_add_kernel<P, NC, CN, FN><<<C.n_cells, 256, 0, s>>>(a);
  } else if (kind == 1) {  // restrict_add: coarse colours share no node
    a.E = t.E.as<double>();
    for (size_t col = 0; col + 1 < C.colour_start.size(); ++col) {
      a.cell_begin = C.colour_start[col];
      a.cell_end = C.colour_start[col + 1];
      if (a.cell_end > a.cell_begin)
        restrict_add_kernel<P, NC, CN, FN><<<a.cell_end - a.cell_begin, 256, 0, s>>>(a);
    }
  } else {
    a.E = t.I.as<double>();
    a.cell_begin = 0;
    a.cell_end = C.n_cells;
    interpolate_kernel<P, NC, CN, FN><<<C.n_cells, 256, 0, s>>>(a);
  }
  MFG_HIP(hipGetLastError());
}

template <int P, int NC>
static void transfer_prec(TransferData &t, int kind, void *fine, const void *ci, void *co, const void *fi) {
  const int cp = t.coarse->precision, fp = t.fine->precision;
  if (cp == MFG_FP64 && fp == MFG_FP64) transfer_launch<P, NC, double, double>(t, kind, fine, ci, co, fi);
  else if (cp == MFG_FP64) transfer_launch<P, NC, double, float>(t, kind, fine, ci, co, fi);
  else if (fp == MFG_FP64) transfer_launch<P, NC, float, double>(t, kind, fine, ci, co, fi);
  else transfer_launch<P, NC, float, float>(t, kind, fine, ci, co, fi);
}

template <int NC>
static void transfer_deg(TransferData &t, int kind, void *fine, const void *ci, void *co, const void *fi) {
  switch (t.p) {
    case 1: transfer_prec<1, NC>(t, kind, fine, ci, co, fi); break;
    case 2: transfer_prec<2, NC>(t, kind, fine, ci, co, fi); break;
    case 3: transfer_prec<3, NC>(t, kind, fine, ci, co, fi); break;
    case 4: transfer_prec<4, NC>(t, kind, fine, ci, co, fi); break;
    case 5: transfer_prec<5, NC>(t, kind, fine, ci, co, fi); break;
    case 6: transfer_prec<6, NC>(t, kind, fine, ci, co, fi); break;
    default: throw Error(MFG_ERR_ARGUMENT, "unsupported degree");
  }
}

void TransferData::run(int kind, void *fine, const void *coarse_in, void *coarse_out, const void *fine_in) {
  if (nc == 3) transfer_deg<3>(*this, kind, fine, coarse_in, coarse_out, fine_in);
  else transfer_deg<1>(*this, kind, fine, coarse_in, coarse_out, fine_in);
}
void TransferData::prolongate_add(void *fine, const void *coarse) { run(0, fine, coarse, nullptr, nullptr); }
void TransferData::restrict_add(void *coarse, const void *fine) { run(1, nullptr, nullptr, coarse, fine); }
void TransferData::interpolate(void *coarse, const void *fine) { run(2, nullptr, nullptr, coarse, fine); }

void TransferData::build(OperatorData *c, OperatorData *f, const int32_t *children) {
  MFG_REQUIRE(c && f && children, MFG_ERR_ARGUMENT, "null argument");
  MFG_REQUIRE(c->p == f->p && c->nc == f->nc && c->kind == f->kind, MFG_ERR_ARGUMENT,
              "transfer levels must share degree, components and kind");
  MFG_REQUIRE((int64_t)f->n_cells == 8 * (int64_t)c->n_cells, MFG_ERR_ARGUMENT,
              "fine level must have 8 children per coarse cell");
  coarse = c;
  fine = f;
  p = c->p;
  nc = c->nc;
  const int N = p + 1, M = 2 * p + 1, N3 = N * N * N, M3 = M * M * M;
  std::vector<uint32_t> lat((size_t)c->n_cells * M3);
  std::vector<int32_t> owner((size_t)f->n_nodes, -1);
  for (int cc = 0; cc < c->n_cells; ++cc) {
    const int64_t oc = c->perm[cc];
    for (int K = 0; K < M; ++K)
      for (int J = 0; J < M; ++J)
        for (int I = 0; I < M; ++I) {
          const int dx = std::min(I / p, 1), dy = std::min(J / p, 1), dz = std::min(K / p, 1);
          const int32_t fc = children[oc * 8 + dx + 2 * dy + 4 * dz];
          MFG_REQUIRE(fc >= 0 && fc < f->n_cells, MFG_ERR_ARGUMENT, "child cell index out of range");
          const int li = I - dx * p, lj = J - dy * p, lk = K - dz * p;
          const int32_t node = f->host_cell_nodes[(int64_t)fc * N3 + li + N * (lj + N * lk)];
          uint32_t e = (uint32_t)node | ((uint32_t)f->host_cmask[node] << kConstraintShift);
          if (owner[node] < 0) {
            owner[node] = cc;
            e |= kOwnerBit;
          }
          lat[(size_t)cc * M3 + I + M * (J + M * K)] = e;
        }
  }
  for (int64_t nd = 0; nd < f->n_nodes; ++nd)
    MFG_REQUIRE(owner[nd] >= 0, MFG_ERR_ARGUMENT, "fine node not covered by any coarse cell");
  upload(lattice, lat.data(), lat.size(), c->stream);
  // 1D embedding E[I][a] = phi_a(eta_I) and interpolation Imat[a][I]
  const std::vector<double> z = gauss_lobatto_points(p);
  std::vector<double> Em((size_t)M * N), Im((size_t)N * M, 0.0), val(N), der(N);
  for (int I = 0; I < M; ++I) {
    const int dx = std::min(I / p, 1);
    const double eta = 0.5 * (dx + z[I - dx * p]);
    lagrange_basis(z, eta, val.data(), der.data());
    for (int a = 0; a < N; ++a) Em[(size_t)I * N + a] = val[a];
  }
  for (int a = 0; a < N; ++a) {
    const int dx = z[a] <= 0.5 ? 0 : 1;
    const double s = 2.0 * z[a] - dx;
    lagrange_basis(z, s, val.data(), der.data());
    for (int l = 0; l < N; ++l) Im[(size_t)a * M + dx * p + l] = val[l];
  }
  upload(E, Em.data(), Em.size(), c->stream);
  upload(I, Im.data(), Im.size(), c->stream);
}

// ---------------------------------------------------- operator helpers ---
static void op_vmult(const OperatorData &op, void *dst, const void *src) { op.apply(src, dst, false); }

// Eigenvalues of the Lanczos tridiagonal matrix (bisection).
static void tridiag_eigs(const std::vector<double> &dg, const std::vector<double> &off, double &lmin, double &lmax) {
  const int n = (int)dg.size();
  if (n == 0) { lmin = lmax = 1.0; return; }
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::abs(off[i - 1]) : 0.0) + (i < n - 1 ? std::abs(off[i]) : 0.0);
    lo = std::min(lo, dg[i] - r);
    hi = std::max(hi, dg[i] + r);
  }
  auto below = [&](double x) {
    int cnt = 0;
    double q = dg[0] - x;
    if (q < 0) ++cnt;
    for (int i = 1; i < n; ++i) {
      const double qq = q == 0.0 ? 1e-300 : q;
      q = dg[i] - x - off[i - 1] * off[i - 1] / qq;
      if (q < 0) ++cnt;
    }
    return cnt;
  };
  auto kth = [&](int k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      const double m = 0.5 * (a + b);
      if (below(m) > k) b = m;
      else a = m;
    }
    return 0.5 * (a + b);
  };
  lmin = kth(0);
  lmax = kth(n - 1);
}

void lanczos_estimate(const std::vector<double> &alphas, const std::vector<double> &betas, double &lmin,
                      double &lmax) {
  const int k = (int)alphas.size();
  std::vector<double> dg(k), off(k > 0 ? k - 1 : 0);
  for (int i = 0; i < k; ++i) {
    dg[i] = 1.0 / alphas[i] + (i > 0 ? betas[i - 1] / alphas[i - 1] : 0.0);
    if (i < k - 1) off[i] = std::sqrt(betas[i]) / alphas[i];
  }
  tridiag_eigs(dg, off, lmin, lmax);
}

// ------------------------------------------------------------ levels ---
void MGLevel::allocate() {
  const size_t b = (size_t)op->n_dofs * op->elem_size();
  for (DeviceBuffer *v : {&sol, &def, &tmp, &d, &t, &cg_r, &cg_p, &cg_h, &cg_z}) v->allocate(b);
}

// CG with the inverse-diagonal preconditioner on A x = v0, recording the
// Lanczos coefficients (PreconditionChebyshev's eigenvalue estimation).
void MGLevel::estimate_eigenvalues(int n_it, Reducer &red) {
  const OperatorData &A = *op;
  const int64_t n = A.n_dofs;
  const int pr = A.precision;
  hipStream_t s = A.stream;
  // b = initial vector (tmp), x = 0 (sol), r = b
  if (pr == MFG_FP64) eig_initial_kernel<double><<<grid_for(n), 256, 0, s>>>(tmp.as<double>(), n);
  else eig_initial_kernel<float><<<grid_for(n), 256, 0, s>>>(tmp.as<float>(), n);
  MFG_HIP(hipGetLastError());
  if (A.n_constrained) {
    if (pr == MFG_FP64)
      set_entries_kernel<double><<<blocks_for(A.n_constrained, 256), 256, 0, s>>>(
          tmp.as<double>(), A.constrained.as<int32_t>(), A.n_constrained, 0.0);
    else
      set_entries_kernel<float><<<blocks_for(A.n_constrained, 256), 256, 0, s>>>(
          tmp.as<float>(), A.constrained.as<int32_t>(), A.n_constrained, 0.0f);
    MFG_HIP(hipGetLastError());
  }
  vec_fill(pr, sol.ptr, n, 0.0, s);
  vec_convert(pr, cg_r.ptr, pr, tmp.ptr, n, s);
  // The iterations run without host round trips: the step lengths live in
  // the reducer's scalar block and history; the stop flag (breakdown, or
  // ||r|| <= 1e-15 ||r_0||) turns the remaining iterations into no-ops.  Long
  // runs (the coarse level's m() iterations) look at the flag every 16
  // iterations to skip the no-ops.
  red.ensure_hist(std::max(n_it, 1));
  red.jacobi_dot_dev(pr, cg_z.ptr, A.inv_diag.ptr, cg_r.ptr, 1.0, n, kRedStart, s);
  vec_convert(pr, cg_p.ptr, pr, cg_z.ptr, n, s);
  red.dot_dev(pr, cg_r.ptr, cg_r.ptr, n, kRedRes0, 0, s, 0.0, 1e-15);
  double sc[kRedScalars];
  for (int it = 0; it < n_it; ++it) {
    op_vmult(A, cg_h.ptr, cg_p.ptr);
    red.dot_dev(pr, cg_p.ptr, cg_h.ptr, n, kRedAlpha, 0, s, 0, 0, true);
    red.cg_update_dev(pr, sol.ptr, cg_r.ptr, cg_p.ptr, cg_h.ptr, n, kRedRRStop, s);
    if (it + 1 == n_it) break;
    red.jacobi_dot_dev(pr, cg_z.ptr, A.inv_diag.ptr, cg_r.ptr, 1.0, n, kRedBeta, s, true);
    red.xpby_dev(pr, cg_p.ptr, cg_z.ptr, n, s);
    if (n_it > 32 && (it + 1) % 16 == 0) {
      red.read(sc, s);
      if (sc[kSlotStop] != 0) break;
    }
  }
  red.read(sc, s);
  std::vector<double> alphas, betas;
  red.read_hist(sc, alphas, betas, s);
  if (alphas.size() > betas.size() + 1) alphas.resize(betas.size() + 1);
  lanczos_estimate(alphas, betas, lmin, lmax);
}

// [1.2 lmax / range, 1.2 lmax]; range < 1: solver mode with the degree from
// the estimate (mf_elasticity.h:1149-1177).
void MGLevel::setup_chebyshev(int deg, double range, int eig_it, Reducer &red) {
  estimate_eigenvalues(eig_it, red);
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

// x = p(A) b (zero initial guess: degree-1 operator applications) or
// x += p(A)(b - A x) (degree applications)
void MGLevel::chebyshev(void *x, const void *b, bool zero_guess) {
  const OperatorData &A = *op;
  const int64_t n = A.n_dofs;
  hipStream_t s = A.stream;
  const double inv_theta = 1.0 / theta;
  if (zero_guess) {
    if (A.precision == MFG_FP64)
      cheb_first_kernel<double><<<grid_for(n), 256, 0, s>>>((double *)x, d.as<double>(), (const double *)b,
                                                            A.inv_diag.as<double>(), inv_theta, n);
    else
      cheb_first_kernel<float><<<grid_for(n), 256, 0, s>>>((float *)x, d.as<float>(), (const float *)b,
                                                           A.inv_diag.as<float>(), inv_theta, n);
  } else {
    op_vmult(A, t.ptr, x);
    if (A.precision == MFG_FP64)
      cheb_first_step_kernel<double><<<grid_for(n), 256, 0, s>>>((double *)x, d.as<double>(), (const double *)b,
                                                                 t.as<double>(), A.inv_diag.as<double>(), inv_theta, n);
    else
      cheb_first_step_kernel<float><<<grid_for(n), 256, 0, s>>>((float *)x, d.as<float>(), (const float *)b,
                                                                t.as<float>(), A.inv_diag.as<float>(), inv_theta, n);
  }
  MFG_HIP(hipGetLastError());
  if (degree < 2 || std::abs(delta) < 1e-40) return;
  double rhok = delta / theta;
  const double sigma = theta / delta;
  for (int k = 0; k < degree - 1; ++k) {
    op_vmult(A, t.ptr, x);
    const double rhokp = 1.0 / (2.0 * sigma - rhok);
    const double f1 = rhokp * rhok, f2 = 2.0 * rhokp / delta;
    rhok = rhokp;
    if (A.precision == MFG_FP64)
      cheb_update_kernel<double><<<grid_for(n), 256, 0, s>>>((double *)x, d.as<double>(), (const double *)b,
                                                             t.as<double>(), A.inv_diag.as<double>(), f1, f2, n);
    else
      cheb_update_kernel<float><<<grid_for(n), 256, 0, s>>>((float *)x, d.as<float>(), (const float *)b,
                                                            t.as<float>(), A.inv_diag.as<float>(), f1, f2, n);
    MFG_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------------- the MG ---
void MGData::setup() {
  drop_graphs();  // new smoother coefficients
  for (auto &L : levels) {
    if (!L.op->has_diagonal) L.op->compute_diagonal();
    if (L.sol.bytes != (size_t)L.op->n_dofs * L.op->elem_size()) L.allocate();
  }
  for (size_t l = 0; l < levels.size(); ++l) {
    MGLevel &L = levels[l];
    if (l == 0) {
      if (coarse_kind() == 0) {
        int64_t it = L.op->n_dofs;
        if (prm.coarse_eig_iterations > 0) it = std::min<int64_t>(it, prm.coarse_eig_iterations);
        L.setup_chebyshev(0, prm.coarse_range, (int)it, red);
      }
      else L.setup_chebyshev(prm.cheb_degree, prm.smoothing_range, prm.eig_cg_iterations, red);
    } else {
      L.setup_chebyshev(prm.cheb_degree, prm.smoothing_range, prm.eig_cg_iterations, red);
    }
  }
  ready = true;
}

// MGCoarseGridIterativeSolver: CG to relative coarse_tolerance (absolute
// 1e-10), max(m, 100) iterations, preconditioned by the level-0 Chebyshev.
void MGData::coarse_solve() {
  MGLevel &L = levels[0];
  if (coarse_kind() == 0) {
    L.chebyshev(L.sol.ptr, L.def.ptr, true);
    return;
  }
  const OperatorData &A = *L.op;
  const int64_t n = A.n_dofs;
  const int pr = A.precision;
  hipStream_t s = A.stream;
  vec_fill(pr, L.sol.ptr, n, 0.0, s);
  vec_convert(pr, L.cg_r.ptr, pr, L.def.ptr, n, s);
  // The step lengths and the stop test (||r|| <= max(1e-10, coarse_tolerance
  // ||r_0||), or a breakdown) stay on the device: once the stop flag is set,
  // the remaining iterations change neither sol nor r.  The host reads the
  // scalar block after a run of iterations as long as the previous coarse
  // solve took, then after runs of 4, so a coarse solve usually costs one
  // read-back instead of one per iteration.
  L.chebyshev(L.cg_z.ptr, L.cg_r.ptr, true);
  red.dot_dev(pr, L.cg_r.ptr, L.cg_z.ptr, n, kRedStart, 0, s);
  vec_convert(pr, L.cg_p.ptr, pr, L.cg_z.ptr, n, s);
  red.dot_dev(pr, L.cg_r.ptr, L.cg_r.ptr, n, kRedRes0, 0, s, 1e-10, prm.coarse_tolerance);
  const int maxit = std::max<int>((int)n, 100);
  double sc[kRedScalars];
  int it = 0, run = std::max(1, coarse_iterations);
  for (;;) {
    for (const int end = std::min(maxit, it + run); it < end; ++it) {
      if (it > 0) {
        L.chebyshev(L.cg_z.ptr, L.cg_r.ptr, true);
        red.dot_dev(pr, L.cg_r.ptr, L.cg_z.ptr, n, kRedBeta, 0, s);
        red.xpby_dev(pr, L.cg_p.ptr, L.cg_z.ptr, n, s);
      }
      op_vmult(A, L.cg_h.ptr, L.cg_p.ptr);
      red.dot_dev(pr, L.cg_p.ptr, L.cg_h.ptr, n, kRedAlpha, 0, s, 0, 0, true);
      red.cg_update_dev(pr, L.sol.ptr, L.cg_r.ptr, L.cg_p.ptr, L.cg_h.ptr, n, kRedRRStop, s);
    }
    red.read(sc, s);
    if (sc[kSlotStop] != 0 || it >= maxit) break;
    run = 4;
  }
  // active iterations (counted while the step-length history has room)
  coarse_iterations = (int)sc[kSlotNAlpha];
}

void MGData::down_sweep() {
  for (int l = (int)levels.size() - 1; l >= 1; --l) {
    MGLevel &L = levels[l];
    MGLevel &C = levels[l - 1];
    const OperatorData &A = *L.op;
    L.chebyshev(L.sol.ptr, L.def.ptr, true);  // pre-smoothing from a zero guess
    op_vmult(A, L.tmp.ptr, L.sol.ptr);
    vec_residual(A.precision, L.tmp.ptr, L.def.ptr, A.n_dofs, A.stream);
    vec_fill(C.op->precision, C.def.ptr, C.op->n_dofs, 0.0, A.stream);
    transfers[l - 1]->restrict_add(C.def.ptr, L.tmp.ptr);
  }
}

void MGData::up_sweep() {
  for (int l = 1; l < (int)levels.size(); ++l) {
    MGLevel &L = levels[l];
    MGLevel &C = levels[l - 1];
    transfers[l - 1]->prolongate_add(L.sol.ptr, C.sol.ptr);
    L.chebyshev(L.sol.ptr, L.def.ptr, false);  // post-smoothing
  }
}

bool MGData::graphs_allowed() const {
  if (!use_graph || levels.size() < 2) return false;
  for (const MGLevel &L : levels)
    if (L.op->timing) return false;  // the timing events belong to eager launches
  return true;
}

void MGData::drop_graphs() {
  if (graph_down) (void)hipGraphExecDestroy(graph_down);
  if (graph_up) (void)hipGraphExecDestroy(graph_up);
  graph_down = graph_up = nullptr;
  graph_versions.clear();
  graph_calls = 0;
  graph_nodes = 0;
}

MGData::~MGData() {
  drop_graphs();
  if (graph_in) (void)hipEventDestroy(graph_in);
  if (graph_out) (void)hipEventDestroy(graph_out);
  if (graph_stream) (void)hipStreamDestroy(graph_stream);
}

static hipGraphExec_t capture_on(hipStream_t s, const std::function<void()> &body, size_t &nodes) {
  hipGraph_t g = nullptr;
  MFG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  MFG_HIP(hipStreamEndCapture(s, &g));
  size_t nn = 0;
  (void)hipGraphGetNodes(g, nullptr, &nn);
  nodes += nn;
  hipGraphExec_t e = nullptr;
  const hipError_t st = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MFG_HIP(st);
  return e;
}

void MGData::run_graphed(hipStream_t caller) {
  if (!graph_stream) {
    MFG_HIP(hipStreamCreateWithFlags(&graph_stream, hipStreamNonBlocking));
    MFG_HIP(hipEventCreateWithFlags(&graph_in, hipEventDisableTiming));
    MFG_HIP(hipEventCreateWithFlags(&graph_out, hipEventDisableTiming));
  }
  std::vector<uint64_t> versions;
  for (const MGLevel &L : levels) versions.push_back(L.op->buffers_version);
  if (versions != graph_versions) {
    if (graph_down) (void)hipGraphExecDestroy(graph_down);
    if (graph_up) (void)hipGraphExecDestroy(graph_up);
    graph_down = graph_up = nullptr;
    graph_nodes = 0;
    graph_versions = versions;
  }
  MFG_HIP(hipEventRecord(graph_in, caller));
  MFG_HIP(hipStreamWaitEvent(graph_stream, graph_in, 0));
  // every level operator (and with it each transfer, which launches on its
  // coarse operator's stream) works on the graph stream meanwhile
  std::vector<hipStream_t> saved;
  for (MGLevel &L : levels) {
    saved.push_back(L.op->stream);
    L.op->stream = graph_stream;
  }
  auto restore = [&] {
    for (size_t l = 0; l < levels.size(); ++l) levels[l].op->stream = saved[l];
  };
  try {
    if (!graph_down) graph_down = capture_on(graph_stream, [&] { down_sweep(); }, graph_nodes);
    MFG_HIP(hipGraphLaunch(graph_down, graph_stream));
    coarse_solve();
    if (!graph_up) graph_up = capture_on(graph_stream, [&] { up_sweep(); }, graph_nodes);
    MFG_HIP(hipGraphLaunch(graph_up, graph_stream));
  } catch (...) {
    restore();
    throw;
  }
  restore();
  MFG_HIP(hipEventRecord(graph_out, graph_stream));
  MFG_HIP(hipStreamWaitEvent(caller, graph_out, 0));
}

// Multigrid::level_v_step from the finest level: the down sweep, the coarse
// solve and the up sweep (the same launches the graphs replay)
void MGData::v_cycle() {
  down_sweep();
  coarse_solve();
  up_sweep();
}

// PreconditionMG::vmult: copy_to_mg, V-cycle, copy_from_mg
void MGData::vmult(void *dst, const void *src, int vec_prec) {
  MFG_REQUIRE(ready, MFG_ERR_STATE, "mfg_mg_setup() not called");
  MGLevel &F = levels.back();
  const int64_t n = F.op->n_dofs;
  hipStream_t caller = F.op->stream;
  vec_convert(F.op->precision, F.def.ptr, vec_prec, src, n, caller);
  if (graphs_allowed() && ++graph_calls >= 2) run_graphed(caller);
  else v_cycle();
  vec_convert(vec_prec, dst, F.op->precision, F.sol.ptr, n, caller);
}

// ------------------------------------------------------------------ CG ---
void cg_solve(const OperatorData &A, int precond, MGData *mg, double omega, const void *b, void *x, double tol_rel,
              int max_iter, mfg_solve_info *info, const CGVectors &v, Reducer &red) {
  const int64_t n = A.n_dofs;
  const int pr = A.precision;
  hipStream_t s = A.stream;
  if (precond == MFG_PRECOND_JACOBI) MFG_REQUIRE(A.has_diagonal, MFG_ERR_STATE, "compute_diagonal() not called");
  if (precond == MFG_PRECOND_MG) MFG_REQUIRE(mg && mg->ready, MFG_ERR_STATE, "multigrid not set up");
  // z = P r and (r, z) through the epilogue `mode`; the Jacobi step and the
  // product in one pass
  auto precond_dot = [&](int mode) {
    if (precond == MFG_PRECOND_JACOBI) {
      red.jacobi_dot_dev(pr, v.z, A.inv_diag.ptr, v.r, omega, n, mode, s);
      return;
    }
    if (precond == MFG_PRECOND_NONE) vec_convert(pr, v.z, pr, v.r, n, s);
    else mg->vmult(v.z, v.r, pr);
    red.dot_dev(pr, v.r, v.z, n, mode, 0, s);
  };
  // The step lengths are computed on the device from reductions that stay in
  // the reducer's scalar block; the host reads the block once per iteration
  // for ||r|| (and the coefficients it keeps for the condition estimate).
  double sc[kRedScalars];
  // r = b - A x
  op_vmult(A, v.h, x);
  vec_add_scaled(pr, v.r, b, -1.0, v.h, n, s);
  red.dot_dev(pr, b, b, n, kRedSet, kSlotB2, s);
  red.dot_dev(pr, v.r, v.r, n, kRedSet, kSlotRR, s);
  red.read(sc, s);
  const double bnorm = std::sqrt(sc[kSlotB2]);
  const double tol = tol_rel * bnorm;
  double res = std::sqrt(sc[kSlotRR]);
  info->initial_residual = res;
  info->iterations = 0;
  info->condition_estimate = 1.0;
  std::vector<double> alphas, betas;
  bool converged = res <= tol;
  if (!converged) {
    precond_dot(kRedStart);
    vec_convert(pr, v.p, pr, v.z, n, s);
    for (int it = 1; it <= max_iter; ++it) {
      op_vmult(A, v.h, v.p);
      red.dot_dev(pr, v.p, v.h, n, kRedAlpha, 0, s);
      red.cg_update_dev(pr, x, v.r, v.p, v.h, n, kRedSet, s);
      red.read(sc, s);
      MFG_REQUIRE(sc[kSlotBreakdown] == 0, MFG_ERR_NUMERIC, "CG breakdown: operator not positive definite");
      if (it > 1) betas.push_back(sc[kSlotBeta]);
      res = std::sqrt(sc[kSlotRR]);
      alphas.push_back(sc[kSlotAlpha]);
      info->iterations = it;
      if (res <= tol) {
        converged = true;
        break;
      }
      precond_dot(kRedBeta);
      red.xpby_dev(pr, v.p, v.z, n, s);
    }
    if (!converged && info->iterations == max_iter) {
      red.read(sc, s);
      betas.push_back(sc[kSlotBeta]);
    }
  }
  info->final_residual = res;
  if (!alphas.empty()) {
    if (alphas.size() > betas.size() + 1) alphas.resize(betas.size() + 1);
    double lmin, lmax;
    lanczos_estimate(alphas, betas, lmin, lmax);
    info->condition_estimate = lmin > 0 ? lmax / lmin : 0.0;
  }
  if (!converged) throw Error(MFG_ERR_NOCONV, "CG did not converge within the iteration limit");
}

double cg_condition_estimate(const OperatorData &A, const void *b, void *x, int n_it, double tol, const CGVectors &v,
                             Reducer &red) {
  const int64_t n = A.n_dofs;
  const int pr = A.precision;
  hipStream_t s = A.stream;
  if (n_it <= 0) return 1.0;
  // x = 0, r = z = p = b; the stop flag ends the run at ||r|| <= tol
  vec_fill(pr, x, n, 0.0, s);
  vec_convert(pr, v.r, pr, b, n, s);
  red.ensure_hist(n_it);
  red.dot_dev(pr, v.r, v.r, n, kRedStart, 0, s);
  vec_convert(pr, v.p, pr, v.r, n, s);
  red.dot_dev(pr, v.r, v.r, n, kRedRes0, 0, s, tol, 0.0);
  for (int it = 0; it < n_it; ++it) {
    op_vmult(A, v.h, v.p);
    red.dot_dev(pr, v.p, v.h, n, kRedAlpha, 0, s, 0, 0, true);
    red.cg_update_dev(pr, x, v.r, v.p, v.h, n, kRedRRStop, s);
    if (it + 1 == n_it) break;
    // (r, r) of the new iterate: the update's partial sums again
    red.repeat_epilogue(kRedBeta, 0, s, 0, 0, true);
    red.xpby_dev(pr, v.p, v.r, n, s);
  }
  double sc[kRedScalars];
  red.read(sc, s);
  std::vector<double> alphas, betas;
  red.read_hist(sc, alphas, betas, s);
  if (alphas.empty()) return 1.0;
  if (alphas.size() > betas.size() + 1) alphas.resize(betas.size() + 1);
  double lmin, lmax;
  lanczos_estimate(alphas, betas, lmin, lmax);
  return lmin > 0 ? lmax / lmin : 0.0;
}

void CGWorkspace::ensure(size_t bytes) {
  if (r.bytes == bytes) return;
  for (DeviceBuffer *v : {&r, &z, &p, &h}) v->allocate(bytes);
}

}  // namespace mfg

using namespace mfg;

// --------------------------------------------------------------- C ABI ---
extern "C" {

int mfg_transfer_create(mfg_operator coarse, mfg_operator fine, const int32_t *children, mfg_transfer *out) {
  return guarded([&] {
    MFG_REQUIRE(out, MFG_ERR_ARGUMENT, "null argument");
    auto t = new mfg_transfer_t();
    try {
      t->build(coarse, fine, children);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int mfg_transfer_destroy(mfg_transfer t) {
  return guarded([&] { delete t; });
}

int mfg_transfer_prolongate_add(mfg_transfer t, void *fine, const void *coarse) {
  return guarded([&] {
    MFG_REQUIRE(t && fine && coarse, MFG_ERR_ARGUMENT, "null argument");
    t->prolongate_add(fine, coarse);
  });
}
int mfg_transfer_restrict_add(mfg_transfer t, void *coarse, const void *fine) {
  return guarded([&] {
    MFG_REQUIRE(t && fine && coarse, MFG_ERR_ARGUMENT, "null argument");
    t->restrict_add(coarse, fine);
  });
}
int mfg_transfer_interpolate(mfg_transfer t, void *coarse, const void *fine) {
  return guarded([&] {
    MFG_REQUIRE(t && fine && coarse, MFG_ERR_ARGUMENT, "null argument");
    t->interpolate(coarse, fine);
  });
}

int mfg_mg_create(int32_t n_levels, const mfg_operator *ops, const mfg_transfer *transfers,
                  const mfg_mg_params *params, mfg_mg *out) {
  return guarded([&] {
    MFG_REQUIRE(n_levels >= 1 && ops && params && out, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(n_levels == 1 || transfers, MFG_ERR_ARGUMENT, "transfers required for more than one level");
    auto mg = new mfg_mg_t();
    mg->prm = *params;
    for (int l = 0; l < n_levels; ++l) {
      if (!ops[l]) { delete mg; throw Error(MFG_ERR_ARGUMENT, "null level operator"); }
      MGLevel L;
      L.op = ops[l];
      mg->levels.push_back(std::move(L));
      if (l > 0) {
        mfg_transfer t = transfers[l - 1];
        if (!t || t->coarse != ops[l - 1] || t->fine != ops[l]) {
          delete mg;
          throw Error(MFG_ERR_ARGUMENT, "transfer does not connect the given levels");
        }
        mg->transfers.push_back(t);
      }
    }
    *out = mg;
  });
}

int mfg_mg_destroy(mfg_mg mg) {
  return guarded([&] { delete mg; });
}

int mfg_mg_setup(mfg_mg mg) {
  return guarded([&] {
    MFG_REQUIRE(mg, MFG_ERR_ARGUMENT, "null multigrid");
    mg->setup();
  });
}

int mfg_mg_vmult(mfg_mg mg, void *dst, const void *src, int32_t vec_precision) {
  return guarded([&] {
    MFG_REQUIRE(mg && dst && src, MFG_ERR_ARGUMENT, "null argument");
    mg->vmult(dst, src, vec_precision);
  });
}

int mfg_mg_use_graph(mfg_mg mg, int32_t enable) {
  return guarded([&] {
    MFG_REQUIRE(mg, MFG_ERR_ARGUMENT, "null argument");
    mg->use_graph = enable != 0;
    mg->drop_graphs();
  });
}

int mfg_mg_graph_info(mfg_mg mg, int64_t *graph_nodes) {
  return guarded([&] {
    MFG_REQUIRE(mg && graph_nodes, MFG_ERR_ARGUMENT, "null argument");
    *graph_nodes = (int64_t)mg->graph_nodes;
  });
}

int mfg_mg_level_info(mfg_mg mg, int32_t level, double *lmax, double *lmin, int32_t *degree) {
  return guarded([&] {
    MFG_REQUIRE(mg && level >= 0 && level < (int)mg->levels.size(), MFG_ERR_ARGUMENT, "bad level");
    const MGLevel &L = mg->levels[level];
    if (lmax) *lmax = L.lmax;
    if (lmin) *lmin = L.lmin;
    if (degree) *degree = L.degree;
  });
}

int mfg_cg_solve(mfg_operator A, int32_t preconditioner, mfg_mg mg, double omega, const void *b, void *x,
                 double tol_rel, int32_t max_iter, mfg_solve_info *info) {
  mfg_solve_info dummy;
  if (!info) info = &dummy;
  std::memset(info, 0, sizeof(*info));
  return guarded([&] {
    MFG_REQUIRE(A && b && x, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(preconditioner >= 0 && preconditioner <= 2, MFG_ERR_ARGUMENT, "unknown preconditioner");
    static thread_local CGWorkspace ws;
    static thread_local Reducer red;
    cg_solve(*A, preconditioner, mg, omega, b, x, tol_rel, max_iter, info, ws, red);
  });
}

int mfg_box_mesh_children(mfg_mesh coarse, mfg_mesh fine, int32_t *children) {
  return guarded([&] {
    MFG_REQUIRE(coarse && fine && children, MFG_ERR_ARGUMENT, "null argument");
    mfg_level_desc dc, df;
    MFG_REQUIRE(mfg_box_mesh_describe(coarse, &dc) == 0 && mfg_box_mesh_describe(fine, &df) == 0, MFG_ERR_ARGUMENT,
                "bad mesh");
    int nc[3], nf[3];
    mfg_box_mesh_cells(coarse, nc);
    mfg_box_mesh_cells(fine, nf);
    for (int d = 0; d < 3; ++d)
      MFG_REQUIRE(nf[d] == 2 * nc[d], MFG_ERR_ARGUMENT, "fine mesh is not the once-refined coarse mesh");
    for (int cz = 0; cz < nc[2]; ++cz)
      for (int cy = 0; cy < nc[1]; ++cy)
        for (int cx = 0; cx < nc[0]; ++cx) {
          const int64_t c = cx + (int64_t)nc[0] * (cy + (int64_t)nc[1] * cz);
          for (int ch = 0; ch < 8; ++ch) {
            const int fx = 2 * cx + (ch & 1), fy = 2 * cy + ((ch >> 1) & 1), fz = 2 * cz + ((ch >> 2) & 1);
            children[c * 8 + ch] = (int32_t)(fx + (int64_t)nf[0] * (fy + (int64_t)nf[1] * fz));
          }
        }
  });
}

}  // extern "C"
